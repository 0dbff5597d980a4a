// parser.cpp -- equation strings -> operator descriptors.
//
// Behaviour follows the reference grammar (src/parser.cpp), including the
// details that decide which system is built:
//   * spaces removed; exactly one '=' (src/parser.cpp:406-447);
//   * terms split at '+'/'-' outside parentheses, except at position 0
//     (:449-474); one parenthesised group expanded per pass until none is
//     left, the expansions appended at the end of the term list, signs
//     folded into the prefix (:477-545); '/(' is an error;
//   * a factor list is the term split at '*' outside parentheses; '/' is
//     rewritten '*1/' first so that a/b reads as a * 1/b (:547-589);
//   * factors: fields (optionally ^n, checked first), q^n (q2n += n/2, integer
//     division), iqx/iqy/iqz[^n], 1/q[^n] (invq), parameters or numbers,
//     each optionally prefixed by 1/ (:125-267);
//   * LHS: 'dt<field>' first makes the equation dynamic and the remaining LHS
//     prefactors are negated; otherwise the LHS must be linear in one field;
//     a LHS that is the bare field stores no implicit operator (:628-738);
//   * RHS terms grouped by their sorted field multiset, one term per group,
//     prefactors summed (:702-764); an empty multiset is a constant term;
//   * parameter use is recorded per field/term for updateParameter
//     (:726-764, 811-846); system files have Fields/Parameters/Equations
//     sections (:11-96).
// Errors print a message and exit(1), as the reference does.
#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>

#include "../../../include/cupss.h"

namespace {

[[noreturn]] void die(const std::string &msg)
{
    std::cout << msg << std::endl;
    std::exit(1);
}

// split at '*' when no parenthesis is open (a flag, not a depth counter, as
// in the reference)
std::vector<std::string> split_factors(const std::string &t)
{
    std::vector<std::string> out;
    size_t start = 0;
    bool open = false;
    for (size_t p = 0; p < t.size(); p++) {
        if (t[p] == '(') open = true;
        if (t[p] == ')') open = false;
        if (t[p] == '*' && !open) {
            out.push_back(t.substr(start, p - start));
            start = p + 1;
        }
        if (p == t.size() - 1) out.push_back(t.substr(start));
    }
    return out;
}

std::string strip_sign(const std::string &t)
{
    if (!t.empty() && (t[0] == '+' || t[0] == '-')) return t.substr(1);
    return t;
}

// a/b -> a*1/b
std::string divisions_as_factors(std::string t)
{
    for (size_t i = 0; i < t.size(); i++)
        if (t[i] == '/') {
            t.insert(i, "*1");
            i += 2;
        }
    return t;
}

}  // namespace

parser::parser(evolver *_system) : verbose(false), system(_system) {}

int parser::populate_system() { return 0; }

int parser::is_number(const std::string &term)
{
    std::string s = term.compare(0, 2, "1/") == 0 ? term.substr(2) : term;
    size_t dots = std::count(s.begin(), s.end(), '.');
    if (dots > 1) die("ERROR: prefactor not a parameter and not a number: " + term);
    if (dots == 1) s.erase(s.find('.'), 1);
    for (char c : s)
        if (!std::isdigit((unsigned char)c)) return 0;
    return 1;
}

int parser::is_field(const std::string &factor)
{
    // power after the last '^', 1 if none
    int power = 1;
    size_t caret = std::string::npos;
    for (size_t i = 0; i < factor.size(); i++)
        if (factor[i] == '^') {
            power = std::atoi(factor.c_str() + i + 1);
            caret = i;
        }
    const std::string name = power != 1 ? factor.substr(0, caret) : factor;
    for (field *f : system->fields)
        if (f->name == name) return power;
    return 0;
}

int parser::is_q2n_factor(const std::string &factor)
{
    if (factor.size() < 3 || factor[0] != 'q' || factor[1] != '^') return 0;
    const std::string num = factor.substr(2);
    if (!is_number(num)) die("ERROR in parser: " + num + " power of q not a number, missing *?");
    return std::atoi(num.c_str());
}

int parser::is_iq_factor(const std::string &factor, const std::string &iq)
{
    if (factor.size() < 3 || factor.compare(0, 3, iq) != 0) return 0;
    if (factor.size() == 3) return 1;
    const std::string num = factor.substr(4);   // skips the '^'
    if (!is_number(num)) die("ERROR in parser: " + num + " power of " + iq + " not a number, missing *?");
    return std::atoi(num.c_str());
}

float parser::is_numerical_factor(const std::string &factor)
{
    if (is_q2n_factor(factor) || is_iq_factor(factor, "iqx") || is_iq_factor(factor, "iqy") ||
        is_iq_factor(factor, "iqz") || is_iq_factor(factor, "1/q") || is_field(factor))
        return 1.0f;
    const bool divides = factor.compare(0, 2, "1/") == 0;
    const std::string body = divides ? factor.substr(2) : factor;
    float v;
    if (exists_parameter(body))
        v = parameters[body];
    else if (is_number(body))
        v = std::stof(body);
    else
        die("ERROR, parameter not found and not a number: " + factor);
    return divides ? 1.0f / v : v;
}

std::string parser::extract_field_name(const std::string &term)
{
    if (is_field(term) == 1) return term;
    size_t caret = 0;
    for (size_t i = 0; i < term.size(); i++)
        if (term[i] == '^') caret = i;
    return term.substr(0, caret);
}

std::string parser::get_field_name(const std::string &term)
{
    std::string name;
    for (const std::string &f : split_factors(strip_sign(term)))
        if (is_field(f) > 0) name = f;
    return name;
}

int parser::number_of_fields(const std::string &term)
{
    int count = 0;
    for (const std::string &f : split_factors(strip_sign(term))) count += is_field(f);
    return count;
}

void parser::get_factors(const std::string &term, std::vector<std::string> &factors)
{
    for (const std::string &f : split_factors(term)) factors.push_back(f);
}

void parser::get_par_terms(const std::string &term, std::vector<std::string> &out) { split_terms(term, out); }

pres parser::add_noise(const std::string &equation)
{
    std::string eq = equation;
    eq.erase(std::remove(eq.begin(), eq.end(), ' '), eq.end());
    pres p;
    p.preFactor = 1.0f;
    for (const std::string &f : split_factors(eq)) {
        p.preFactor *= is_numerical_factor(f);
        p.q2n += is_q2n_factor(f) / 2;
        p.invq += is_iq_factor(f, "1/q");
    }
    return p;
}

int parser::split_equation_sides(std::string input, std::string &lhs, std::string &rhs)
{
    const size_t n_eq = std::count(input.begin(), input.end(), '=');
    if (n_eq != 1) {
        std::cout << "Error processing" << std::endl << input << std::endl;
        die("More than one equal sign");
    }
    const size_t pos = input.find('=');
    lhs = input.substr(0, pos);
    rhs = input.substr(pos + 1);
    return 0;
}

int parser::remove_spaces(std::string &input)
{
    input.erase(std::remove(input.begin(), input.end(), ' '), input.end());
    return 0;
}

int parser::split_terms(std::string input, std::vector<std::string> &terms)
{
    size_t current = 0;
    int level = 0;
    for (size_t p = 0; p < input.size(); p++) {
        if (input[p] == '(') level++;
        if (input[p] == ')') level--;
        if (p > 0 && level == 0 && (input[p] == '+' || input[p] == '-')) {
            terms.push_back(input.substr(current, p - current));
            current = p;
        }
    }
    terms.push_back(input.substr(current));
    return 0;
}

int parser::expand_once(std::string input, std::vector<std::string> &terms)
{
    int level = 0;
    bool found = false;
    size_t open_at = 0;
    for (size_t p = 0; p < input.size(); p++) {
        if (input[p] == '(') {
            if (p > 0 && input[p - 1] == '/') die("ERROR: No dividing over parenthesis allowed!");
            found = true;
            if (++level == 1) open_at = p;
        }
        if (input[p] == ')') level--;
        if (level == 0 && found) {
            std::vector<std::string> inner;
            split_terms(input.substr(open_at + 1, p - open_at - 1), inner);
            const std::string pre = input.substr(0, open_at);
            const std::string post = input.substr(p + 1);
            for (std::string sub : inner) {
                std::string prefix = pre;
                if (!sub.empty() && sub[0] == '+') sub = sub.substr(1);
                if (!sub.empty() && sub[0] == '-') {
                    // fold the minus sign into the prefix
                    if (!prefix.empty() && prefix[0] == '-')
                        prefix = "+" + prefix.substr(1);
                    else if (!prefix.empty() && prefix[0] == '+')
                        prefix = "-" + pre.substr(1);
                    else
                        prefix = "-" + prefix;
                    sub = sub.substr(1);
                }
                terms.push_back(prefix + sub + post);
            }
            return 1;
        }
    }
    return 0;
}

int parser::expand_all(std::vector<std::string> &terms)
{
    for (size_t k = 0; k < terms.size(); k++) {
        std::vector<std::string> sub;
        if (expand_once(terms[k], sub) > 0) {
            terms.erase(terms.begin() + k);
            terms.insert(terms.end(), sub.begin(), sub.end());
            k--;
        }
    }
    return 0;
}

int parser::get_fields(std::string term, std::vector<std::string> &fields) { return get_field_vector(term, fields); }

pres parser::get_prefactor(const std::string &term_in)
{
    pres p;
    p.preFactor = 1.0f;
    std::string t = term_in;
    if (!t.empty() && t[0] == '+') t = t.substr(1);
    if (!t.empty() && t[0] == '-') {
        p.preFactor = -1.0f;
        t = t.substr(1);
    }
    for (const std::string &f : split_factors(divisions_as_factors(t))) {
        if (is_field(f) > 0) continue;
        p.preFactor *= is_numerical_factor(f);
        p.q2n += is_q2n_factor(f) / 2;
        p.iqx += is_iq_factor(f, "iqx");
        p.iqy += is_iq_factor(f, "iqy");
        p.iqz += is_iq_factor(f, "iqz");
        p.invq += is_iq_factor(f, "1/q");
    }
    return p;
}

int parser::get_field_vector(const std::string &term_in, std::vector<std::string> &fields)
{
    std::string t = term_in;
    if (!t.empty() && t[0] == '+') t = t.substr(1);
    if (!t.empty() && t[0] == '-') t = t.substr(1);
    for (const std::string &f : split_factors(divisions_as_factors(t))) {
        const int power = is_field(f);
        if (power > 0) {
            const std::string name = extract_field_name(f);
            for (int k = 0; k < power; k++) fields.push_back(name);
        }
    }
    return 0;
}

int parser::add_equation(const std::string &equation_in)
{
    std::string equation = equation_in;
    remove_spaces(equation);
    if (verbose) std::cout << "Processing\n" << equation << "\n";

    std::string lhs, rhs;
    split_equation_sides(equation, lhs, rhs);
    std::vector<std::string> lhs_terms, rhs_terms;
    split_terms(lhs, lhs_terms);
    split_terms(rhs, rhs_terms);
    expand_all(lhs_terms);
    expand_all(rhs_terms);

    bool dynamic = false;
    std::string field_name;
    if (lhs_terms[0].compare(0, 2, "dt") == 0) {
        dynamic = true;
        field_name = lhs_terms[0].substr(2);
        const int power = is_field(field_name);
        if (!power) die("Field not found: " + field_name);
        if (power > 1) die("Nonlinearity in lhs: " + field_name);
        lhs_terms.erase(lhs_terms.begin());
    } else {
        if (number_of_fields(lhs_terms[0]) != 1) {
            std::cout << "ERROR: Nonlinear term in left hand side" << std::endl;
            die(lhs_terms[0]);
        }
        field_name = get_field_name(lhs_terms[0]);
    }

    std::vector<pres> implicits;
    for (const std::string &t : lhs_terms) {
        std::vector<std::string> fl;
        get_field_vector(t, fl);
        if (fl.size() != 1) die("Nonlinear term in lhs");
        if (fl[0] != field_name) die("ERROR: " + field_name + " and " + fl[0] + " incompatible in lhs");
        pres p = get_prefactor(t);
        if (dynamic) p.preFactor *= -1.0f;
        implicits.push_back(p);
    }

    // group the right-hand side by field multiset, in order of appearance
    std::vector<std::vector<std::string>> groups;
    std::vector<std::vector<pres>> group_pres;
    std::vector<std::vector<std::string>> group_strings;
    for (const std::string &t : rhs_terms) {
        std::vector<std::string> fl;
        get_field_vector(t, fl);
        std::sort(fl.begin(), fl.end());
        const pres p = get_prefactor(t);
        bool found = false;
        for (size_t k = 0; k < groups.size(); k++)
            if (groups[k] == fl) {
                found = true;
                group_pres[k].push_back(p);
                group_strings[k].push_back(t);
            }
        if (!found) {
            groups.push_back(fl);
            group_pres.push_back({p});
            group_strings.push_back({t});
        }
    }

    field *target = system->fieldsMap[field_name];
    for (const auto &kv : parameters) target->usedParameters[kv.first] = 0;
    const bool bare = implicits.size() == 1 && implicits[0].preFactor == 1.0f && implicits[0].q2n == 0 &&
                      implicits[0].iqx == 0 && implicits[0].iqy == 0 && implicits[0].iqz == 0 &&
                      implicits[0].invq == 0;
    if (!bare) {
        for (size_t i = 0; i < implicits.size(); i++) {
            target->implicit.push_back(implicits[i]);
            target->addImplicitString(lhs_terms[i]);
            for (const auto &kv : parameters)
                if (isParameterInString(lhs_terms[i], kv.first)) target->usedParameters[kv.first] = 1;
        }
    }

    for (size_t g = 0; g < groups.size(); g++) {
        system->createTerm(field_name, group_pres[g], groups[g]);
        term *t = target->terms.back();
        t->setPrefactorString(group_strings[g]);
        for (const auto &kv : parameters) t->usedParameters[kv.first] = 0;
        for (const std::string &s : group_strings[g])
            for (const auto &kv : parameters)
                if (isParameterInString(s, kv.first)) t->usedParameters[kv.first] = 1;
    }
    return 0;
}

int parser::createFromFile(const std::string &file_name)
{
    std::cout << "Trying to read system from " << file_name << std::endl;
    std::ifstream in(file_name);
    enum { NONE = -1, FIELDS = 0, PARAMS = 1, EQUATIONS = 2 } section = NONE;
    auto header = [](const std::string &l) {
        if (l.compare(0, 6, "Fields") == 0 || l.compare(0, 6, "fields") == 0) return (int)FIELDS;
        if (l.compare(0, 10, "Parameters") == 0 || l.compare(0, 10, "parameters") == 0) return (int)PARAMS;
        // the reference misses the lower-case header after the first section
        // (src/parser.cpp:49); both spellings are accepted here
        if (l.compare(0, 9, "Equations") == 0 || l.compare(0, 9, "equations") == 0) return (int)EQUATIONS;
        return (int)NONE;
    };
    std::vector<std::string> equations;
    std::string line;
    while (std::getline(in, line)) {
        std::cout << line << std::endl;
        if (line.empty() || line[0] == '#') continue;
        const int h = header(line);
        if (h != NONE) {
            section = (decltype(section))h;
            continue;
        }
        if (section == NONE) {
            std::cout << "ERROR: reading file " << file_name << std::endl;
            std::cout << "First line must be either fields, parameters, or equations, it is:" << std::endl;
            std::cout << line << std::endl;
            return -1;
        }
        std::istringstream iss(line);
        if (section == FIELDS) {
            std::string name;
            int dyn, output;
            if (!(iss >> name >> dyn >> output)) {
                std::cout << "Error reading field line: " << line << std::endl;
                std::cout << "Must be: field_name dynamic_value output" << std::endl;
                return -1;
            }
            std::cout << "Creating field: " << name << ", dynamic: " << dyn << std::endl;
            system->createField(name, dyn);
            system->setOutputField(name, output);
        } else if (section == PARAMS) {
            std::string name;
            float value;
            if (!(iss >> name >> value)) {
                std::cout << "Error reading parameter line: " << line << std::endl;
                std::cout << "Must be: param_name value" << std::endl;
                return -1;
            }
            std::cout << "Creating parameter: " << name << " = " << value << std::endl;
            insert_parameter(name, value);
        } else {
            equations.push_back(line);
        }
    }
    for (const std::string &e : equations) add_equation(e);
    return 0;
}

int parser::insert_parameter(const std::string &name, float value)
{
    if (exists_parameter(name)) {
        std::cout << "Duplicate parameter " << name << std::endl;
        return -1;
    }
    parameters[name] = value;
    return 0;
}

int parser::exists_parameter(const std::string &name) { return parameters.count(name) ? 1 : 0; }

void parser::writeParamsToFile(const std::string &file_name)
{
    std::ofstream out(file_name);
    for (const auto &kv : parameters) out << kv.first << "\t" << kv.second << "\n";
}

float parser::getParameter(const std::string &name)
{
    if (!exists_parameter(name)) die("ERROR: getParameter " + name + " not found");
    return parameters[name];
}

int parser::changeParameter(const std::string &name, float value)
{
    if (!exists_parameter(name)) die("ERROR: changeParameter " + name + " not found");
    parameters[name] = value;
    return 0;
}

int parser::isParameterInString(const std::string &s, const std::string &param)
{
    std::string local = strip_sign(s);
    std::replace(local.begin(), local.end(), '/', '*');
    int n = 0;
    for (const std::string &f : split_factors(local))
        if (f == param) n++;
    return n;
}

int parser::recalculateImplicits(const std::vector<std::string> &strings, std::vector<pres> &out, int dynamic)
{
    for (size_t i = 0; i < strings.size(); i++) {
        out[i] = get_prefactor(strings[i]);
        if (dynamic) out[i].preFactor *= -1.0f;
    }
    return 0;
}
