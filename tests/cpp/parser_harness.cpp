// parser_harness.cpp -- CPU test harness for the product parser
// (cupss_amd/csrc/host/parser.cpp).  It links parser.cpp with minimal
// recording definitions of the evolver/field/term members the parser calls
// and prints the parsed system as JSON lines, one per field:
//   {"field": name, "dynamic": d, "implicit": [[pre,q2n,iqx,iqy,iqz,invq],...],
//    "terms": [{"product": [...], "pres": [[...]...], "used": [...]}], "used": [...]}
// Usage: parser_harness <file with lines "field NAME DYN", "param NAME VALUE",
//        "eq EQUATION", "noise FIELD EXPR", "update NAME VALUE">
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>

#include "../../include/cupss.h"

// ---- recording stand-ins (test only) ---------------------------------------
field::field(int _sx, int _sy, int _sz, float _dx, float _dy, float _dz)
    : sx(_sx), sy(_sy), sz(_sz), dx(_dx), dy(_dy), dz(_dz), stepqx(0), stepqy(0), stepqz(0)
{
    dynamic = false;
    isNoisy = false;
    outputToFile = false;
    device_id = -1;
}
field::~field() {}
int field::addImplicitString(const std::string &s)
{
    implicit_prefactor_strings.push_back(s);
    return 0;
}
term::term(int _sx, int _sy, int _sz, float _dx, float _dy, float _dz)
    : sx(_sx), sy(_sy), sz(_sz), dx(_dx), dy(_dy), dz(_dz), stepqx(0), stepqy(0), stepqz(0)
{
}
term::~term() {}
int term::setPrefactorString(const std::vector<std::string> &s)
{
    for (auto &x : s) prefactor_strings.push_back(x);
    return 0;
}
evolver::evolver(bool, int _sx, int _sy, float _dx, float _dy, float _dt, int _ses)
    : sx(_sx), sy(_sy), sz(1), dx(_dx), dy(_dy), dz(1), writeEveryNSteps(_ses), dt(_dt)
{
    _parser = new parser(this);
}
evolver::~evolver() {}
int evolver::createField(std::string name, bool dyn)
{
    field *f = new field(sx, sy, sz, dx, dy, dz);
    f->name = name;
    f->dynamic = dyn;
    fields.push_back(f);
    fieldsMap[name] = f;
    return 0;
}
void evolver::setOutputField(const std::string &, int) {}
int evolver::createTerm(const std::string &fname, const std::vector<pres> &ps, const std::vector<std::string> &prod)
{
    term *t = new term(sx, sy, sz, dx, dy, dz);
    for (auto &n : prod)
        for (field *f : fields)
            if (f->name == n) t->product.push_back(f);
    t->prefactors_h = ps;
    fieldsMap[fname]->terms.push_back(t);
    return 0;
}
// ---------------------------------------------------------------------------

static void pr(const pres &p)
{
    std::printf("[%.9g,%d,%d,%d,%d,%d]", p.preFactor, p.q2n, p.iqx, p.iqy, p.iqz, p.invq);
}

static void used(const std::map<std::string, int> &m)
{
    std::printf("[");
    bool first = true;
    for (auto &kv : m)
        if (kv.second) {
            std::printf("%s\"%s\"", first ? "" : ",", kv.first.c_str());
            first = false;
        }
    std::printf("]");
}

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    evolver e(RUN_GPU, 16, 16, 1.0f, 1.0f, 0.1f, 1);
    std::ifstream in(argv[1]);
    std::string line;
    std::map<std::string, pres> noise;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string kind;
        ss >> kind;
        std::string rest;
        std::getline(ss, rest);
        if (!rest.empty() && rest[0] == ' ') rest = rest.substr(1);
        if (kind == "field") {
            std::istringstream r(rest);
            std::string n;
            int d;
            r >> n >> d;
            e.createField(n, d);
        } else if (kind == "param") {
            std::istringstream r(rest);
            std::string n;
            float v;
            r >> n >> v;
            e._parser->insert_parameter(n, v);
        } else if (kind == "eq") {
            e._parser->add_equation(rest);
        } else if (kind == "noise") {
            std::istringstream r(rest);
            std::string f, ex;
            r >> f;
            std::getline(r, ex);
            noise[f] = e._parser->add_noise(ex);
        } else if (kind == "update") {
            std::istringstream r(rest);
            std::string n;
            float v;
            r >> n >> v;
            e._parser->changeParameter(n, v);
            for (field *f : e.fields) {
                if (f->usedParameters[n]) {
                    // same recalculation as field::updateParameter
                    std::vector<std::string> strs;
                    (void)strs;
                }
            }
        }
    }
    for (field *f : e.fields) {
        std::printf("{\"field\":\"%s\",\"dynamic\":%d,\"implicit\":[", f->name.c_str(), f->dynamic ? 1 : 0);
        for (size_t i = 0; i < f->implicit.size(); i++) {
            if (i) std::printf(",");
            pr(f->implicit[i]);
        }
        std::printf("],\"terms\":[");
        for (size_t t = 0; t < f->terms.size(); t++) {
            term *T = f->terms[t];
            std::printf("%s{\"product\":[", t ? "," : "");
            for (size_t k = 0; k < T->product.size(); k++)
                std::printf("%s\"%s\"", k ? "," : "", T->product[k]->name.c_str());
            std::printf("],\"pres\":[");
            for (size_t k = 0; k < T->prefactors_h.size(); k++) {
                if (k) std::printf(",");
                pr(T->prefactors_h[k]);
            }
            std::printf("],\"strings\":[");
            for (size_t k = 0; k < T->prefactor_strings.size(); k++)
                std::printf("%s\"%s\"", k ? "," : "", T->prefactor_strings[k].c_str());
            std::printf("],\"used\":");
            used(T->usedParameters);
            std::printf("}");
        }
        std::printf("],\"used\":");
        used(f->usedParameters);
        if (noise.count(f->name)) {
            std::printf(",\"noise\":");
            pr(noise[f->name]);
        }
        std::printf("}\n");
    }
    return 0;
}
