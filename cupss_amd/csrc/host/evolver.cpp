// evolver.cpp -- the system driver on top of the C-ABI (reference:
// src/evolver.cpp:1-394).  Set-up, parameters, output cadence and
// information printing follow the reference; the per-step work is one call
// to cupss_system_advance (phases A-D of src/evolver.cpp:199-226 run on the
// device).
#include <sys/stat.h>

#include <cmath>
#include <cstdlib>
#include <ctime>
#include <iomanip>
#include <iostream>
#include <sstream>

#include "../../../include/cupss.h"
#include "../../../include/cupss_hip.h"
#include "output.hpp"

namespace {

void abi_check(int rc, const char *what)
{
    if (rc != CUPSS_OK) {
        std::cerr << "cupss: " << what << " failed: " << cupss_last_error() << std::endl;
        std::exit(1);
    }
}

cupss_pres to_abi(const pres &p)
{
    cupss_pres c;
    c.preFactor = p.preFactor;
    c.q2n = p.q2n;
    c.iqx = p.iqx;
    c.iqy = p.iqy;
    c.iqz = p.iqz;
    c.invq = p.invq;
    return c;
}

// Bridges the C-ABI hooks to the reference callback signature.  On RUN_CPU
// systems the callback gets a host copy, as the reference CPU mode does.
void real_hook(void *user, int, void *real2, int which)
{
    field *f = static_cast<field *>(user);
    evolver *e = f->system_p;
    const int sx = e->getSystemSizeX(), sy = e->getSystemSizeY(), sz = e->getSystemSizeZ();
    if (!f->hasCB || !f->callback) return;
    if (f->isCUDA) {
        f->callback(e, static_cast<float2 *>(real2), sx, sy, sz);
        return;
    }
    const size_t n = (size_t)sx * sy * sz;
    float2 *host = which == 0 ? f->real_array : f->real_dealiased;
    abi_check(cupss_memcpy_d2h(host, real2, n * sizeof(float2)), "callback copy");
    f->callback(e, host, sx, sy, sz);
    abi_check(cupss_memcpy_h2d(real2, host, n * sizeof(float2)), "callback copy");
}

void spec_hook(void *user, int, void *full2)
{
    field *f = static_cast<field *>(user);
    evolver *e = f->system_p;
    const int sx = e->getSystemSizeX(), sy = e->getSystemSizeY(), sz = e->getSystemSizeZ();
    if (!f->hasCBFourier || !f->callbackFourier) return;
    if (f->isCUDA) {
        f->callbackFourier(e, static_cast<float2 *>(full2), sx, sy, sz);
        return;
    }
    const size_t n = (size_t)sx * sy * sz;
    abi_check(cupss_memcpy_d2h(f->comp_array, full2, n * sizeof(float2)), "callback copy");
    f->callbackFourier(e, f->comp_array, sx, sy, sz);
    abi_check(cupss_memcpy_h2d(full2, f->comp_array, n * sizeof(float2)), "callback copy");
}

// device-side callbacks: the library's buffers and stream, passed through
void real_hook_dev(void *user, int, float *real, int sx, int sy, int sz, int which, void *stream)
{
    field *f = static_cast<field *>(user);
    if (f->hasCBDevice && f->callbackDevice)
        f->callbackDevice(f->system_p, real, sx, sy, sz, which, static_cast<hipStream_t>(stream));
}

void spec_hook_dev(void *user, int, void *half2, int sx, int sy, int sz, void *stream)
{
    field *f = static_cast<field *>(user);
    if (f->hasCBFourierDevice && f->callbackFourierDevice)
        f->callbackFourierDevice(f->system_p, static_cast<float2 *>(half2), sx, sy, sz,
                                 static_cast<hipStream_t>(stream));
}

}  // namespace

void evolver::common_constructor()
{
    currentTime = 0.0f;
    currentTimeStep = 0;
    dtsqrt = std::sqrt(dt);
    writeParametersOnUpdate = true;
    writePrecision = 6;
    _parser = new parser(this);
    verbose = false;
    noise_seed = (uint64_t)time(NULL) * 0x9E3779B97F4A7C15ull;
    device_system = nullptr;

    // launch geometry of the reference (src/evolver.cpp:22-36); kept for user
    // callbacks that launch their own kernels with blocks/threads_per_block
    if (sz == 1 && sy == 1) {
        dimension = 1;
        blocks = dim3(1);
        threads_per_block = dim3(sx < 1024 ? sx : 1024);
        blocks = dim3((sx + threads_per_block.x - 1) / threads_per_block.x);
    } else if (sz == 1) {
        dimension = 2;
        threads_per_block = dim3(32, 32);
        blocks = dim3((sx + 31) / 32, (sy + 31) / 32);
    } else {
        dimension = 3;
        threads_per_block = dim3(16, 8, 8);
        blocks = dim3((sx + 15) / 16, (sy + 7) / 8, (sz + 7) / 8);
    }
    abi_check(cupss_system_create(sx, sy, sz, dx, dy, dz, dt, &device_system), "cupss_system_create");
}

evolver::~evolver()
{
    cupss_host::output_flush();  // files of this system are complete when it is gone
    for (field *f : fields) delete f;
    delete _parser;
    cupss_system_destroy(device_system);
}

evolver::evolver(bool _with_cuda, int _sx, float _dx, float _dt, int _ses)
    : sx(_sx), sy(1), sz(1), dx(_dx), dy(1.0f), dz(1.0f), writeEveryNSteps(_ses), dt(_dt)
{
    std::srand(time(0));
    with_cuda = _with_cuda;
    common_constructor();
}

evolver::evolver(bool _with_cuda, int _sx, int _sy, float _dx, float _dy, float _dt, int _ses)
    : sx(_sx), sy(_sy), sz(1), dx(_dx), dy(_dy), dz(1.0f), writeEveryNSteps(_ses), dt(_dt)
{
    std::srand(time(NULL));
    with_cuda = _with_cuda;
    common_constructor();
}

evolver::evolver(bool _with_cuda, int _sx, int _sy, int _sz, float _dx, float _dy, float _dz, float _dt, int _ses)
    : sx(_sx), sy(_sy), sz(_sz), dx(_dx), dy(_dy), dz(_dz), writeEveryNSteps(_ses), dt(_dt)
{
    std::srand(time(NULL));
    with_cuda = _with_cuda;
    common_constructor();
}

void evolver::setSeed(uint64_t seed) { noise_seed = seed; }

cupss_system_t *evolver::deviceSystem() { return device_system; }

int evolver::createFromFile(const std::string &file)
{
    _parser->createFromFile(file);
    return 0;
}

void evolver::attachField(field *f)
{
    int id = -1;
    abi_check(cupss_field_create(device_system, f->dynamic ? 1 : 0, &id), "cupss_field_create");
    f->device_id = id;
}

void evolver::syncHooks()
{
    for (field *f : fields) {
        const bool r = f->hasCB && f->callback;
        const bool c = f->hasCBFourier && f->callbackFourier;
        abi_check(cupss_field_set_hooks(device_system, f->device_id, r ? real_hook : nullptr,
                                        c ? spec_hook : nullptr, f),
                  "cupss_field_set_hooks");
        // device-side callbacks act on device buffers: GPU systems only
        const bool rd = with_cuda && f->hasCBDevice && f->callbackDevice;
        const bool cd = with_cuda && f->hasCBFourierDevice && f->callbackFourierDevice;
        abi_check(cupss_field_set_device_hooks(device_system, f->device_id, rd ? real_hook_dev : nullptr,
                                               cd ? spec_hook_dev : nullptr, f),
                  "cupss_field_set_device_hooks");
    }
}

void evolver::prepareProblem()
{
    if (with_cuda && verbose) check_device();
    struct stat info;
    const std::string pathname = "data";
    if (stat(pathname.c_str(), &info) != 0) {
        if (verbose) std::cout << "data directory not found, creating it.\n";
        if (mkdir(pathname.c_str(), S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH) == -1) {
            std::cout << "Error creating data directory\n";
            std::exit(1);
        }
    } else if (info.st_mode & S_IFDIR) {
        if (verbose) std::cout << "data directory already found, might rewrite output data.\n";
    } else {
        std::cout << "Can't create data directory, is there a file called data?\n";
        std::exit(1);
    }
    _parser->writeParamsToFile("data/parameter_list.txt.0");

    if (verbose) {
        std::cout << "Preparing problem." << std::endl;
        std::cout << "Copying initial states to device if necessary." << std::endl;
    }
    for (field *f : fields) f->copyHostToDevice();   // + toComp
    if (verbose) std::cout << "Preparing device and precalculating implicit matrices." << std::endl;
    for (field *f : fields) {
        f->system_p = this;
        f->prepareDevice();
        f->precalculateImplicit(dt);
        if (f->isNoisy)
            abi_check(cupss_field_set_noise(device_system, f->device_id, to_abi(f->noise_amplitude),
                                            noise_seed ^ (0x632BE59BD9B4E019ull * (uint64_t)(f->device_id + 1))),
                      "cupss_field_set_noise");
    }
    syncHooks();
    abi_check(cupss_system_prepare(device_system), "cupss_system_prepare");
    for (field *f : fields) {
        const int order = cupss_field_aliasing_order(device_system, f->device_id);
        f->needsaliasing = order > 0;
        f->aliasing_order = order > 0 ? order : 1;
    }
}

void evolver::setOutputField(const std::string &_name, int _output)
{
    for (field *f : fields)
        if (f->name == _name) {
            f->outputToFile = _output != 0;
            return;
        }
    std::cout << "setOutputField EROR: " << _name << " not found." << std::endl;
}

int evolver::addParameter(const std::string &_name, float value)
{
    _parser->insert_parameter(_name, value);
    return 0;
}

int evolver::addEquation(const std::string &equation)
{
    _parser->add_equation(equation);
    return 0;
}

int evolver::existsField(const std::string &_name)
{
    for (size_t i = 0; i < fields.size(); i++)
        if (fields[i]->name == _name) return (int)i;
    return -1;
}

int evolver::addNoise(const std::string &_name, const std::string &equation)
{
    if (existsField(_name) == -1) {
        std::cout << "Adding noise to non existing field! (" << _name << ")" << std::endl;
        return -1;
    }
    pres prefactor = _parser->add_noise(equation);
    fieldsMap[_name]->isNoisy = true;
    fieldsMap[_name]->noise_amplitude = prefactor;
    return 0;
}

void evolver::addField(field *newField)
{
    // like the reference, the field is owned but not registered in the maps
    fields.push_back(newField);
    newField->isCUDA = with_cuda;
    attachField(newField);
}

int evolver::createField(std::string name, bool dynamic)
{
    for (field *f : fields)
        if (f->name == name) {
            std::cout << "Trying to create field with name that already exists" << std::endl;
            return 1;
        }
    field *newField = new field(sx, sy, sz, dx, dy, dz);
    newField->name = name;
    newField->isCUDA = with_cuda;
    newField->dynamic = dynamic;
    newField->blocks = blocks;
    newField->threads_per_block = threads_per_block;
    newField->system_p = this;
    fields.push_back(newField);
    attachField(newField);
    fieldsMap[name] = newField;
    fieldsReal[name] = newField->real_array;
    fieldsFourier[name] = newField->comp_array;
    return 0;
}

int evolver::advanceTime()
{
    // output before stepping when step % writeEveryNSteps == 0
    // (src/evolver.cpp:200-203); a cadence <= 0 disables output instead of
    // dividing by zero
    if (writeEveryNSteps > 0 && currentTimeStep % writeEveryNSteps == 0) writeOut();
    syncHooks();
    abi_check(cupss_system_advance(device_system, 1), "cupss_system_advance");
    currentTime += dt;
    currentTimeStep += 1;
    if (!with_cuda) copyAllDataToHost();   // host arrays are authoritative in CPU mode
    return 0;
}

void evolver::writeOut()
{
    for (field *f : fields) f->writeToFile(currentTimeStep, dimension, writePrecision);
}

void evolver::setVerbose() { verbose = true; }
void evolver::unsetVerbose() { verbose = false; }

// A summary of the system as parsed: the grid, then one block per field
// with its kind, its dealiasing and its equation written back in the
// parser's own grammar (operator monomials as c*q^2n*iqx^a*iqy^b*iqz^c*1/q^k,
// implicit operators on the left, each term as [operator sum](factors)).
void evolver::printInformation()
{
    auto monomial = [](const pres &p) {
        std::ostringstream os;
        os << p.preFactor;
        if (p.q2n) os << "*q^" << 2 * p.q2n;
        if (p.iqx) os << "*iqx^" << p.iqx;
        if (p.iqy) os << "*iqy^" << p.iqy;
        if (p.iqz) os << "*iqz^" << p.iqz;
        if (p.invq) os << "*1/q^" << p.invq;
        return os.str();
    };
    auto sum = [&](const std::vector<pres> &ps) {
        std::string out;
        for (size_t i = 0; i < ps.size(); i++) out += (i ? " + " : "") + monomial(ps[i]);
        return "[" + out + "]";
    };
    std::ostringstream os;
    os << "evolver: " << dimension << "D grid " << sx << " x " << sy << " x " << sz << ", cells " << dx << " x " << dy
       << " x " << dz << ", dt " << dt << ", step " << currentTimeStep << ", " << (with_cuda ? "GPU" : "CPU mode")
       << ", " << fields.size() << " field(s)\n";
    for (size_t i = 0; i < fields.size(); i++) {
        const field *f = fields[i];
        os << "  [" << i << "] " << f->name << ": " << (f->dynamic ? "dynamic" : "constraint") << ", "
           << f->terms.size() << " term(s), " << f->implicit.size() << " implicit operator(s), ";
        if (f->needsaliasing)
            os << "dealiased for products of order " << f->aliasing_order << " (keeps |n| <= N/"
               << f->aliasing_order + 1 << ")";
        else
            os << "not dealiased";
        os << "\n      " << (f->dynamic ? "dt " : "") << f->name;
        if (!f->implicit.empty()) {
            // the parser stores a dynamic field's implicit operators negated
            std::vector<pres> lhs = f->implicit;
            if (f->dynamic)
                for (pres &p : lhs) p.preFactor = -p.preFactor;
            os << (f->dynamic ? " + " : " * ") << sum(lhs) << (f->dynamic ? " " + f->name : "");
        }
        os << " =";
        if (f->terms.empty() && !f->isNoisy) os << " (unchanged)";
        for (size_t j = 0; j < f->terms.size(); j++) {
            const term *t = f->terms[j];
            os << (j ? " + " : " ") << sum(t->prefactors_h) << "(";
            for (size_t c = 0; c < t->product.size(); c++) os << (c ? "*" : "") << t->product[c]->name;
            os << (t->product.empty() ? "1" : "") << ")";
        }
        if (f->isNoisy) os << (f->terms.empty() ? " " : " + ") << "noise of variance " << monomial(f->noise_amplitude);
        os << "\n";
    }
    std::cout << os.str() << std::flush;
}

int evolver::createTerm(const std::string &_field, const std::vector<pres> &_prefactors,
                        const std::vector<std::string> &_product)
{
    int field_index = existsField(_field);
    if (field_index == -1) {
        std::cout << "Field " << _field << " not found trying to create term" << std::endl;
        return 1;
    }
    term *t = new term(sx, sy, sz, dx, dy, dz);
    t->isCUDA = with_cuda;
    t->owner = this;
    for (const std::string &name : _product)
        for (field *f : fields)
            if (f->name == name) t->product.push_back(f);
    for (const pres &p : _prefactors) t->prefactors_h.push_back(p);
    t->blocks = blocks;
    t->threads_per_block = threads_per_block;
    fields[field_index]->terms.push_back(t);

    std::vector<int> ids;
    for (field *f : t->product) ids.push_back(f->device_id);
    std::vector<cupss_pres> ps;
    for (const pres &p : t->prefactors_h) ps.push_back(to_abi(p));
    abi_check(cupss_term_create(device_system, fields[field_index]->device_id, ids.data(), (int)ids.size(), ps.data(),
                                (int)ps.size(), &t->device_id),
              "cupss_term_create");
    return 0;
}

void evolver::copyAllDataToHost()
{
    for (field *f : fields) f->copyDeviceToHost();
}

int evolver::getSystemSizeX() { return sx; }
int evolver::getSystemSizeY() { return sy; }
int evolver::getSystemSizeZ() { return sz; }

float evolver::getSystemPhysicalSizeX() { return ((float)sx) * dx; }
float evolver::getSystemPhysicalSizeY() { return ((float)sy) * dy; }
float evolver::getSystemPhysicalSizeZ() { return ((float)sz) * dz; }

int evolver::getCurrentTimestep() { return currentTimeStep; }
float evolver::getCurrentTime() { return currentTime; }
bool evolver::getCuda() { return with_cuda; }

float evolver::getParameter(const std::string &name) { return _parser->getParameter(name); }

int evolver::updateParameter(const std::string &name, float new_value)
{
    _parser->changeParameter(name, new_value);
    for (field *f : fields) f->updateParameter(name, new_value);
    if (writeParametersOnUpdate) _parser->writeParamsToFile("data/parameter_list.txt." + std::to_string(currentTimeStep));
    return 0;
}
