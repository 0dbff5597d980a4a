// field.cpp -- host handle of one field (reference: src/field.cpp,
// src/field_init.cpp).  The state lives on the device (half spectrum, see
// cupss_hip.h); this class keeps the reference's host arrays and public
// members and moves data between them and the device.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <vector>

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

[[noreturn]] void per_field_step(const char *fn)
{
    std::cout << "field::" << fn
              << ": per-field stepping is not available in this library; evolver::advanceTime advances the "
                 "whole system on the device"
              << std::endl;
    std::exit(1);
}

}  // namespace

void field::common_constructor()
{
    const size_t n = (size_t)sx * sy * sz;
    real_array = new float2[n];
    comp_array = new float2[n];
    comp_dealiased = new float2[n];
    real_dealiased = new float2[n];
    for (size_t i = 0; i < n; i++) {
        real_array[i] = make_float2(0.0f, 0.0f);
        comp_array[i] = make_float2(0.0f, 0.0f);
        comp_dealiased[i] = make_float2(0.0f, 0.0f);
        real_dealiased[i] = make_float2(0.0f, 0.0f);
    }
    real_array_d = comp_array_d = nullptr;
    terms_h = terms_d = nullptr;
    comp_dealiased_d = real_dealiased_d = nullptr;
    gen_noise = nullptr;
    noise_comp = noise_gend = noise_real = noise_fourier = nullptr;
    noise_comp_d_r = noise_comp_d_i = nullptr;
    precomp_implicit = precomp_implicit_d = nullptr;
    precomp_noise = precomp_noise_d = nullptr;
    implicit_terms = nullptr;
    needsaliasing = false;
    aliasing_order = 1;
    outputToFile = false;
    hasCB = false;
    callback = NULL;
    hasCBFourier = false;
    callbackFourier = NULL;
    hasCBDevice = false;
    callbackDevice = NULL;
    hasCBFourierDevice = false;
    callbackFourierDevice = NULL;
    isNoisy = false;
    noiseType = GaussianWhite;
    integrator = EULER;
    isCUDA = true;
    dynamic = false;
    system_p = nullptr;
    device_id = -1;
}

field::~field()
{
    delete[] real_array;
    delete[] comp_array;
    delete[] comp_dealiased;
    delete[] real_dealiased;
    for (term *t : terms) delete t;
}

// src/field_init.cpp:191-213: stepq = 2*PI/(d*s) in float
field::field(int _sx, float _dx)
    : sx(_sx), sy(1), sz(1), dx(_dx), dy(1), dz(1), stepqx(2.0f * PI / (_dx * (float)_sx)), stepqy(2.0f * PI),
      stepqz(2.0f * PI), rng(rd()), dist(std::normal_distribution<>(0, 1.0))
{
    common_constructor();
}

field::field(int _sx, int _sy, float _dx, float _dy)
    : sx(_sx), sy(_sy), sz(1), dx(_dx), dy(_dy), dz(1.0f), stepqx(2.0f * PI / (_dx * (float)_sx)),
      stepqy(2.0f * PI / (_dy * (float)_sy)), stepqz(2.0f * PI), rng(rd()), dist(std::normal_distribution<>(0, 1.0))
{
    common_constructor();
}

field::field(int _sx, int _sy, int _sz, float _dx, float _dy, float _dz)
    : sx(_sx), sy(_sy), sz(_sz), dx(_dx), dy(_dy), dz(_dz), stepqx(2.0f * PI / (_dx * (float)_sx)),
      stepqy(2.0f * PI / (_dy * (float)_sy)), stepqz(2.0f * PI / (_dz * (float)_sz)), rng(rd()),
      dist(std::normal_distribution<>(0, 1.0))
{
    common_constructor();
}

static cupss_system_t *sys_of(field *f)
{
    if (!f->system_p || f->deviceId() < 0) {
        std::cout << "field " << f->name << " is not attached to an evolver" << std::endl;
        std::exit(1);
    }
    return f->system_p->deviceSystem();
}

void field::prepareDevice()
{
    for (term *t : terms) t->prepareDevice();
}

// field::precalculateImplicit (src/field_init.cpp:237-284): the device
// evaluates D(q) and the noise amplitude in-kernel from these descriptors.
void field::precalculateImplicit(float)
{
    std::vector<cupss_pres> ps;
    for (const pres &p : implicit) {
        cupss_pres c;
        c.preFactor = p.preFactor;
        c.q2n = p.q2n;
        c.iqx = p.iqx;
        c.iqy = p.iqy;
        c.iqz = p.iqz;
        c.invq = p.invq;
        ps.push_back(c);
    }
    abi_check(cupss_field_set_implicits(sys_of(this), device_id, ps.data(), (int)ps.size()),
              "cupss_field_set_implicits");
}

void field::copyHostToDevice()
{
    const size_t n = (size_t)sx * sy * sz;
    std::vector<float> x(n);
    for (size_t i = 0; i < n; i++) x[i] = real_array[i].x;
    abi_check(cupss_field_set_real(sys_of(this), device_id, x.data()), "cupss_field_set_real");
}

void field::copyRealHostToDevice() { copyHostToDevice(); }

void field::copyRealDeviceToHost()
{
    const size_t n = (size_t)sx * sy * sz;
    std::vector<float> x(n);
    abi_check(cupss_field_get_real(sys_of(this), device_id, x.data()), "cupss_field_get_real");
    for (size_t i = 0; i < n; i++) real_array[i] = make_float2(x[i], 0.0f);
    if (needsaliasing) {
        abi_check(cupss_field_get_dealiased(sys_of(this), device_id, x.data()), "cupss_field_get_dealiased");
        for (size_t i = 0; i < n; i++) real_dealiased[i] = make_float2(x[i], 0.0f);
    }
}

// comp_array in the reference's full C2C layout, from the half spectrum
void field::copyDeviceToHost()
{
    copyRealDeviceToHost();
    const int nc1 = sx / 2 + 1;
    std::vector<float> h(2 * (size_t)nc1 * sy * sz);
    abi_check(cupss_field_get_spectrum(sys_of(this), device_id, h.data()), "cupss_field_get_spectrum");
    for (int k = 0; k < sz; k++)
        for (int j = 0; j < sy; j++)
            for (int i = 0; i < sx; i++) {
                const size_t idx = (size_t)k * sx * sy + (size_t)j * sx + i;
                if (i < nc1) {
                    const size_t hi = ((size_t)k * sy + j) * nc1 + i;
                    comp_array[idx] = make_float2(h[2 * hi], h[2 * hi + 1]);
                } else {
                    const int mk = (sz - k) % sz, mj = (sy - j) % sy, mi = sx - i;
                    const size_t hi = ((size_t)mk * sy + mj) * nc1 + mi;
                    comp_array[idx] = make_float2(h[2 * hi], -h[2 * hi + 1]);
                }
            }
}

void field::toComp() { copyHostToDevice(); }
void field::toReal() { copyRealDeviceToHost(); }
void field::normalize() {}   // the inverse transform of this library is normalised
void field::dealias() {}     // the mask is applied on the device where products need it

void field::setToZero()
{
    const size_t n = (size_t)sx * sy * sz;
    for (size_t i = 0; i < n; i++) comp_array[i] = make_float2(0.0f, 0.0f);
}

int field::setRHS(float) { per_field_step("setRHS"); }
int field::updateTerms() { per_field_step("updateTerms"); }
void field::createNoise() { per_field_step("createNoise"); }
void field::setNotDynamic(float) { per_field_step("setNotDynamic"); }
void field::setDynamic(float) { per_field_step("setDynamic"); }
void field::stepEuler(float) { per_field_step("stepEuler"); }
void field::stepRK2(float) { std::cout << "RK2 not implemented" << std::endl; }
void field::stepRK4(float) { std::cout << "RK4 not implemented" << std::endl; }

// field::writeToFile (src/field.cpp:350-402): data/<name>.csv.<step>,
// header "x, [y, [z, ]]name", rows "%i, ... %.<precision>f"; NaN aborts
// (checked here); the file itself is written by the background writer
// (output.hpp).
void field::writeToFile(int currentTimeStep, int dim, int writePrecision)
{
    if (!outputToFile) return;
    if (isCUDA) copyRealDeviceToHost();
    const size_t n = (size_t)sx * sy * sz;
    cupss_host::OutputJob job;
    job.file = "data/" + name + ".csv." + std::to_string(currentTimeStep);
    job.name = name;
    job.dim = dim;
    job.sx = sx;
    job.sy = sy;
    job.sz = sz;
    job.precision = writePrecision;
    job.values.resize(n);
    for (size_t i = 0; i < n; i++) {
        const float v = real_array[i].x;
        if (std::isnan(v)) {
            std::cout << "NaN found in field " << name << std::endl;
            std::exit(1);
        }
        job.values[i] = v;
    }
    // formatting and writing continue on the background writer
    cupss_host::output_submit(std::move(job));
}

float field::getStepqx() { return stepqx; }
float field::getStepqy() { return stepqy; }
float field::getStepqz() { return stepqz; }

int field::addImplicitString(const std::string &s)
{
    implicit_prefactor_strings.push_back(s);
    return 0;
}

void field::printImplicitString()
{
    for (const std::string &s : implicit_prefactor_strings) std::cout << s << std::endl;
    for (const auto &kv : usedParameters) std::cout << kv.first << " " << kv.second << std::endl;
}

// field::updateParameter (src/field.cpp:422-441)
int field::updateParameter(const std::string &pname, float)
{
    if (usedParameters[pname]) {
        system_p->_parser->recalculateImplicits(implicit_prefactor_strings, implicit, dynamic ? 1 : 0);
        precalculateImplicit(system_p->dt);
    }
    for (term *t : terms)
        if (t->usedParameters[pname]) {
            system_p->_parser->recalculateImplicits(t->prefactor_strings, t->prefactors_h, 0);
            t->precomputePrefactors();
        }
    return 0;
}
