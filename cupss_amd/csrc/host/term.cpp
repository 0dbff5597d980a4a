// term.cpp -- host handle of one right-hand-side term (reference:
// src/term.cpp, src/term_init.cpp).  Its multiplier
// M(q) = sum_p (-1)^((I_p-m)/2) pre_p q^(2 q2n_p) qx^a qy^b qz^c |q|^-invq,
// times i when m is odd (src/term_init.cpp:135-200), is evaluated in-kernel
// from prefactors_h; precomputePrefactors() pushes them to the device.
#include <cstdlib>
#include <iostream>
#include <vector>

#include "../../../include/cupss.h"
#include "../../../include/cupss_hip.h"

void term::common_constructor()
{
    isCUDA = true;
    prefactors_d = nullptr;
    term_real = term_comp = term_real_d = term_comp_d = nullptr;
    product_h = product_d = nullptr;
    precomp_prefactor_h = precomp_prefactor_d = nullptr;
    multiply_by_i_pre = 0;
    device_id = -1;
    owner = nullptr;
}

term::~term() {}

term::term(int _sx, float _dx)
    : sx(_sx), sy(1), sz(1), dx(_dx), dy(1.0f), dz(1.0f), stepqx(2.0f * PI / (_dx * (float)_sx)),
      stepqy(2.0f * PI), stepqz(2.0f * PI)
{
    common_constructor();
}

term::term(int _sx, int _sy, float _dx, float _dy)
    : sx(_sx), sy(_sy), sz(1), dx(_dx), dy(_dy), dz(1.0f), stepqx(2.0f * PI / (_dx * (float)_sx)),
      stepqy(2.0f * PI / (_dy * (float)_sy)), stepqz(2.0f * PI)
{
    common_constructor();
}

term::term(int _sx, int _sy, int _sz, float _dx, float _dy, float _dz)
    : sx(_sx), sy(_sy), sz(_sz), dx(_dx), dy(_dy), dz(_dz), stepqx(2.0f * PI / (_dx * (float)_sx)),
      stepqy(2.0f * PI / (_dy * (float)_sy)), stepqz(2.0f * PI / (_dz * (float)_sz))
{
    common_constructor();
}

int term::prepareDevice()
{
    // dealias orders are assigned by the device system in cupss_system_prepare
    return precomputePrefactors();
}

int term::precomputePrefactors()
{
    if (prefactors_h.empty()) return 0;
    const int m = (prefactors_h[0].iqx + prefactors_h[0].iqy + prefactors_h[0].iqz) % 2;
    multiply_by_i_pre = m;
    if (!owner || device_id < 0) return 0;
    std::vector<cupss_pres> ps;
    for (const pres &p : prefactors_h) {
        cupss_pres c;
        c.preFactor = p.preFactor;
        c.q2n = p.q2n;
        c.iqx = p.iqx;
        c.iqy = p.iqy;
        c.iqz = p.iqz;
        c.invq = p.invq;
        ps.push_back(c);
    }
    int rc = cupss_term_set_prefactors(owner->deviceSystem(), device_id, ps.data(), (int)ps.size());
    if (rc != CUPSS_OK) {
        std::cerr << "cupss: cupss_term_set_prefactors failed: " << cupss_last_error() << std::endl;
        std::exit(1);
    }
    return 0;
}

int term::update()
{
    std::cout << "term::update: terms are evaluated by evolver::advanceTime on the device in this library"
              << std::endl;
    std::exit(1);
}

int term::setPrefactorString(const std::vector<std::string> &s)
{
    for (const std::string &x : s) prefactor_strings.push_back(x);
    return 0;
}

void term::printPrefactorString()
{
    for (const std::string &s : prefactor_strings) std::cout << s << std::endl;
    for (const auto &kv : usedParameters) std::cout << kv.first << " " << kv.second << std::endl;
}
