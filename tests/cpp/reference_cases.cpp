// reference_cases.cpp -- the scenarios of the reference test suite
// (tests/tests.cpp:64-422) through this library's C++ API: droplet initial
// conditions (1D/2D/3D) and one step of the spectral operators -q^2, iqx,
// iqy, iqz, 1/q (1D/3D), for RUN_CPU and RUN_GPU systems, against the golden
// files.  The iqy/iqz fields are checked against their own files.
// Usage: reference_cases <dir with base_truths/>   (exit status = failures)
#include <cmath>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/cupss.h"

static std::string g_dir;
static int g_fail = 0;

static std::vector<float> load(const std::string &name, int dim)
{
    std::vector<float> v;
    std::ifstream in(g_dir + "/base_truths/" + name);
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        std::stringstream ss(line);
        std::string tok;
        for (int i = 0; i <= dim; i++) std::getline(ss, tok, ',');
        v.push_back(std::stof(tok));
    }
    return v;
}

static void expect_near(const char *what, const float2 *got, const std::vector<float> &truth, size_t n)
{
    if (truth.size() != n) {
        std::cout << "FAIL " << what << ": fixture has " << truth.size() << " values" << std::endl;
        g_fail++;
        return;
    }
    float worst = 0.0f;
    for (size_t i = 0; i < n; i++) worst = std::fmax(worst, std::fabs(got[i].x - truth[i]));
    const bool ok = worst < 1e-4f;
    std::cout << (ok ? "ok   " : "FAIL ") << what << " max|diff| = " << worst << std::endl;
    if (!ok) g_fail++;
}

static void init_case(bool mode, int dim)
{
    const std::string f = "phi_" + std::string(mode == RUN_CPU ? "cpu" : "gpu") + "_" + std::to_string(dim) + "d";
    evolver *e = dim == 1 ? new evolver(mode, 16, 1.0, 1.0, 1)
                          : dim == 2 ? new evolver(mode, 16, 16, 1.0, 1.0, 1.0, 1)
                                     : new evolver(mode, 16, 16, 16, 1.0, 1.0, 1.0, 1.0, 1);
    e->createField(f, true);
    e->initializeDroplet(f, 0, 1, 16 / 8, 4, 16 / 2, dim > 1 ? 16 / 2 : 0, 0);
    e->prepareProblem();
    e->copyAllDataToHost();
    const size_t n = dim == 1 ? 16 : dim == 2 ? 256 : 4096;
    expect_near(("init " + f).c_str(), e->fieldsMap[f]->real_array, load("phi_" + std::to_string(dim) + "d", dim), n);
    delete e;
}

static void operator_case(bool mode, int dim)
{
    const std::string sfx = std::string(mode == RUN_CPU ? "_cpu_" : "_gpu_") + std::to_string(dim) + "d";
    evolver *e = dim == 1 ? new evolver(mode, 16, 1.0, 1.0, 1) : new evolver(mode, 16, 16, 16, 1.0, 1.0, 1.0, 1.0, 1);
    std::vector<std::string> ops = {"lapphi", "iqxphi"};
    if (dim == 3) {
        ops.push_back("iqyphi");
        ops.push_back("iqzphi");
    }
    ops.push_back("invqphi");
    e->createField("phi" + sfx, true);
    for (auto &o : ops) e->createField(o + sfx, false);
    e->addEquation("dt phi" + sfx + " +q^2*phi" + sfx + " = iqxphi" + sfx + "^2");
    e->addEquation("lapphi" + sfx + " = -q^2*phi" + sfx);
    e->addEquation("iqxphi" + sfx + " =  iqx*phi" + sfx);
    if (dim == 3) {
        e->addEquation("iqyphi" + sfx + " =  iqy*phi" + sfx);
        e->addEquation("iqzphi" + sfx + " =  iqz*phi" + sfx);
    }
    e->addEquation("invqphi" + sfx + " =  1/q*phi" + sfx);
    e->initializeDroplet("phi" + sfx, 0, 1, 16 / 8, 4, 16 / 2, dim == 3 ? 16 / 2 : 0, 0);
    e->prepareProblem();
    e->advanceTime();
    e->copyAllDataToHost();
    const size_t n = dim == 1 ? 16 : 4096;
    for (auto &o : ops)
        for (const char *dev : {"cpu", "gpu"}) {
            const std::string fx = o + "_" + dev + "_" + std::to_string(dim) + "d";
            expect_near(("op " + o + sfx + " vs " + fx).c_str(), e->fieldsMap[o + sfx]->real_array, load(fx, dim), n);
        }
    delete e;
}

int main(int argc, char **argv)
{
    g_dir = argc > 1 ? argv[1] : "tests/golden";
    for (bool mode : {(bool)RUN_CPU, (bool)RUN_GPU}) {
        for (int dim = 1; dim <= 3; dim++) init_case(mode, dim);
        operator_case(mode, 1);
        operator_case(mode, 3);
    }
    std::cout << (g_fail ? "FAILED " : "PASSED ") << g_fail << std::endl;
    return g_fail;
}
