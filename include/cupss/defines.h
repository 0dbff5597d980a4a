// defines.h -- shared definitions (reference: inc/cupss/defines.h:1-62).
#ifndef HELPS_H
#define HELPS_H

#include <hip/hip_runtime.h>

#include <iostream>
#include <map>
#include <string>
#include <vector>

struct system_constants
{
    // Integration constants
    int sx, sy, sz;
    float dx, dy, dz;
    float dt;

    // Output constants
    int writeEveryNSteps;
};

// Operator descriptor: preFactor * q^(2*q2n) * (iqx)^iqx (iqy)^iqy (iqz)^iqz * |q|^(-invq).
// Same layout as cupss_pres in cupss_hip.h.
struct pres
{
    float preFactor = 0.0f;
    int q2n = 0;
    int iqx = 0;
    int iqy = 0;
    int iqz = 0;
    int invq = 0;
};

struct full_term
{
    std::vector<pres> prefactors;
    std::vector<std::string> fields;
};

#endif

#ifndef DEFS_H
#define DEFS_H

// The reference's PI is a float literal (defines.h:54); the q grid uses it.
#define PI 3.1415926535f

// RUN_CPU is accepted for compatibility; this library executes every system on
// the GPU (see evolver.h).
enum DeviceToRunOn {RUN_CPU, RUN_GPU};

enum integrators { EULER, RK2, RK4 };

enum NoiseType {GaussianWhite};

#endif
