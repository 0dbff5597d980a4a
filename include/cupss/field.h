// field.h -- one field of a system (reference: inc/cupss/field.h:21-145).
//
// Public members are those of the reference.  Host arrays: real_array and
// comp_array (float2[N], the full C2C layout of the reference) are
// authoritative until prepareProblem(); afterwards the device is, and the
// host copies are refreshed by writeOut()/copyAllDataToHost() (and after
// every step for RUN_CPU systems).  The *_d device pointers and the
// precomputed tables of the reference are not used by this library (null):
// spectra live on the device as half spectra and the multipliers are
// computed in-kernel.  Callbacks keep the reference signature and receive
// a float2[N] array of real values (device pointer on RUN_GPU systems,
// host pointer on RUN_CPU systems).
#ifndef FIELD_H
#define FIELD_H

#include <map>
#include <random>
#include <string>
#include <vector>

#include "defines.h"

class term;

struct pres;

class evolver;

class field {
    private:
    const int sx, sy, sz;
    const float dx, dy, dz;
    const float stepqx, stepqy, stepqz;
    // for dynamic change of parameters
    std::vector<std::string> implicit_prefactor_strings;
    int device_id;   // id of this field in the evolver's device system (-1 = not attached)
    friend class evolver;

    public:
    field(int, float);
    field(int, int, float, float);
    field(int, int, int, float, float, float);
    ~field();
    void common_constructor();

    std::string name;
    std::map<std::string, int> usedParameters; // only implicits
    bool isCUDA;
    bool outputToFile;
    evolver *system_p;
    bool dynamic;
    int integrator;

    // Noise objects
    bool isNoisy;
    NoiseType noiseType;
    std::random_device rd;
    std::mt19937 rng;
    std::normal_distribution<> dist;

    float2 *real_array;
    float2 *comp_array;
    float2 *real_array_d;
    float2 *comp_array_d;
    float2 **terms_h;
    float2 **terms_d;

    bool needsaliasing;
    int aliasing_order;
    float2 *comp_dealiased;
    float2 *real_dealiased;
    float2 *comp_dealiased_d;
    float2 *real_dealiased_d;

    float *gen_noise;
    float2 *noise_comp;
    float2 *noise_gend;
    float2 *noise_real;
    float2 *noise_fourier;
    float *noise_comp_d_r;
    float *noise_comp_d_i;

    std::vector<term *> terms;
    std::vector<pres> implicit;

    float *precomp_implicit;
    float *precomp_implicit_d;

    pres noise_amplitude;
    float *precomp_noise;
    float *precomp_noise_d;

    pres *implicit_terms;

    // callback functions
    bool hasCB;
    void (*callback) (evolver *, float2 *, int, int, int);
    bool hasCBFourier;
    void (*callbackFourier) (evolver *, float2 *, int, int, int);
    // Device-side callbacks, the fast path of the two above (no counterpart
    // in the reference): they get the library's own device buffers -- the
    // real values as float[sz][sy][sx] (which = 0: the field, 1: its
    // dealiased copy), or the half spectrum float2[sz][sy][sx/2+1] -- and
    // the evolver's HIP stream, on which they must launch their work.  No
    // float2 copies and no device synchronisation per call.
    bool hasCBDevice;
    void (*callbackDevice) (evolver *, float *, int, int, int, int, hipStream_t);
    bool hasCBFourierDevice;
    void (*callbackFourierDevice) (evolver *, float2 *, int, int, int, hipStream_t);

    dim3 threads_per_block;
    dim3 blocks;

    // The per-field step functions of the reference.  In this library a step
    // runs for the whole system on the device (evolver::advanceTime); these
    // keep their names for source compatibility and act on the host copies
    // or forward to the system as documented in field.cpp.
    int setRHS(float);
    int updateTerms();
    void createNoise();
    void setToZero();
    void setNotDynamic(float);
    void setDynamic(float);

    void stepEuler(float);
    void stepRK2(float);
    void stepRK4(float);

    void toReal();
    void toComp();
    void normalize();
    void dealias();

    void copyHostToDevice();
    void copyDeviceToHost();
    void copyRealHostToDevice();
    void copyRealDeviceToHost();

    void writeToFile(int , int , int );

    void prepareDevice();

    void precalculateImplicit(float dt);

    float getStepqx();
    float getStepqy();
    float getStepqz();

    int addImplicitString(const std::string &);
    void printImplicitString();
    int updateParameter(const std::string &, float);

    // Extension of this library: id of the field in the device system.
    int deviceId() const { return device_id; }
};

#endif // FIELD_H
