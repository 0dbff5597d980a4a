// term.h -- one right-hand-side group: a product of fields with summed
// spectral prefactors (reference: inc/cupss/term.h:14-70).  In this library a
// term is evaluated on the device by the system's step program; the public
// members keep the reference names (device pointers and tables are null).
#ifndef TERM_H
#define TERM_H

#include <map>
#include <string>
#include <vector>

#include "defines.h"

class field;
class evolver;

class term {
private:
    const int sx, sy, sz;             // system size
    const float dx, dy, dz;           // cell size
    const float stepqx, stepqy, stepqz;
    int device_id;                    // id in the device system (-1 = not attached)
    evolver *owner;
    friend class evolver;

public:
    term(int, float);
    term(int, int, float, float);
    term(int, int, int, float, float, float);
    ~term();
    void common_constructor();
    int prepareDevice();
    int precomputePrefactors();
    int update();

    bool isCUDA;

    pres prefactors;
    std::vector<pres> prefactors_h;

    pres *prefactors_d;

    float2 *term_real;
    float2 *term_comp;
    float2 *term_real_d;
    float2 *term_comp_d;

    std::vector<field *> product;
    float2 **product_h;
    float2 **product_d;
    float *precomp_prefactor_h;
    float *precomp_prefactor_d;

    int multiply_by_i_pre;

    dim3 threads_per_block;
    dim3 blocks;

    // dynamic update of parameters
    std::map<std::string, int> usedParameters;
    std::vector<std::string> prefactor_strings;
    int setPrefactorString(const std::vector<std::string> &);
    void printPrefactorString();
};

#endif
