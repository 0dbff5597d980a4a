// cupss.h -- source-compatible umbrella header of the MI355X build of cuPSS
// (reference: inc/cupss.h:1-6).  A solver written against the reference
// (#include "cupss.h", evolver, createField, addEquation, advanceTime, ...)
// recompiles with hipcc/amdclang++ against these headers and links
// libcupss.so.  The HIP runtime types float2 and dim3 replace CUDA's.
#include "cupss/defines.h"
#include "cupss/cu_utils.h"
#include "cupss/evolver.h"
#include "cupss/field.h"
#include "cupss/term.h"
#include "cupss/parser.h"
