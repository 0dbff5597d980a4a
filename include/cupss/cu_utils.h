// cu_utils.h -- device error check and query (reference: inc/cupss/cu_utils.h,
// src/cu_utils.cpp:4-37), on the HIP runtime.
#ifndef CU_UTILS
#define CU_UTILS

#include "defines.h"

void check_error(hipError_t err);
void check_device();

#endif
