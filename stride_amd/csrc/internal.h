// Accessors shared between the solver and the processing kernels (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/stride_amd.h"

hipStream_t sa_solver_stream(sa_solver *s);
int sa_solver_device(sa_solver *s);
int sa_fail(int code, const char *msg);
