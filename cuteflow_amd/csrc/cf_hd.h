// Host/device qualifier: the step's numerics compile for gfx950 (hipcc) and,
// for host-side checks of the derivative code, for the CPU (g++).
#pragma once
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CF_HD __host__ __device__
#else
#include <math.h>
#define CF_HD
#endif
