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

// v, hidden from the GPU compiler's common-subexpression elimination (an
// empty asm that may have changed the register; it emits no instruction):
// an address formed from it is formed again where it is used instead of
// being kept, or spilled to private memory, across a loop.  Identity on the
// host.
CF_HD inline int cf_opaque(int v) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(v));
#endif
    return v;
}
