// ep_runtime.h -- the HIP runtime for libelastoplasm_hip.so; the tests'
// CPU checking build (-DEP_CPU_CHECK, see ep_cpu_check.h) substitutes a
// host-thread runtime to run the same kernel sources under AddressSanitizer.
#pragma once
#ifdef EP_CPU_CHECK
#include "ep_cpu_check.h"
#define EP_BLOCK 64
#define EP_LDS
inline void lds_atomic_add(double *p, double v) { std::atomic_ref<double>(*p).fetch_add(v); }
#else
#include <hip/hip_runtime.h>
#define EP_BLOCK 256
// LDS-qualified pointers make every block-local access a ds_* instruction
// (a generic pointer into LDS compiles to slower flat instructions).
#define EP_LDS __attribute__((address_space(3)))
__device__ __forceinline__ void lds_atomic_add(EP_LDS double *p, double v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
#define EP_LAUNCH(kern, grid, block, stream, ...) hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, stream, __VA_ARGS__)
#endif
