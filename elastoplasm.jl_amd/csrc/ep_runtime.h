// ep_runtime.h -- the HIP runtime for libelastoplasm_hip.so; the tests'
// CPU checking build (-DEP_CPU_CHECK, see ep_cpu_check.h) substitutes a
// host-thread runtime to run the same kernel sources under AddressSanitizer.
#pragma once
#ifdef EP_CPU_CHECK
#include "ep_cpu_check.h"
#define EP_BLOCK 64
#else
#define EP_BLOCK 256
#include <hip/hip_runtime.h>
#define EP_LAUNCH(kern, grid, block, stream, ...) hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, stream, __VA_ARGS__)
#endif
