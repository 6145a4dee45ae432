// ep_runtime.h -- the HIP runtime for libelastoplasm_hip.so; the tests'
// CPU checking build (-DEP_CPU_CHECK, see ep_cpu_check.h) substitutes a
// host-thread runtime to run the same kernel sources under AddressSanitizer.
#pragma once
#ifdef EP_CPU_CHECK
#include "ep_cpu_check.h"
#define EP_BLOCK 64
#define EP_LDS
inline void lds_atomic_add(double *p, double v) { std::atomic_ref<double>(*p).fetch_add(v); }
#define EP_MFMA_32X32X8_BF16(a, b, c) emu_mfma_32x32x8bf16(a, b, c)
#define EP_MFMA_32X32X16_BF16(a, b, c) emu_mfma_32x32x16bf16(a, b, c)
// bf16 bits of f rounded to nearest even (quiet NaN for NaN), 2^x, 1/x
inline unsigned short ep_bf16_bits(float f) {
    unsigned u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
inline float ep_exp2(float x) { return std::exp2(x); }
inline float ep_rcp(float x) { return 1.0f / x; }
inline int uniform(int v) { return v; }
#define EP_WAVES_PER_EU(n)
#else
#include <hip/hip_runtime.h>
#define EP_BLOCK 256
// LDS-qualified pointers make every block-local access a ds_* instruction
// (a generic pointer into LDS compiles to slower flat instructions).
#define EP_LDS __attribute__((address_space(3)))
__device__ __forceinline__ void lds_atomic_add(EP_LDS double *p, double v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// register budget of a kernel: at least n waves per SIMD (at most 512/n VGPRs)
#define EP_WAVES_PER_EU(n) __attribute__((amdgpu_waves_per_eu(n)))
#define EP_LAUNCH(kern, grid, block, stream, ...) hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, stream, __VA_ARGS__)
// a value every lane of the wave holds alike, moved to a scalar register
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
// MFMA operand types: 4 bf16 (as raw 16-bit patterns) and 16 FP32 accumulators
typedef short v4s __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));
#define EP_MFMA_32X32X8_BF16(a, b, c) __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0)
// gfx950: K = 16 per instruction, 8 bf16 per lane and operand, twice the
// work of the 32x32x8 form in the same issue slots
#define EP_MFMA_32X32X16_BF16(a, b, c) \
    __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(v8bf, (a)), __builtin_bit_cast(v8bf, (b)), c, 0, 0, 0)
// bf16 bits of f rounded to nearest even: gfx950's v_cvt_pk_bf16_f32
__device__ __forceinline__ unsigned short ep_bf16_bits(float f) { return __builtin_bit_cast(unsigned short, (__bf16)f); }
// 2^x and 1/x by the transcendental unit (v_exp_f32, v_rcp_f32: about 1 ulp)
__device__ __forceinline__ float ep_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float ep_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
#endif
