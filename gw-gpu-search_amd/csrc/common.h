// Shared helpers for the gw-gpu-search MI355X kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GWGS_OK 0
#define GWGS_EINVAL -1     // bad argument (null pointer, negative size, misaligned leading dimension)
#define GWGS_ESIZE -2      // a size beyond what the kernel supports (K too large, too many segments, ...)
#define GWGS_EUNSUP -3     // unsupported combination (dtype, mode)

#define GWGS_API extern "C" __attribute__((visibility("default")))

static inline int gwgs_status(void)
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GWGS_OK : (int)e;
}

#define GWGS_CHECK(cond, code) do { if (!(cond)) return (code); } while (0)

// first index wins on ties: (v, i) replaces (bv, bi) when v > bv, or v == bv and i < bi
__device__ __forceinline__ void take_max_first(float& bv, int& bi, const float v, const int i)
{
    if (v > bv || (v == bv && i < bi)) { bv = v; bi = i; }
}

// max and its first index over the 64 lanes of a wave (butterfly, every lane ends with the result)
__device__ __forceinline__ void wave_argmax_first(float& bv, int& bi)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float ov = __shfl_xor(bv, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        take_max_first(bv, bi, ov, oi);
    }
}

__device__ __forceinline__ double wave_sum_f64(double x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

__device__ __forceinline__ float wave_sum_f32(float x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// round-to-nearest-even float -> bfloat16 bits (NaN kept quiet)
__device__ __forceinline__ uint16_t f32_to_bf16(const float x)
{
    const uint32_t u = __float_as_uint(x);
    const uint32_t rn = (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
    const bool nan = (u & 0x7fffffffu) > 0x7f800000u;
    return (uint16_t)(nan ? ((u >> 16) | 0x40u) : rn);     // a select, not a branch
}

// float -> fp8 e4m3 (OCP, gfx950 v_cvt_pk_fp8_f32: round to nearest even), saturated to +-448 first
__device__ __forceinline__ uint8_t f32_to_fp8(float x)
{
    x = fminf(fmaxf(x, -448.f), 448.f);
    return (uint8_t)(__builtin_amdgcn_cvt_pk_fp8_f32(x, 0.f, 0, false) & 0xff);
}

__device__ __forceinline__ float fp8_to_f32(const uint8_t b)
{
    return __builtin_amdgcn_cvt_f32_fp8((int)b, 0);
}
