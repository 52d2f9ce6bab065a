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

// fp16 (IEEE binary16) operand storage: a 2-byte type of its own, so that templates over the operand type take the
// f16 MFMA and conversions where they take the bf16 ones for uint16_t
struct f16s { uint16_t v; };

// float -> fp16 bits, round to nearest even (v_cvt_f16_f32); overflow goes to +-inf
__device__ __forceinline__ uint16_t f32_to_f16(const float x) { return __builtin_bit_cast(uint16_t, (_Float16)x); }
__device__ __forceinline__ uint32_t f32x2_to_f16x2(const float a, const float b)
{
    return (uint32_t)f32_to_f16(a) | ((uint32_t)f32_to_f16(b) << 16);
}
__device__ __forceinline__ float f16_to_f32(const uint16_t b) { return (float)__builtin_bit_cast(_Float16, b); }

// two floats -> a packed bfloat16 pair (a in the low half), round to nearest even: one v_cvt_pk_bf16_f32 on gfx950
__device__ __forceinline__ uint32_t f32x2_to_bf16x2(const float a, const float b)
{
    typedef float f32x2_t __attribute__((ext_vector_type(2)));
    typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
    f32x2_t v;
    v.x = a;
    v.y = b;
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_t));
}

// round-to-nearest-even float -> bfloat16 bits
__device__ __forceinline__ uint16_t f32_to_bf16(const float x)
{
    return (uint16_t)(f32x2_to_bf16x2(x, 0.f) & 0xffffu);
}

// float -> fp8 e4m3 (OCP, gfx950 v_cvt_pk_fp8_f32: round to nearest even), saturated to +-448 first
__device__ __forceinline__ uint8_t f32_to_fp8(float x)
{
    x = fminf(fmaxf(x, -448.f), 448.f);
    return (uint8_t)(__builtin_amdgcn_cvt_pk_fp8_f32(x, 0.f, 0, false) & 0xff);
}

// two floats -> two fp8 e4m3 bytes (a in the low byte), each as f32_to_fp8, in one conversion
__device__ __forceinline__ uint16_t f32x2_to_fp8x2(const float a, const float b)
{
    return (uint16_t)(__builtin_amdgcn_cvt_pk_fp8_f32(fminf(fmaxf(a, -448.f), 448.f), fminf(fmaxf(b, -448.f), 448.f),
                                                      0, false) & 0xffff);
}

__device__ __forceinline__ float fp8_to_f32(const uint8_t b)
{
    return __builtin_amdgcn_cvt_f32_fp8((int)b, 0);
}

constexpr int kRsrcWord3 = 0x00020000;   // gfx9 buffer descriptor word 3 (32-bit data format, no swizzle)

// A raw buffer of nbytes bytes at p (capped at 2^31 - 1; callers' offsets stay below that). Loads at or past its
// size return zeros and stores there are dropped by the hardware, so a range check needs no branch: a load under a
// per-lane branch makes the compiler wait for it before the next one is issued.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_of(const void* p, const long nbytes)
{
    const int nrec = nbytes <= 0 ? 0 : nbytes > 0x7fffffffL ? 0x7fffffff : (int)nbytes;
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, nrec, kRsrcWord3);
}

// 8 bytes at byte offset off of a buffer (zeros when off is at or past its size)
__device__ __forceinline__ float2 load_f2(const __amdgpu_buffer_rsrc_t rs, const unsigned off)
{
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    const u32x2 w = __builtin_amdgcn_raw_buffer_load_b64(rs, (int)off, 0, 0);
    return make_float2(__uint_as_float(w.x), __uint_as_float(w.y));
}

// 4 bytes at byte offset off of a buffer (zero when off is at or past its size)
__device__ __forceinline__ float load_f1(const __amdgpu_buffer_rsrc_t rs, const unsigned off)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, (int)off, 0, 0));
}
