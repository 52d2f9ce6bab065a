// Basis-series and operand preparation around the fused GEMM (the reference's fp_filter.py basis_series :597-613
// and stacked_A :615-621), in the layouts the MFMA kernel reads.
//
// Layouts. A tile's r complex basis filters are padded to rp (a multiple of 16) and stored as
// Vc[j, k] = conj(Vh[j, k]) / sw[k] (rp, flen), zero rows j >= r. Per segment the transform input is
// q[j, k] = Vc[j, k] D[k] for k < flen and 0 above (rp, tlen; the zero tail is written once and kept): one batched
// inverse FFT along k gives the basis series Zc (rp, tlen). pack_zt transposes the analysed span through LDS into
// the GEMM's time-major operand Zt[t, 2j] = Re Zc[j, a0 + t], Zt[t, 2j + 1] = Im Zc[j, a0 + t] (fp32 or bf16),
// zero-padding rows up to a multiple of the GEMM's 128-sample tile.
#include "common.h"

namespace {

constexpr int NT = 256;

// q[j, k] = Vc[j, k] * D[k] for k < flen (columns >= flen of q are left untouched: zero once, reuse)
__global__ void form_q_k(const float2* __restrict__ Vc, const long flen, const float2* __restrict__ D,
                         const long ldq, float2* __restrict__ q)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (k >= flen) return;
    const float2 v = Vc[(size_t)j * flen + k], d = D[k];
    q[(size_t)j * ldq + k] = make_float2(v.x * d.x - v.y * d.y, v.x * d.y + v.y * d.x);
}

// Zt[t, 2j] = Re Zc[j, a0 + t] cs[j], Zt[t, 2j + 1] = Im Zc[j, a0 + t] cs[j] for t < Tn; 0 for Tn <= t < Tpad.
// Tile of 64 samples x 16 basis vectors per block through LDS: coalesced reads along t, writes along j.
template <typename T>
__device__ __forceinline__ T cvt_out(const float x);
template <>
__device__ __forceinline__ float cvt_out<float>(const float x) { return x; }
template <>
__device__ __forceinline__ uint16_t cvt_out<uint16_t>(const float x) { return f32_to_bf16(x); }
template <>
__device__ __forceinline__ uint8_t cvt_out<uint8_t>(const float x) { return f32_to_fp8(x); }
template <>
__device__ __forceinline__ f16s cvt_out<f16s>(const float x) { return f16s{f32_to_f16(x)}; }

template <typename T>
__global__ void __launch_bounds__(NT) pack_zt_k(const float2* __restrict__ Zc, const long ldc, const int rp,
                                                const long a0, const long Tn, const long Tpad,
                                                const float* __restrict__ cs, T* __restrict__ Zt)
{
    __shared__ float2 tile[16][65];
    const long t0 = (long)blockIdx.x * 64;
    const int j0 = blockIdx.y * 16;
    const int K = 2 * rp;
    for (int i = threadIdx.x; i < 16 * 64; i += NT) {
        const int jl = i / 64, tl = i % 64;
        const long t = t0 + tl;
        float2 v = make_float2(0.f, 0.f);
        if (t < Tn) {
            v = Zc[(size_t)(j0 + jl) * ldc + a0 + t];
            if (cs) { const float c = cs[j0 + jl]; v.x *= c; v.y *= c; }
        }
        tile[jl][tl] = v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 16 * 64; i += NT) {
        const int tl = i / 16, jl = i % 16;
        const long t = t0 + tl;
        if (t >= Tpad) continue;
        const float2 v = tile[jl][tl];
        T* o = Zt + (size_t)t * K + 2 * (j0 + jl);
        o[0] = cvt_out<T>(v.x);
        o[1] = cvt_out<T>(v.y);
    }
}

// Ar, Ai (Mpad, K) from A (n, r) complex: As = A[m, j] rs[m] cs[j]; Ar[m, 2j] = Re As, Ar[m, 2j+1] = -Im As,
// Ai[m, 2j] = Im As, Ai[m, 2j+1] = Re As; zero for j >= r or m >= n.
template <typename T>
__device__ __forceinline__ T from_f32(const float x);
template <>
__device__ __forceinline__ float from_f32<float>(const float x) { return x; }
template <>
__device__ __forceinline__ uint16_t from_f32<uint16_t>(const float x) { return f32_to_bf16(x); }
template <>
__device__ __forceinline__ f16s from_f32<f16s>(const float x) { return f16s{f32_to_f16(x)}; }

template <typename T>
__global__ void stack_a_k(const float2* __restrict__ A, const int n, const int r, const float* __restrict__ rs,
                          const float* __restrict__ cs, const int Mpad, const int K, T* __restrict__ Ar,
                          T* __restrict__ Ai)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)Mpad * (K / 2)) return;
    const int m = (int)(g / (K / 2)), j = (int)(g % (K / 2));
    float re = 0.f, im = 0.f;
    if (m < n && j < r) {
        const float2 a = A[(size_t)m * r + j];
        const float s = rs[m] * (cs ? cs[j] : 1.f);
        re = a.x * s; im = a.y * s;
    }
    Ar[(size_t)m * K + 2 * j] = from_f32<T>(re);
    Ar[(size_t)m * K + 2 * j + 1] = from_f32<T>(-im);
    Ai[(size_t)m * K + 2 * j] = from_f32<T>(im);
    Ai[(size_t)m * K + 2 * j + 1] = from_f32<T>(re);
}

// fp8 / fp16 operands for one template row per block: values As = A[m, j] rs[m] cs[j] (Re/Im stacked as stack_a_k),
// a row scale ra = TGT / max |As| (224: half the e4m3 range; 1024 for fp16), operands As ra in the output type, and
// pw_scale[m] = 1 / ra^2, which the GEMM applies to the row's |z|^2.
template <typename T, int TGT>
__global__ void __launch_bounds__(NT) stack_a_scaled_k(const float2* __restrict__ A, const int n, const int r,
                                                       const float* __restrict__ rs, const float* __restrict__ cs,
                                                       const int K, T* __restrict__ Ar, T* __restrict__ Ai,
                                                       float* __restrict__ pw_scale)
{
    const int m = blockIdx.x;
    float mx = 0.f;
    for (int j = threadIdx.x; j < r && m < n; j += NT) {
        const float2 a = A[(size_t)m * r + j];
        const float s = rs[m] * (cs ? cs[j] : 1.f);
        mx = fmaxf(mx, fmaxf(fabsf(a.x * s), fabsf(a.y * s)));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    __shared__ float sh[NT / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
    __syncthreads();
    mx = fmaxf(fmaxf(sh[0], sh[1]), fmaxf(sh[2], sh[3]));
    const float ra = mx > 0.f ? (float)TGT / mx : 1.f;
    if (threadIdx.x == 0) pw_scale[m] = 1.f / (ra * ra);
    for (int j = threadIdx.x; j < K / 2; j += NT) {
        float re = 0.f, im = 0.f;
        if (m < n && j < r) {
            const float2 a = A[(size_t)m * r + j];
            const float s = rs[m] * (cs ? cs[j] : 1.f) * ra;
            re = a.x * s; im = a.y * s;
        }
        Ar[(size_t)m * K + 2 * j] = cvt_out<T>(re);
        Ar[(size_t)m * K + 2 * j + 1] = cvt_out<T>(-im);
        Ai[(size_t)m * K + 2 * j] = cvt_out<T>(im);
        Ai[(size_t)m * K + 2 * j + 1] = cvt_out<T>(re);
    }
}

// Per basis vector j: sum over t in [a0, a0 + Tn) of |Zc[j, t]|^2 in double (the rms scale of fp_filter.py:609-610
// and per-basis scales). One block per basis vector.
__global__ void __launch_bounds__(NT) basis_power_k(const float2* __restrict__ Zc, const long ldc, const long a0,
                                                    const long Tn, double* __restrict__ out)
{
    const int j = blockIdx.x;
    double s = 0.0;
    for (long t = threadIdx.x; t < Tn; t += NT) {
        const float2 v = Zc[(size_t)j * ldc + a0 + t];
        s += (double)v.x * v.x + (double)v.y * v.y;
    }
    s = wave_sum_f64(s);
    __shared__ double sh[NT / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[j] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

} // namespace

GWGS_API int gwgs_form_q(const void* Vc, int rp, long flen, const void* D, long ldq, void* q, void* stream)
{
    GWGS_CHECK(Vc && D && q && rp > 0 && flen > 0 && ldq >= flen, GWGS_EINVAL);
    hipLaunchKernelGGL(form_q_k, dim3((unsigned)((flen + NT - 1) / NT), rp), dim3(NT), 0, (hipStream_t)stream,
                       (const float2*)Vc, flen, (const float2*)D, ldq, (float2*)q);
    return gwgs_status();
}

// dtype 0: float32 Zt, 1: bfloat16, 2: fp8 e4m3 (values clamped to +-448), 3: fp16. Zc (rp, ldc) complex64; Zt
// (Tpad, 2 rp).
// colscale: per basis vector or null.
GWGS_API int gwgs_pack_zt(int dtype, const void* Zc, long ldc, int rp, long a0, long Tn, long Tpad,
                          const float* colscale, void* Zt, void* stream)
{
    GWGS_CHECK(Zc && Zt && rp > 0 && rp % 16 == 0 && a0 >= 0 && Tn >= 0 && Tpad >= Tn && a0 + Tn <= ldc,
               GWGS_EINVAL);
    if (Tpad == 0) return GWGS_OK;
    const dim3 grid((unsigned)((Tpad + 63) / 64), rp / 16);
    if (dtype == 3)
        hipLaunchKernelGGL(pack_zt_k<f16s>, grid, dim3(NT), 0, (hipStream_t)stream, (const float2*)Zc, ldc, rp, a0,
                           Tn, Tpad, colscale, (f16s*)Zt);
    else if (dtype == 2)
        hipLaunchKernelGGL(pack_zt_k<uint8_t>, grid, dim3(NT), 0, (hipStream_t)stream, (const float2*)Zc, ldc, rp, a0,
                           Tn, Tpad, colscale, (uint8_t*)Zt);
    else if (dtype == 1)
        hipLaunchKernelGGL(pack_zt_k<uint16_t>, grid, dim3(NT), 0, (hipStream_t)stream, (const float2*)Zc, ldc, rp, a0,
                           Tn, Tpad, colscale, (uint16_t*)Zt);
    else if (dtype == 0)
        hipLaunchKernelGGL(pack_zt_k<float>, grid, dim3(NT), 0, (hipStream_t)stream, (const float2*)Zc, ldc, rp, a0, Tn,
                           Tpad, colscale, (float*)Zt);
    else
        return GWGS_EUNSUP;
    return gwgs_status();
}

GWGS_API int gwgs_stack_a(int dtype, const void* A, int n, int r, const float* rowscale, const float* colscale,
                          int Mpad, int K, void* Ar, void* Ai, void* stream)
{
    GWGS_CHECK(A && rowscale && Ar && Ai && n >= 0 && r > 0 && Mpad >= n && K >= 2 * r && K % 2 == 0, GWGS_EINVAL);
    const long tot = (long)Mpad * (K / 2);
    if (tot == 0) return GWGS_OK;
    const dim3 grid((unsigned)((tot + NT - 1) / NT));
    if (dtype == 3)
        hipLaunchKernelGGL(stack_a_k<f16s>, grid, dim3(NT), 0, (hipStream_t)stream, (const float2*)A, n, r,
                           rowscale, colscale, Mpad, K, (f16s*)Ar, (f16s*)Ai);
    else if (dtype == 1)
        hipLaunchKernelGGL(stack_a_k<uint16_t>, grid, dim3(NT), 0, (hipStream_t)stream, (const float2*)A, n, r,
                           rowscale, colscale, Mpad, K, (uint16_t*)Ar, (uint16_t*)Ai);
    else if (dtype == 0)
        hipLaunchKernelGGL(stack_a_k<float>, grid, dim3(NT), 0, (hipStream_t)stream, (const float2*)A, n, r, rowscale,
                           colscale, Mpad, K, (float*)Ar, (float*)Ai);
    else
        return GWGS_EUNSUP;
    return gwgs_status();
}

// fp8 operands: Ar, Ai (n, K) bytes and pw_scale (n floats); see stack_a_scaled_k.
GWGS_API int gwgs_stack_a_fp8(const void* A, int n, int r, const float* rowscale, const float* colscale, int K,
                              void* Ar, void* Ai, float* pw_scale, void* stream)
{
    GWGS_CHECK(A && rowscale && Ar && Ai && pw_scale && n >= 0 && r > 0 && K >= 2 * r && K % 2 == 0, GWGS_EINVAL);
    if (n == 0) return GWGS_OK;
    hipLaunchKernelGGL((stack_a_scaled_k<uint8_t, 224>), dim3(n), dim3(NT), 0, (hipStream_t)stream, (const float2*)A,
                       n, r, rowscale, colscale, K, (uint8_t*)Ar, (uint8_t*)Ai, pw_scale);
    return gwgs_status();
}

// fp16 operands with a per-template scale to max |As| = 1024 (the GEMM's fp16 form): as gwgs_stack_a_fp8.
GWGS_API int gwgs_stack_a_f16(const void* A, int n, int r, const float* rowscale, const float* colscale, int K,
                              void* Ar, void* Ai, float* pw_scale, void* stream)
{
    GWGS_CHECK(A && rowscale && Ar && Ai && pw_scale && n >= 0 && r > 0 && K >= 2 * r && K % 2 == 0, GWGS_EINVAL);
    if (n == 0) return GWGS_OK;
    hipLaunchKernelGGL((stack_a_scaled_k<f16s, 1024>), dim3(n), dim3(NT), 0, (hipStream_t)stream, (const float2*)A,
                       n, r, rowscale, colscale, K, (f16s*)Ar, (f16s*)Ai, pw_scale);
    return gwgs_status();
}

GWGS_API int gwgs_basis_power(const void* Zc, long ldc, int rp, long a0, long Tn, double* out, void* stream)
{
    GWGS_CHECK(Zc && out && rp > 0 && a0 >= 0 && Tn > 0 && a0 + Tn <= ldc, GWGS_EINVAL);
    hipLaunchKernelGGL(basis_power_k, dim3(rp), dim3(NT), 0, (hipStream_t)stream, (const float2*)Zc, ldc, a0, Tn, out);
    return gwgs_status();
}
