// Data conditioning kernels (the reference's gpu_prep.py:87-231, 405-454): Welch PSD with median average in
// float64, interpolation, inverse-spectrum truncation, whitening D = stilde / psd and IP = 1 / psd, and the
// autogating whitening and inverted-Tukey gates. FFTs are hipFFT calls made by the host; these kernels are the
// memory-bound passes between them.
#include "common.h"

namespace {

constexpr int NT = 256;

// segs[r, s, i] = x[r, start + s stride + i] * w[i], w the symmetric Hann window 0.5 - 0.5 cos(2 pi i / (L - 1))
__global__ void welch_segments_k(const float* __restrict__ x, const long ldx, const long start, const int nseg,
                                 const int L, const int stride, double* __restrict__ segs)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (g >= (long)nseg * L) return;
    const int s = (int)(g / L), i = (int)(g % L);
    const double w = 0.5 - 0.5 * cospi(2.0 * (double)i / (double)(L - 1));
    segs[((size_t)r * nseg + s) * L + i] = (double)x[(size_t)r * ldx + start + (size_t)s * stride + i] * w;
}

// Per (row, bin): p_s = |X_s|^2 dt^2 (halved at DC and Nyquist), median over the nseg segments (numpy's: the middle
// order statistic, or the mean of the two middle ones), / bias, * fac; or the mean (summed in segment order).
// A block of 256 threads takes WB = 8192 / NP bins: their nseg <= NP values are staged in LDS (64 KB, bins fastest, so
// a row of segments is read coalesced), padded to NP with +inf, and each bin's column is sorted by a bitonic network
// (NP / 2 compare-exchanges per stage over all WB columns at once, one barrier per stage). The order statistics are
// the same values whatever the algorithm, so the result is the one the selection by counting gave.
template <int NP>
__global__ void __launch_bounds__(256) welch_reduce_k(const double2* __restrict__ X, const int nseg, const int nf,
                                                     const double dt2, const int mean, const double bias,
                                                     const double fac, double* __restrict__ out)
{
    constexpr int WB = 8192 / NP;
    __shared__ double v[NP * WB];               // v[s WB + bin]
    const int r = blockIdx.y, b0 = blockIdx.x * WB, tid = threadIdx.x;
    const double2* __restrict__ xr = X + (size_t)r * nseg * nf;
    for (int e = tid; e < NP * WB; e += 256) {
        const int s = e / WB, b = b0 + e % WB;
        const bool ok = s < nseg && b < nf;
        const double2 z = xr[ok ? (size_t)s * nf + b : 0];     // unconditional load of a valid element
        double p = (z.x * z.x + z.y * z.y) * dt2;
        if (b == 0 || b == nf - 1) p /= 2;
        v[e] = ok ? p : s < nseg ? 0.0 : __builtin_inf();
    }
    __syncthreads();
    if (mean) {
        if (tid < WB && b0 + tid < nf) {
            double acc = 0.0;
            for (int s = 0; s < nseg; s++) acc += v[s * WB + tid];
            out[(size_t)r * nf + b0 + tid] = acc / nseg * fac;
        }
        return;
    }
    for (int k = 2; k <= NP; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int e = tid; e < (NP / 2) * WB; e += 256) {
                const int bl = e % WB, pi = e / WB;
                const int i = 2 * j * (pi / j) + pi % j, l = i + j;    // pair (i, i + j), bit j of i clear
                const double a = v[i * WB + bl], c = v[l * WB + bl];
                if ((a > c) == ((i & k) == 0)) {
                    v[i * WB + bl] = c;
                    v[l * WB + bl] = a;
                }
            }
            __syncthreads();
        }
    if (tid < WB && b0 + tid < nf) {
        const int m = nseg / 2;
        double res = (nseg & 1) ? v[m * WB + tid] : 0.5 * (v[(m - 1) * WB + tid] + v[m * WB + tid]);
        res /= bias;
        out[(size_t)r * nf + b0 + tid] = res * fac;
    }
}

// The same reduction for more than 512 segments (the autogating Welch over hours of data): one workgroup per (bin,
// row), the bin's nseg powers in LDS padded with +inf to NP, a bitonic sort over all NP (NP / 2 compare-exchanges
// per stage over 1024 threads, one barrier per stage), then the same median (or the mean in segment order).
template <int NP>
__global__ void __launch_bounds__(1024) welch_reduce_big_k(const double2* __restrict__ X, const int nseg,
                                                          const int nf, const double dt2, const int mean,
                                                          const double bias, const double fac,
                                                          double* __restrict__ out)
{
    __shared__ double v[NP];
    const int b = blockIdx.x, r = blockIdx.y, tid = threadIdx.x;
    const double2* __restrict__ xr = X + (size_t)r * nseg * nf;
    for (int s = tid; s < NP; s += 1024) {
        double p = __builtin_inf();
        if (s < nseg) {
            const double2 z = xr[(size_t)s * nf + b];
            p = (z.x * z.x + z.y * z.y) * dt2;
            if (b == 0 || b == nf - 1) p /= 2;
        }
        v[s] = p;
    }
    __syncthreads();
    if (mean) {
        if (tid == 0) {
            double acc = 0.0;
            for (int s = 0; s < nseg; s++) acc += v[s];
            out[(size_t)r * nf + b] = acc / nseg * fac;
        }
        return;
    }
    for (int k = 2; k <= NP; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int pi = tid; pi < NP / 2; pi += 1024) {
                const int i = 2 * j * (pi / j) + pi % j, l = i + j;
                const double a = v[i], c = v[l];
                if ((a > c) == ((i & k) == 0)) {
                    v[i] = c;
                    v[l] = a;
                }
            }
            __syncthreads();
        }
    if (tid == 0) {
        const int m = nseg / 2;
        double res = (nseg & 1) ? v[m] : 0.5 * (v[m - 1] + v[m]);
        res /= bias;
        out[(size_t)r * nf + b] = res * fac;
    }
}

// np.interp of P (rows, nold) sampled at k df_old onto i df_new, i < nnew (right end held)
__global__ void interp_k(const double* __restrict__ P, const int nold, const double df_old, const double df_new,
                         const long nnew, double* __restrict__ out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (i >= nnew) return;
    const double* p = P + (size_t)r * nold;
    const double x = (double)i * df_new;
    const double u = x / df_old;
    long j = (long)floor(u);
    double y;
    if (j >= nold - 1) y = p[nold - 1];
    else {
        if (j < 0) j = 0;
        const double x0 = (double)j * df_old, x1 = (double)(j + 1) * df_old;
        const double t = (x - x0) / (x1 - x0);
        y = p[j] + (p[j + 1] - p[j]) * t;
    }
    out[(size_t)r * nnew + i] = y;
}

// inv_asd = P^-1/2 on [kmin, khi), 0 elsewhere, as the complex input of an inverse real FFT
__global__ void isqrt_band_k(const double* __restrict__ P, const long flen, const long kmin, const long khi,
                             double2* __restrict__ out)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (k >= flen) return;
    double v = 0.0;
    if (k >= kmin && k < khi) v = 1.0 / sqrt(P[(size_t)r * flen + k]);
    out[(size_t)r * flen + k] = make_double2(v, 0.0);
}

// q *= df; zero [L/2, N - L/2); with hann, taper the kept ends by the halves of a symmetric Hann window of length L
__global__ void trunc_k(double* __restrict__ q, const long N, const long L, const int hann, const double df)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (i >= N) return;
    const long ts = L / 2, te = N - L / 2;
    double v = q[(size_t)r * N + i] * df;
    if (hann) {
        if (i < ts) v *= 0.5 - 0.5 * cospi(2.0 * (double)(L - ts + i) / (double)(L - 1));
        else if (i >= te) v *= 0.5 - 0.5 * cospi(2.0 * (double)(i - te) / (double)(L - 1));
    }
    if (ts < te && i >= ts && i < te) v = 0.0;
    q[(size_t)r * N + i] = v;
}

// out = 1 / |X / (df N)|^2
__global__ void inv_abs2_k(const double2* __restrict__ X, const long flen, const double scale, double* __restrict__ out)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (k >= flen) return;
    const double2 z = X[(size_t)r * flen + k];
    const double a = z.x * scale, b = z.y * scale;
    out[(size_t)r * flen + k] = 1.0 / (a * a + b * b);
}

// D[s, k] = (stilde[s, k] * (1/sr as float32)) * ip[s, k] rounded to complex64, ip = 1/P[pick[s], k] (0 where P <= 0
// or k < kmin), IP = float32(ip)
__global__ void whiten_k(const float2* __restrict__ X, const float inv_sr, const double* __restrict__ P,
                         const int* __restrict__ pick, const long flen, const long kmin, float2* __restrict__ D,
                         float* __restrict__ IP)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (k >= flen) return;
    const double p = P[(size_t)pick[s] * flen + k];
    const double ip = (k >= kmin && p > 0.0) ? 1.0 / p : 0.0;
    const float2 x = X[(size_t)s * flen + k];
    const float xr = x.x * inv_sr, xi = x.y * inv_sr;
    D[(size_t)s * flen + k] = make_float2((float)((double)xr * ip), (float)((double)xi * ip));
    IP[(size_t)s * flen + k] = (float)ip;
}

// Autogating: taper the corrupt ends (ramp i / cl) of s and centre it in a zero buffer of length Lp (double)
__global__ void gate_pad_k(const float* __restrict__ s, const long n, const long cl, const long pad_start,
                           const long Lp, double* __restrict__ out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Lp) return;
    double v = 0.0;
    const long j = i - pad_start;
    if (j >= 0 && j < n) {
        v = (double)s[j];
        if (j < cl) v *= (double)j / (double)cl;
        else if (j >= n - cl) v *= (double)(n - 1 - j) / (double)cl;
    }
    out[i] = v;
}

// st[k] *= dt (psd[k] norm)^-1/2, 0 where psd is +inf (k < kmin)
__global__ void gate_whiten_k(double2* __restrict__ st, const double* __restrict__ psd, const long flen,
                              const long kmin, const double dt, const double norm)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= flen) return;
    double2 z = st[k];
    const double f = k < kmin ? 0.0 : dt / sqrt(psd[k] * norm);
    st[k] = make_double2(z.x * f, z.y * f);
}

// mag[j] = |wt[pad_start + j] * df| for j in [cl, n - cl), 0 at the corrupt ends (float32 for the compaction)
__global__ void gate_mag_k(const double* __restrict__ wt, const long pad_start, const long n, const long cl,
                           const double df, float* __restrict__ mag)
{
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    mag[j] = (j < cl || j >= n - cl) ? 0.f : (float)fabs(wt[pad_start + j] * df);
}

// x[offset + i] *= window[i] for i in [i1, i2) (product in double, rounded to float32); window = inverted Tukey:
// cosine ramps of pad samples around win - 2 pad zeros (gpu_prep.py:212-231)
__global__ void gate_apply_k(float* __restrict__ x, const long offset, const long i1, const long i2, const long pad,
                             const long win)
{
    const long i = i1 + (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= i2) return;
    double w;
    if (i < pad) w = 0.5 * (1.0 + cospi((double)i / (double)pad));
    else if (i >= win - pad) w = 0.5 * (1.0 + cospi((double)(win - 1 - i) / (double)pad));
    else w = 0.0;
    x[offset + i] = (float)((double)x[offset + i] * w);
}

inline dim3 grid1(long n) { return dim3((unsigned)((n + NT - 1) / NT)); }
inline dim3 grid2(long n, int rows) { return dim3((unsigned)((n + NT - 1) / NT), rows); }

// 'ldas' FIR decimation (pycbc resample_to_delta_t, gpu_prep.py:67-84): out[i] = sum_k c[k] x[n - k] with
// n = factor i + shift, zero where n < 2 (ntaps / 2) or n >= L (the filter's start-up samples, and the wrapped
// indices of the rolled series that land there); float64 taps and sum in tap order without fused multiply-add (the
// reference's float64 numpy arithmetic, bit for bit), one thread per output sample.
__global__ void fir_decimate_k(const double* __restrict__ x, const long L, const double* __restrict__ c,
                               const int ntaps, const int factor, const long shift, const long nout,
                               double* __restrict__ out)
{
#pragma clang fp contract(off)
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nout) return;
    const long n = (long)factor * i + shift;
    double acc = 0.0;
    if (n >= 2L * (ntaps / 2) && n < L)
        for (int k = 0; k < ntaps; k++) acc += c[k] * x[n - k];          // contraction off: as numpy
    out[i] = acc;
}

} // namespace

GWGS_API int gwgs_welch_segments(const float* x, long ldx, int rows, long start, int nseg, int L, int stride,
                                 double* segs, void* stream)
{
    GWGS_CHECK(x && segs && rows > 0 && nseg > 0 && L > 1 && stride > 0 && start >= 0, GWGS_EINVAL);
    hipLaunchKernelGGL(welch_segments_k, grid2((long)nseg * L, rows), dim3(NT), 0, (hipStream_t)stream, x, ldx, start,
                       nseg, L, stride, segs);
    return gwgs_status();
}

GWGS_API int gwgs_welch_reduce(const void* X, int rows, int nseg, int nf, double dt, int mean, double bias, double fac,
                               double* out, void* stream)
{
    GWGS_CHECK(X && out && rows > 0 && nseg > 0 && nf > 0, GWGS_EINVAL);
    GWGS_CHECK(nseg <= 16384, GWGS_ESIZE);
    const double2* x = (const double2*)X;
    hipStream_t st = (hipStream_t)stream;
    const double dt2 = dt * dt;
    if (nseg > 512) {
        const dim3 g((unsigned)nf, (unsigned)rows);
        if (nseg <= 1024)
            hipLaunchKernelGGL(welch_reduce_big_k<1024>, g, dim3(1024), 0, st, x, nseg, nf, dt2, mean, bias, fac, out);
        else if (nseg <= 2048)
            hipLaunchKernelGGL(welch_reduce_big_k<2048>, g, dim3(1024), 0, st, x, nseg, nf, dt2, mean, bias, fac, out);
        else if (nseg <= 4096)
            hipLaunchKernelGGL(welch_reduce_big_k<4096>, g, dim3(1024), 0, st, x, nseg, nf, dt2, mean, bias, fac, out);
        else if (nseg <= 8192)
            hipLaunchKernelGGL(welch_reduce_big_k<8192>, g, dim3(1024), 0, st, x, nseg, nf, dt2, mean, bias, fac, out);
        else
            hipLaunchKernelGGL(welch_reduce_big_k<16384>, g, dim3(1024), 0, st, x, nseg, nf, dt2, mean, bias, fac,
                               out);
        return gwgs_status();
    }
    if (nseg <= 64)
        hipLaunchKernelGGL(welch_reduce_k<64>, dim3((nf + 127) / 128, rows), dim3(256), 0, st, x, nseg, nf, dt2, mean,
                           bias, fac, out);
    else if (nseg <= 128)
        hipLaunchKernelGGL(welch_reduce_k<128>, dim3((nf + 63) / 64, rows), dim3(256), 0, st, x, nseg, nf, dt2, mean,
                           bias, fac, out);
    else if (nseg <= 256)
        hipLaunchKernelGGL(welch_reduce_k<256>, dim3((nf + 31) / 32, rows), dim3(256), 0, st, x, nseg, nf, dt2, mean,
                           bias, fac, out);
    else
        hipLaunchKernelGGL(welch_reduce_k<512>, dim3((nf + 15) / 16, rows), dim3(256), 0, st, x, nseg, nf, dt2, mean,
                           bias, fac, out);
    return gwgs_status();
}

GWGS_API int gwgs_interp(const double* P, int rows, int nold, double df_old, double df_new, long nnew, double* out,
                         void* stream)
{
    GWGS_CHECK(P && out && rows > 0 && nold > 1 && nnew > 0 && df_old > 0 && df_new > 0, GWGS_EINVAL);
    hipLaunchKernelGGL(interp_k, grid2(nnew, rows), dim3(NT), 0, (hipStream_t)stream, P, nold, df_old, df_new, nnew,
                       out);
    return gwgs_status();
}

GWGS_API int gwgs_isqrt_band(const double* P, int rows, long flen, long kmin, long khi, void* out, void* stream)
{
    GWGS_CHECK(P && out && rows > 0 && flen > 0, GWGS_EINVAL);
    hipLaunchKernelGGL(isqrt_band_k, grid2(flen, rows), dim3(NT), 0, (hipStream_t)stream, P, flen, kmin, khi,
                       (double2*)out);
    return gwgs_status();
}

GWGS_API int gwgs_trunc(double* q, int rows, long N, long L, int hann, double df, void* stream)
{
    GWGS_CHECK(q && rows > 0 && N > 0 && L > 1, GWGS_EINVAL);
    hipLaunchKernelGGL(trunc_k, grid2(N, rows), dim3(NT), 0, (hipStream_t)stream, q, N, L, hann, df);
    return gwgs_status();
}

GWGS_API int gwgs_inv_abs2(const void* X, int rows, long flen, double scale, double* out, void* stream)
{
    GWGS_CHECK(X && out && rows > 0 && flen > 0, GWGS_EINVAL);
    hipLaunchKernelGGL(inv_abs2_k, grid2(flen, rows), dim3(NT), 0, (hipStream_t)stream, (const double2*)X, flen, scale,
                       out);
    return gwgs_status();
}

GWGS_API int gwgs_whiten(const void* X, float inv_sr, const double* P, const int* pick, int nseg, long flen, long kmin,
                         void* D, float* IP, void* stream)
{
    GWGS_CHECK(X && P && pick && D && IP && nseg > 0 && flen > 0, GWGS_EINVAL);
    hipLaunchKernelGGL(whiten_k, grid2(flen, nseg), dim3(NT), 0, (hipStream_t)stream, (const float2*)X, inv_sr, P,
                       pick, flen, kmin, (float2*)D, IP);
    return gwgs_status();
}

GWGS_API int gwgs_gate_pad(const float* s, long n, long cl, long pad_start, long Lp, double* out, void* stream)
{
    GWGS_CHECK(s && out && n > 0 && Lp >= n && pad_start >= 0 && pad_start + n <= Lp, GWGS_EINVAL);
    hipLaunchKernelGGL(gate_pad_k, grid1(Lp), dim3(NT), 0, (hipStream_t)stream, s, n, cl, pad_start, Lp, out);
    return gwgs_status();
}

GWGS_API int gwgs_gate_whiten(void* st, const double* psd, long flen, long kmin, double dt, double norm, void* stream)
{
    GWGS_CHECK(st && psd && flen > 0, GWGS_EINVAL);
    hipLaunchKernelGGL(gate_whiten_k, grid1(flen), dim3(NT), 0, (hipStream_t)stream, (double2*)st, psd, flen, kmin,
                       dt, norm);
    return gwgs_status();
}

GWGS_API int gwgs_gate_mag(const double* wt, long pad_start, long n, long cl, double df, float* mag, void* stream)
{
    GWGS_CHECK(wt && mag && n > 0, GWGS_EINVAL);
    hipLaunchKernelGGL(gate_mag_k, grid1(n), dim3(NT), 0, (hipStream_t)stream, wt, pad_start, n, cl, df, mag);
    return gwgs_status();
}

GWGS_API int gwgs_gate_apply(float* x, long n, long offset, long i1, long i2, long pad, long win, void* stream)
{
    GWGS_CHECK(x && pad > 0 && win >= 2 * pad && i1 >= 0 && i2 <= win && offset + i1 >= 0 && offset + i2 <= n,
               GWGS_EINVAL);
    if (i2 <= i1) return GWGS_OK;
    hipLaunchKernelGGL(gate_apply_k, grid1(i2 - i1), dim3(NT), 0, (hipStream_t)stream, x, offset, i1, i2, pad, win);
    return gwgs_status();
}

// Decimation by `factor` of x (L samples, float64) with the ntaps FIR taps c (float64, device): out has
// ceil(L / factor) samples (gpu_prep.py:67-84 resample_ldas, the shift -ntaps // 2 rounded as Python does).
GWGS_API int gwgs_fir_decimate(const double* x, long L, const double* c, int ntaps, int factor, double* out,
                               void* stream)
{
    GWGS_CHECK(x && c && out && L > 0 && ntaps > 0 && factor > 0, GWGS_EINVAL);
    const long nout = (L + factor - 1) / factor;
    const long shift = (ntaps + 1) / 2;      // -(-ntaps // 2)
    hipLaunchKernelGGL(fir_decimate_k, grid1(nout), dim3(NT), 0, (hipStream_t)stream, x, L, c, ntaps, factor, shift,
                       nout, out);
    return gwgs_status();
}
