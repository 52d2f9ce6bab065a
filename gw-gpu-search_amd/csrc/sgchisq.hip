// Sine-Gaussian chi-squared tile sums (the reference's sg_kernel.py sg_tiles and fp_filter.py:463-486
// sg_chisq_one): for every (peak, tile) b, over bins k in [kk[b], k1[b]) with k >= kg[b],
//   zsum[b] = sum_k g_k D_k e^{2 pi i (k t mod N) / N},   gsum[b] = sum_k g_k^2 IP_k,
//   g_k = amp exp(-(pi tau (f - fc))^2) (1 + exp(-q^2 f / fc)),  f = k df,  amp = sqrt(pi) / 2 tau,
// the host finishing z_j = 4 df zsum / sqrt(4 df gsum) and the value in float64. The tile ranges come from the host
// in float64 (the reference's index arithmetic, bit for bit).
//
// One 256-thread workgroup per (peak, tile): thread i takes bins kk + i, kk + i + 256, ... (coalesced loads of D
// and IP, which stay in L2 across the peaks of one segment). Each thread's phasor is seeded exactly every R of its
// bins (the 64-bit integer phase k t mod N, then sincospif(2 ph / N) in float, exact in its argument for
// power-of-two N <= 2^24) and advanced in between by the fixed rotation e^{2 pi i (256 t mod N) / N}, so its
// error stays within about R float roundings; g in float with expf; float partial sums over those R bins are added
// in float64, then reduced in a fixed order (wave butterflies, then the four waves in order): the result does not
// depend on scheduling.
#include "common.h"

namespace {

constexpr int NT = 256;
constexpr int R = 16;            // bins of one thread between exact phasor seeds

__global__ void __launch_bounds__(NT) sg_tiles_k(const float2* __restrict__ D, const float* __restrict__ IP,
                                                 const long N, const int pow2, const double df,
                                                 const long* __restrict__ kk, const long* __restrict__ k1,
                                                 const long* __restrict__ kg, const double* __restrict__ fcen,
                                                 const double* __restrict__ tau, const double* __restrict__ q2,
                                                 const long* __restrict__ t, const int ntile,
                                                 double2* __restrict__ zsum, double* __restrict__ gsum)
{
    const int b = blockIdx.x, p = b / ntile;
    const long a = kk[b], e = k1[b], g0 = kg[b];
    const float fc = (float)fcen[b], ta = (float)tau[b], qq = (float)(q2[b] / fcen[b]);
    const float amp = 0.886226925452758f * ta;               // sqrt(pi) / 2 * tau
    const float pit = 3.14159265358979f * ta;
    const long tp = t[p];
    const float inv = 2.0f / (float)N;
    const long dstep = pow2 ? (((long)NT * tp) & (N - 1)) : (((long)NT * tp) % N);
    float ws, wc;                                           // the phasor's rotation from one of its bins to the next
    sincospif((float)dstep * inv, &ws, &wc);
    double sre = 0.0, sim = 0.0, s2 = 0.0;
    for (long k0 = max(a, g0) + threadIdx.x; k0 < e; k0 += (long)NT * R) {
        const long prod = k0 * tp;
        const long ph = pow2 ? (prod & (N - 1)) : (prod % N);
        float sn, cs;
        sincospif((float)ph * inv, &sn, &cs);
        const long k1e = min(e, k0 + (long)NT * R);
        float fr = 0.f, fi = 0.f, f2 = 0.f;
        for (long k = k0; k < k1e; k += NT) {
            const float f = (float)((double)k * df);
            const float x = pit * (f - fc);
            const float g = amp * expf(-x * x) * (1.0f + expf(-qq * f));
            const float2 d = D[k];
            const float wr = d.x * cs - d.y * sn, wi = d.x * sn + d.y * cs;     // D_k e^{i phase}
            fr += g * wr;
            fi += g * wi;
            f2 += g * g * IP[k];
            const float nc = cs * wc - sn * ws;
            sn = cs * ws + sn * wc;
            cs = nc;
        }
        sre += (double)fr;
        sim += (double)fi;
        s2 += (double)f2;
    }
    sre = wave_sum_f64(sre);
    sim = wave_sum_f64(sim);
    s2 = wave_sum_f64(s2);
    __shared__ double red[3][NT / 64];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[0][w] = sre; red[1][w] = sim; red[2][w] = s2; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0, i = 0.0, s = 0.0;
        for (int j = 0; j < NT / 64; j++) { r += red[0][j]; i += red[1][j]; s += red[2][j]; }
        zsum[b] = make_double2(r, i);
        gsum[b] = s;
    }
}

}  // namespace

GWGS_API int gwgs_sg_tiles(const void* D, const float* IP, long flen, long N, double df, const long* kk,
                           const long* k1, const long* kg, const double* fcen, const double* tau, const double* q2,
                           const long* t, int npk, int ntile, void* zsum, double* gsum, void* stream)
{
    GWGS_CHECK(npk >= 0 && ntile >= 1 && N >= 2 && flen >= 1 && flen <= N / 2 + 1 && df > 0, GWGS_EINVAL);
    if (npk == 0) return GWGS_OK;
    GWGS_CHECK(D && IP && kk && k1 && kg && fcen && tau && q2 && t && zsum && gsum, GWGS_EINVAL);
    GWGS_CHECK((long)npk * ntile < (1L << 31), GWGS_ESIZE);
    const int pow2 = (N & (N - 1)) == 0;
    hipLaunchKernelGGL(sg_tiles_k, dim3(npk * ntile), dim3(NT), 0, (hipStream_t)stream, (const float2*)D, IP, N, pow2,
                       df, kk, k1, kg, fcen, tau, q2, t, ntile, (double2*)zsum, gsum);
    return gwgs_status();
}
