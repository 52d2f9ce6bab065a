// Differentiable peak SNR of a compressed template (no reference counterpart: the backward kernel; the forward
// quantity is the reference's rebuilt SNR, fp_filter.py:759-764, evaluated off the sample grid and off the bank).
//
// For a template given by its coefficient vector a (r complex, z = sum_j a_j Zc_j, the rows of A in the basis of one
// tile) at the continuous time tau = t + delta (t an integer sample, delta a fraction of a sample), with the basis
// series evaluated exactly from the frequency domain, Zc_j(tau) = sum_{k in [k0, k1)} Vc[j, k] D[k] e^{2 pi i k tau / n}:
//   z = sum_j a_j R_j,  R_j = Zc_j(tau),  dz/ddelta = sum_j a_j R'_j,  R'_j = sum_k (2 pi i k / n) Vc[j, k] D[k] e^{...}
//   sigma^2 = 4 df a^T G conj(a),  G[j, j'] = sum_k Vc[j, k] conj(Vc[j', k]) IP[k]   (the template's norm)
//   rho^2 = |z|^2 (4 df)^2 / sigma^2  and, for a phase phi, rho_phi = Re(e^{-i phi} z) 4 df / sigma.
// phasor_data_k writes E[b, k] = D[k] e^{2 pi i k tau_b / n} and E'[b, k] = (2 pi i k / n) E[b, k] (the phase from
// the exact 64-bit integer k t mod n plus k delta in double); R = E Vc^T is a plain GEMM. snr_grad_k (one wave per
// peak, double accumulation) returns rho^2, z, dz/ddelta, and the gradients
//   d rho^2 / d a_j = (16 df^2 / sigma^2) (z conj(R_j) - (|z|^2 / sigma^2) 4 df (G^T a)_j)   as d/dRe + i d/dIm,
//   d rho^2 / d delta = (32 df^2 / sigma^2) Re(conj(z) dz/ddelta),
//   d rho_phi^2 / d phi = 2 (4 df)^2 / sigma^2 Re(e^{-i phi} z) Im(e^{-i phi} z).
#include "common.h"

namespace {

constexpr int NT = 256;

__global__ void __launch_bounds__(NT) phasor_data_k(const float2* __restrict__ D, const long k0, const long kb,
                                                    const long long* __restrict__ t, const double* __restrict__ dl,
                                                    const unsigned long long n, float2* __restrict__ E,
                                                    float2* __restrict__ Ed, const long lde)
{
    const int b = blockIdx.y;
    const long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i >= kb) return;
    const unsigned long long k = (unsigned long long)(k0 + i);
    const unsigned long long tt = (unsigned long long)(((t[b] % (long long)n) + (long long)n) % (long long)n);
    const double ph = (double)((k * tt) % n) / (double)n + (double)k * dl[b] / (double)n;   // turns
    double s, c;
    sincospi(2.0 * (ph - rint(ph)), &s, &c);
    const float2 d = D[k];
    const double er = (double)d.x * c - (double)d.y * s, ei = (double)d.x * s + (double)d.y * c;
    E[(size_t)b * lde + i] = make_float2((float)er, (float)ei);
    const double w = 2.0 * M_PI * (double)k / (double)n;                                   // d/ddelta: times i w
    Ed[(size_t)b * lde + i] = make_float2((float)(-w * ei), (float)(w * er));
}

// one wave per peak; lanes stride over j
__global__ void __launch_bounds__(NT) snr_grad_k(const float2* __restrict__ R, const float2* __restrict__ Rd,
                                                 const long ldr, const float2* __restrict__ a, const long lda,
                                                 const float2* __restrict__ G, const long ldg, const int r,
                                                 const int npk, const double df, const float* __restrict__ phi,
                                                 float* __restrict__ rho2, float2* __restrict__ zout,
                                                 float* __restrict__ drho2_dt, float* __restrict__ drho2_dphi,
                                                 float2* __restrict__ grad, const long ldgr)
{
    const int lane = threadIdx.x & 63;
    const int b = blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
    if (b >= npk) return;
    const float2* Rb = R + (size_t)b * ldr;
    const float2* Rdb = Rd + (size_t)b * ldr;
    const float2* ab = a + (size_t)b * lda;
    double zr = 0, zi = 0, dr = 0, di = 0, s = 0;
    for (int j = lane; j < r; j += 64) {
        const double ar = ab[j].x, ai = ab[j].y;
        zr += ar * Rb[j].x - ai * Rb[j].y;
        zi += ar * Rb[j].y + ai * Rb[j].x;
        dr += ar * Rdb[j].x - ai * Rdb[j].y;
        di += ar * Rdb[j].y + ai * Rdb[j].x;
        // (G^T a)_j = sum_j' G[j', j] a_j'; s = sum_j conj(a_j) (G^T a)_j = a^T G conj(a)
        double gr = 0, gi = 0;
        for (int jj = 0; jj < r; jj++) {
            const float2 g = G[(size_t)jj * ldg + j];
            const double xr = ab[jj].x, xi = ab[jj].y;
            gr += (double)g.x * xr - (double)g.y * xi;
            gi += (double)g.x * xi + (double)g.y * xr;
        }
        s += ar * gr + ai * gi;
        if (grad) grad[(size_t)b * ldgr + j] = make_float2((float)gr, (float)gi);   // G^T a, finished below
    }
    zr = wave_sum_f64(zr); zi = wave_sum_f64(zi);
    dr = wave_sum_f64(dr); di = wave_sum_f64(di);
    s = wave_sum_f64(s);
    const double c = 4.0 * df;
    const double sig2 = c * s;
    const double z2 = zr * zr + zi * zi;
    const double f = c * c / sig2;                     // (4 df)^2 / sigma^2
    if (lane == 0) {
        rho2[b] = (float)(z2 * f);
        zout[b] = make_float2((float)zr, (float)zi);
        drho2_dt[b] = (float)(2.0 * f * (zr * dr + zi * di));
        const double p = phi ? (double)phi[b] : atan2(zi, zr);
        const double cp = cos(p), sp = sin(p);
        const double re = cp * zr + sp * zi, im = cp * zi - sp * zr;   // e^{-i phi} z
        drho2_dphi[b] = (float)(2.0 * f * re * im);
    }
    if (grad) {
        // d rho^2 = f d|z|^2 - rho^2 dsigma^2 / sigma^2; complex gradients 2 z conj(R_j) and 2 c (G^T a)_j
        const double rho2v = z2 * f;
        for (int j = lane; j < r; j += 64) {
            const float2 ga = grad[(size_t)b * ldgr + j];
            const double Rr = Rb[j].x, Ri = Rb[j].y;
            const double gzr = 2.0 * (zr * Rr + zi * Ri), gzi = 2.0 * (zi * Rr - zr * Ri);   // 2 z conj(R_j)
            const double k2 = rho2v / sig2 * 2.0 * c;
            grad[(size_t)b * ldgr + j] = make_float2((float)(f * gzr - k2 * ga.x), (float)(f * gzi - k2 * ga.y));
        }
    }
}

} // namespace

// E[b, i] = D[k0 + i] e^{2 pi i (k0 + i) (t[b] + delta[b]) / n}, Ed[b, i] = (2 pi i (k0 + i) / n) E[b, i], i < kb;
// t int64 samples, delta double (fraction of a sample), on the device. E, Ed (npk, lde >= kb) complex64.
GWGS_API int gwgs_phasor_data(const void* D, long k0, long kb, const long long* t, const double* delta, long long n,
                              int npk, void* E, void* Ed, long lde, void* stream)
{
    GWGS_CHECK(D && t && delta && E && Ed && k0 >= 0 && kb >= 0 && n > 0 && npk >= 0 && lde >= kb, GWGS_EINVAL);
    GWGS_CHECK(npk < 65536, GWGS_ESIZE);
    if (npk == 0 || kb == 0) return GWGS_OK;
    hipLaunchKernelGGL(phasor_data_k, dim3((unsigned)((kb + NT - 1) / NT), npk), dim3(NT), 0, (hipStream_t)stream,
                       (const float2*)D, k0, kb, t, delta, (unsigned long long)n, (float2*)E, (float2*)Ed, lde);
    return gwgs_status();
}

// Per peak b < npk (one wave each): rho^2, z, d rho^2 / d delta, d rho_phi^2 / d phi (phi: per peak, or NULL for
// phi = arg z, where it vanishes), and (grad != NULL) d rho^2 / d a as complex gradients (d/dRe + i d/dIm), from
// R, Rd (npk, ldr) = the basis responses and their delta derivatives, a (npk, lda) the coefficient vectors, G (r, ldg)
// the template-norm metric (sigma^2 = 4 df a^T G conj(a)). Complex arrays complex64; accumulation in double.
GWGS_API int gwgs_snr_grad(const void* R, const void* Rd, long ldr, const void* a, long lda, const void* G, long ldg,
                           int r, int npk, double df, const float* phi, float* rho2, void* z, float* drho2_dt,
                           float* drho2_dphi, void* grad, long ldgr, void* stream)
{
    GWGS_CHECK(R && Rd && a && G && rho2 && z && drho2_dt && drho2_dphi && r > 0 && npk >= 0 && ldr >= r &&
               lda >= r && ldg >= r && df > 0 && (!grad || ldgr >= r), GWGS_EINVAL);
    if (npk == 0) return GWGS_OK;
    hipLaunchKernelGGL(snr_grad_k, dim3((unsigned)((npk + NT / 64 - 1) / (NT / 64))), dim3(NT), 0,
                       (hipStream_t)stream, (const float2*)R, (const float2*)Rd, ldr, (const float2*)a, lda,
                       (const float2*)G, ldg, r, npk, df, phi, rho2, (float2*)z, drho2_dt, drho2_dphi, (float2*)grad,
                       ldgr);
    return gwgs_status();
}
