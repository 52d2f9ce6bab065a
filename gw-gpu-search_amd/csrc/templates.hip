// Frequency-domain chirp templates and their sigma^2 against each segment's inverse PSD.
//
// Templates: stationary-phase chirps h(f) = f^(-7/6) exp(-i psi(f)) on k df for k in [kmin, kmax), kmax =
// min(int(f_isco / df), flen - 1), with psi the TaylorF2 phase to 1PN (t_c = phi_c = 0) plus the leading eccentric
// term of Yunes, Arun, Berti & Will 2009 (PRD 80 084001):
//     psi = 3 / (128 eta v^5) [1 + (3715/756 + 55 eta / 9) v^2 - (2355/1462) e0^2 (f / f0)^(-19/9)] - pi / 4,
//     v = (pi M f)^(1/3).
// The phase reaches 1e5 rad at the low-frequency end of light binaries, so it is evaluated in double and only the
// final complex value is rounded to float32.
//
// sigma^2 (fp_filter.py:710-714): sig[i, s] = 4 df sum_k |h_ik|^2 IP_sk, accumulated in double.
#include "common.h"

namespace {

constexpr int NT = 256;
constexpr double MSUN_S = 4.925490947641267e-06;
constexpr int MAXSEG = 16;

__global__ void spa_templates_k(const double* __restrict__ m1, const double* __restrict__ m2,
                                const double* __restrict__ e0, const int n, const long flen, const double df,
                                const double flow, const double f0, float2* __restrict__ H, const long ldh,
                                int* __restrict__ kmax_out)
{
    const int i = blockIdx.y;
    const double M = (m1[i] + m2[i]) * MSUN_S;
    const double eta = m1[i] * m2[i] / ((m1[i] + m2[i]) * (m1[i] + m2[i]));
    const double fisco = 1.0 / (pow(6.0, 1.5) * M_PI * M);
    const long kmin = (long)(flow / df);
    const long kmax = min((long)(fisco / df), flen - 1);
    if (blockIdx.x == 0 && threadIdx.x == 0 && kmax_out) kmax_out[i] = (int)kmax;
    const double ecc = e0 ? e0[i] : 0.0;
    const double c1 = 3715.0 / 756.0 + 55.0 * eta / 9.0;
    const double pref = 3.0 / (128.0 * eta);
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < flen; k += (long)gridDim.x * blockDim.x) {
        float2 out = make_float2(0.f, 0.f);
        if (k >= kmin && k < kmax) {
            const double f = (double)k * df;
            const double v = cbrt(M_PI * M * f);
            const double v2 = v * v;
            double pn = 1.0 + c1 * v2;
            if (ecc != 0.0) pn -= (2355.0 / 1462.0) * ecc * ecc * pow(f / f0, -19.0 / 9.0);
            const double psi = pref / (v2 * v2 * v) * pn - 0.25 * M_PI;
            const double amp = pow(f, -7.0 / 6.0);
            double s, c;
            sincos(psi, &s, &c);
            out = make_float2((float)(amp * c), (float)(-amp * s));
        }
        H[(size_t)i * ldh + k] = out;
    }
}

// Partial sums over chunks of CHK bins: block (row i, chunk c) -> part[(c n + i) nseg + s] = sum over its bins of
// |h_ik|^2 IP_sk in double. Rows vary fastest in the grid so the blocks in flight share the IP chunk in L2.
constexpr int CHK = 8192;

// NS >= nseg segments are summed; row i and each segment's IP over [k0, k1) are read as buffers, so bins past k1 and
// segments s >= nseg read zeros (adding exact zeros): the CHK / NT steps are the same for every lane and carry no
// branch, and their loads issue back to back.
template <int NS>
__global__ void __launch_bounds__(NT) sigmasq_part_k(const float2* __restrict__ H, const long ldh, const long kmin,
                                                     const long kend, const float* __restrict__ IP, const long ldip,
                                                     const int nseg, double* __restrict__ part)
{
    const int i = blockIdx.x, c = blockIdx.y, n = gridDim.x;
    const long k0 = kmin + (long)c * CHK, k1 = min(kend, k0 + CHK);
    double acc[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) acc[s] = 0.0;
    const __amdgpu_buffer_rsrc_t rh = buffer_of(H + (size_t)i * ldh + k0, (k1 - k0) * 8);
    __amdgpu_buffer_rsrc_t ri[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) ri[s] = buffer_of(IP + (size_t)s * ldip + k0, s < nseg ? (k1 - k0) * 4 : 0);
#pragma unroll 2
    for (int j = 0; j < CHK / NT; j++) {
        const unsigned o = (unsigned)(threadIdx.x + NT * j);
        const float2 v = load_f2(rh, o * 8u);
        float ip[NS];
#pragma unroll
        for (int s = 0; s < NS; s++) ip[s] = load_f1(ri[s], o * 4u);
        const double p = (double)v.x * v.x + (double)v.y * v.y;
#pragma unroll
        for (int s = 0; s < NS; s++) acc[s] += p * (double)ip[s];
    }
    __shared__ double sh[NT / 64][NS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const double x = wave_sum_f64(acc[s]);
        if (lane == 0) sh[wave][s] = x;
    }
    __syncthreads();
    if (threadIdx.x < nseg) {
        const int s = threadIdx.x;
        part[((size_t)c * n + i) * nseg + s] = ((sh[0][s] + sh[1][s]) + sh[2][s]) + sh[3][s];
    }
}

// sig[i, s] = 4 df sum over chunks in order (deterministic)
__global__ void sigmasq_sum_k(const double* __restrict__ part, const int n, const int nch, const int nseg,
                              const double df, double* __restrict__ sig, const long lds)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)n * nseg) return;
    const int i = (int)(g / nseg), s = (int)(g % nseg);
    double t = 0.0;
    for (int c = 0; c < nch; c++) t += part[((size_t)c * n + i) * nseg + s];
    sig[(size_t)i * lds + s] = 4.0 * df * t;
}

// Coarse-node form of the same chirps (the interpolated templates of the reference's --templates interp,
// imr_interp.py:53-85): at nodes x_0 < ... < x_J the amplitude f^(-7/6) (0 at or above f_isco) and, per interval
// j, the phase -psi as a cubic in d = f - x_j, c3 d^3 + c2 d^2 + c1 d + c0, the Hermite cubic through -psi and its
// analytic derivative at both ends (psi as in spa_templates_k). One thread per (template, node).
__device__ __forceinline__ void spa_phase(const double f, const double M, const double pref, const double c1,
                                          const double ke, const double f0, double& ph, double& dph)
{
    const double pmf = M_PI * M * f;
    const double v5 = pow(pmf, -5.0 / 3.0), v2i = 1.0 / pmf;
    const double fe = ke != 0.0 ? pow(f / f0, -19.0 / 9.0) : 0.0;
    const double psi = pref * (v5 + c1 * v2i - ke * v5 * fe) - 0.25 * M_PI;
    const double dpsi = pref * (-5.0 / 3.0 * v5 - c1 * v2i + (5.0 / 3.0 + 19.0 / 9.0) * ke * v5 * fe) / f;
    ph = -psi;
    dph = -dpsi;
}

__global__ void spa_nodes_k(const double* __restrict__ m1, const double* __restrict__ m2,
                            const double* __restrict__ e0, const int n, const double* __restrict__ x, const int nx,
                            const double f0, double* __restrict__ amp, double* __restrict__ coef)
{
    const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || j >= nx) return;
    const double M = (m1[i] + m2[i]) * MSUN_S;
    const double eta = m1[i] * m2[i] / ((m1[i] + m2[i]) * (m1[i] + m2[i]));
    const double fisco = 1.0 / (pow(6.0, 1.5) * M_PI * M);
    const double ecc = e0 ? e0[i] : 0.0;
    const double c1 = 3715.0 / 756.0 + 55.0 * eta / 9.0, pref = 3.0 / (128.0 * eta);
    const double ke = (2355.0 / 1462.0) * ecc * ecc;
    const double fa = x[j];
    amp[(size_t)i * nx + j] = fa < fisco ? pow(fa, -7.0 / 6.0) : 0.0;
    if (j == nx - 1) return;
    const double fb = x[j + 1], h = fb - fa;
    double pa, da, pb, db;
    spa_phase(fa, M, pref, c1, ke, f0, pa, da);
    spa_phase(fb, M, pref, c1, ke, f0, pb, db);
    const double sl = (pb - pa) / h;
    double* c = coef + ((size_t)i * (nx - 1) + j) * 4;
    c[0] = (da + db - 2.0 * sl) / (h * h);
    c[1] = (3.0 * sl - 2.0 * da - db) / h;
    c[2] = da;
    c[3] = pa;
}

// Dense rows of the interpolated templates on [kmin, kend): for local bin q, node[q] = its interval j, d[q] = f - x_j,
// t[q] = d / (x_j+1 - x_j); h = (A_j (1 - t) + A_j+1 t) exp(i p(d)), evaluated in double, rounded to complex64
// (imr_interp.py:63-72). Zeros elsewhere.
__global__ void interp_eval_k(const double* __restrict__ amp, const double* __restrict__ coef, const int nx,
                              const int* __restrict__ node, const double* __restrict__ d, const double* __restrict__ t,
                              const long kmin, const long kend, const long flen, float2* __restrict__ H,
                              const long ldh)
{
    const int i = blockIdx.y;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < flen; k += (long)gridDim.x * blockDim.x) {
        float2 out = make_float2(0.f, 0.f);
        if (k >= kmin && k < kend) {
            const long q = k - kmin;
            const int j = node[q];
            const double* c = coef + ((size_t)i * (nx - 1) + j) * 4;
            const double dd = d[q], tt = t[q];
            const double p = ((c[0] * dd + c[1]) * dd + c[2]) * dd + c[3];
            const double a = amp[(size_t)i * nx + j] * (1.0 - tt) + amp[(size_t)i * nx + j + 1] * tt;
            double s, co;
            sincos(p, &s, &co);
            out = make_float2((float)(a * co), (float)(a * s));
        }
        H[(size_t)i * ldh + k] = out;
    }
}

} // namespace

GWGS_API int gwgs_spa_templates(const double* m1, const double* m2, const double* e0, int n, long flen, double df,
                                double flow, double f0, void* H, long ldh, int* kmax, void* stream)
{
    GWGS_CHECK(m1 && m2 && H && n >= 0 && flen > 0 && df > 0 && ldh >= flen, GWGS_EINVAL);
    if (n == 0) return GWGS_OK;
    GWGS_CHECK(n < 65536, GWGS_ESIZE);
    const unsigned bx = (unsigned)min((flen + NT - 1) / NT, 512L);
    hipLaunchKernelGGL(spa_templates_k, dim3(bx, n), dim3(NT), 0, (hipStream_t)stream, m1, m2, e0, n, flen, df, flow,
                       f0, (float2*)H, ldh, kmax);
    return gwgs_status();
}

// Workspace doubles for gwgs_sigmasq over [kmin, kend) with n rows and nseg segments.
GWGS_API long gwgs_sigmasq_workspace(int n, long kmin, long kend, int nseg)
{
    if (n <= 0 || kend <= kmin || nseg <= 0) return 0;
    return (long)n * ((kend - kmin + CHK - 1) / CHK) * nseg;
}

// sig[i, s] = 4 df sum_{kmin <= k < kend} |H[i, k]|^2 IP[s, k] (double); H is zero outside each row's band, so
// [kmin, kend) covering every row's band gives the full-row sum. work: gwgs_sigmasq_workspace doubles.
GWGS_API int gwgs_sigmasq(const void* H, long ldh, long kmin, long kend, int n, const float* IP, long ldip, int nseg,
                          double df, double* work, double* sig, long lds, void* stream)
{
    GWGS_CHECK(H && IP && sig && n >= 0 && nseg > 0 && lds >= nseg && kmin >= 0 && kend >= kmin && kend <= ldh &&
               kend <= ldip, GWGS_EINVAL);
    GWGS_CHECK(nseg <= MAXSEG, GWGS_ESIZE);
    if (n == 0) return GWGS_OK;
    hipStream_t st = (hipStream_t)stream;
    const int nch = (int)((kend - kmin + CHK - 1) / CHK);
    if (nch > 0) {
        GWGS_CHECK(work, GWGS_EINVAL);
        GWGS_CHECK(nch < 65536 && n < (1 << 30), GWGS_ESIZE);
        const dim3 grid(n, nch);
        const float2* h = (const float2*)H;
        if (nseg == 1) hipLaunchKernelGGL(sigmasq_part_k<1>, grid, dim3(NT), 0, st, h, ldh, kmin, kend, IP, ldip, nseg, work);
        else if (nseg == 2) hipLaunchKernelGGL(sigmasq_part_k<2>, grid, dim3(NT), 0, st, h, ldh, kmin, kend, IP, ldip, nseg, work);
        else if (nseg <= 4) hipLaunchKernelGGL(sigmasq_part_k<4>, grid, dim3(NT), 0, st, h, ldh, kmin, kend, IP, ldip, nseg, work);
        else if (nseg <= 8) hipLaunchKernelGGL(sigmasq_part_k<8>, grid, dim3(NT), 0, st, h, ldh, kmin, kend, IP, ldip, nseg, work);
        else hipLaunchKernelGGL(sigmasq_part_k<MAXSEG>, grid, dim3(NT), 0, st, h, ldh, kmin, kend, IP, ldip, nseg, work);
    }
    const long tot = (long)n * nseg;
    hipLaunchKernelGGL(sigmasq_sum_k, dim3((unsigned)((tot + NT - 1) / NT)), dim3(NT), 0, st, work, n, nch, nseg, df,
                       sig, lds);
    return gwgs_status();
}

// Coarse nodes of n chirps at the nx nodes x (ascending, device float64): amp (n, nx) and the phase cubics coef
// (n, nx - 1, 4) as (c3, c2, c1, c0) in float64 (see spa_nodes_k).
GWGS_API int gwgs_spa_nodes(const double* m1, const double* m2, const double* e0, int n, const double* x, int nx,
                            double f0, double* amp, double* coef, void* stream)
{
    GWGS_CHECK(m1 && m2 && x && amp && coef && n >= 0 && nx >= 2, GWGS_EINVAL);
    if (n == 0) return GWGS_OK;
    GWGS_CHECK(n < 65536, GWGS_ESIZE);
    hipLaunchKernelGGL(spa_nodes_k, dim3((nx + NT - 1) / NT, n), dim3(NT), 0, (hipStream_t)stream, m1, m2, e0, n, x,
                       nx, f0, amp, coef);
    return gwgs_status();
}

// Dense interpolated templates (n, flen) complex64 from nodes (amp (n, nx), coef (n, nx - 1, 4) float64) and the
// fine-bin grid of [kmin, kend) (node int32, d and t float64, kend - kmin entries each).
GWGS_API int gwgs_interp_templates(const double* amp, const double* coef, int n, int nx, const int* node,
                                   const double* d, const double* t, long kmin, long kend, long flen, void* H,
                                   long ldh, void* stream)
{
    GWGS_CHECK(amp && coef && node && d && t && H && n >= 0 && nx >= 2 && kmin >= 0 && kend >= kmin &&
               kend <= flen && ldh >= flen, GWGS_EINVAL);
    if (n == 0) return GWGS_OK;
    GWGS_CHECK(n < 65536, GWGS_ESIZE);
    const unsigned bx = (unsigned)min((flen + NT - 1) / NT, 512L);
    hipLaunchKernelGGL(interp_eval_k, dim3(bx, n), dim3(NT), 0, (hipStream_t)stream, amp, coef, nx, node, d, t, kmin,
                       kend, flen, (float2*)H, ldh);
    return gwgs_status();
}
