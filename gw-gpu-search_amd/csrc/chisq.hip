// Exact complex SNR numerator and power chi-squared bin sums at peaks (the reference's chisq_kernel.py:14-121), and
// equal-power chi-squared bin edges (chisq_edges.py:90-143), redesigned for wave64.
#include "common.h"

namespace {

constexpr int NT = 256;
constexpr int P = 8;       // peaks per block (one template row's peaks share its read of q / h)
constexpr int R = 32;      // consecutive bins per thread run: the phasor is reseeded exactly at each run start

// exp(2 pi i ((k t) mod N) / N), exact integer phase, double-precision sincos
__device__ __forceinline__ float2 phasor(const unsigned long long k, const unsigned long long t,
                                         const unsigned long long N)
{
    const unsigned long long m = (k * t) % N;
    double s, c;
    sincospi(2.0 * (double)m / (double)N, &s, &c);
    return make_float2((float)c, (float)s);
}

// One block per group of <= P peaks of one row. For each bin j = [e_j, e_j+1) the threads take runs of R consecutive
// bins (run r of the bin -> thread r mod NT), reseed each peak's phasor at the run start and advance it by a complex
// rotation per bin. Per bin: wave butterfly sums, then the 4 waves in fixed order, so the result is deterministic.
// z = sum_j Z_j and pw = sum_j |Z_j|^2 accumulate in double. FUSE: the summand is conj(h_k) d_k formed here.
template <bool FUSE>
__global__ void __launch_bounds__(NT) chisq_pts_k(const float2* __restrict__ q, const long ldq,
                                                  const float2* __restrict__ dv, const unsigned long long N,
                                                  const int* __restrict__ grow, const int* __restrict__ gstart,
                                                  const int* __restrict__ gcount, const long long* __restrict__ ptime,
                                                  const int* __restrict__ edges, const int nedge,
                                                  float* __restrict__ out, float2* __restrict__ zout)
{
    const int g = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int row = grow[g], p0 = gstart[g], np_ = gcount[g];
    const float2* qr = q + (size_t)row * ldq;
    const int* e = edges + (size_t)g * nedge;
    unsigned long long t[P];
    float2 rot[P];
#pragma unroll
    for (int p = 0; p < P; p++) {
        t[p] = p < np_ ? (unsigned long long)ptime[p0 + p] % N : 0ull;
        rot[p] = phasor(1ull, t[p], N);
    }
    __shared__ float2 red[NT / 64][P];
    double tot = 0.0, zx = 0.0, zy = 0.0;    // meaningful in thread p < np_
    for (int j = 0; j < nedge - 1; j++) {
        const int k0 = e[j], k1 = e[j + 1];
        if (k0 >= k1) continue;              // uniform over the block
        float2 acc[P];
#pragma unroll
        for (int p = 0; p < P; p++) acc[p] = make_float2(0.f, 0.f);
        const int nrun = (k1 - k0 + R - 1) / R;
        for (int r = tid; r < nrun; r += NT) {
            const int ka = k0 + r * R, kb = min(ka + R, k1);
            float2 w[P];
#pragma unroll
            for (int p = 0; p < P; p++) w[p] = p < np_ ? phasor((unsigned long long)ka, t[p], N) : make_float2(0.f, 0.f);
            for (int k = ka; k < kb; k++) {
                float2 v = qr[k];
                if (FUSE) {
                    const float2 dk = dv[k];
                    v = make_float2(v.x * dk.x + v.y * dk.y, v.x * dk.y - v.y * dk.x);
                }
#pragma unroll
                for (int p = 0; p < P; p++) {
                    acc[p].x = fmaf(v.x, w[p].x, fmaf(-v.y, w[p].y, acc[p].x));
                    acc[p].y = fmaf(v.x, w[p].y, fmaf(v.y, w[p].x, acc[p].y));
                    const float wx = fmaf(w[p].x, rot[p].x, -w[p].y * rot[p].y);
                    w[p].y = fmaf(w[p].x, rot[p].y, w[p].y * rot[p].x);
                    w[p].x = wx;
                }
            }
        }
#pragma unroll
        for (int p = 0; p < P; p++) {
            const float ax = wave_sum_f32(acc[p].x), ay = wave_sum_f32(acc[p].y);
            if (lane == 0) red[wave][p] = make_float2(ax, ay);
        }
        __syncthreads();
        if (tid < np_) {
            float ax = 0.f, ay = 0.f;
            for (int w = 0; w < NT / 64; w++) { ax += red[w][tid].x; ay += red[w][tid].y; }
            tot += (double)ax * ax + (double)ay * ay;
            zx += ax; zy += ay;
        }
        __syncthreads();
    }
    if (tid < np_) {
        out[p0 + tid] = (float)tot;
        if (zout) zout[p0 + tid] = make_float2((float)zx, (float)zy);
    }
}

// Exclusive scan of one double per thread over a 256-thread block, in a fixed order (deterministic).
__device__ __forceinline__ double block_exscan_f64(const double x, double* sh, double& total)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double inc = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    if (lane == 63) sh[wave] = inc;
    __syncthreads();
    double base = 0.0;
    for (int w = 0; w < wave; w++) base += sh[w];
    total = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return base + inc - x;
}

// One block per requested row: float64 sums of |h|^2 IP over contiguous thread chunks of [0, k1), a block scan, then
// each thread walks its own chunk for the edges whose target j/nb * total falls in it (first k with cumsum > target,
// multi_filter.py:264-282; chisq_edges.py:211-245). Columns j >= nb and the last column hold k1; edges are clipped
// to [kmin, k1].
__global__ void __launch_bounds__(NT) edges_fused_k(const float2* __restrict__ h, const long ldh,
                                                    const int* __restrict__ rows, const float* __restrict__ ips,
                                                    const int* __restrict__ k1s, const int* __restrict__ nbs,
                                                    const int nedge, const int kmin, int* __restrict__ out)
{
    const int g = blockIdx.x, t = threadIdx.x;
    const float2* hr = h + (size_t)rows[g] * ldh;
    const int k1 = k1s[g], nb = nbs[g];
    const int chunk = (k1 + NT - 1) / NT, a = min(t * chunk, k1), b = min(a + chunk, k1);
    double s = 0.0;
    for (int k = a; k < b; k++) {
        const float2 v = hr[k];
        s += ((double)v.x * v.x + (double)v.y * v.y) * (double)ips[k];
    }
    __shared__ double sh[NT / 64];
    double tot;
    const double lo = block_exscan_f64(s, sh, tot);
    const double hi = lo + s;
    int* eo = out + (size_t)g * nedge;
    for (int i = t; i < nedge; i += NT) eo[i] = k1;
    __syncthreads();
    if (b <= a) return;
    int j = 0;
    while (j < nb && ((double)j / (double)nb) * tot < lo) j++;
    if (j >= nb || ((double)j / (double)nb) * tot >= hi) return;
    double cs = lo;
    for (int k = a; k < b && j < nb; k++) {
        const float2 v = hr[k];
        cs += ((double)v.x * v.x + (double)v.y * v.y) * (double)ips[k];
        while (j < nb && cs > ((double)j / (double)nb) * tot) { eo[j] = min(max(k, kmin), k1); j++; }
    }
    while (j < nb && ((double)j / (double)nb) * tot < hi) { eo[j] = min(max(b, kmin), k1); j++; }
}

} // namespace

// Groups: for g in [0, ngroups): row grow[g] of q (or h in fused mode), peaks gstart[g] .. gstart[g] + gcount[g]
// (gcount <= 8) of ptime, bin edges edges[g, 0:nedge]. N: the time length of the phase factor (any positive value;
// the phase k t mod N is exact in 64-bit integers). d: data (fused mode) or null. zout may be null.
GWGS_API int gwgs_chisq_points(const void* q, long ldq, const void* d, long long N, const int* grow, const int* gstart,
                               const int* gcount, const long long* ptime, const int* edges, int nedge, int ngroups,
                               float* out, void* zout, void* stream)
{
    GWGS_CHECK(q && grow && gstart && gcount && ptime && edges && out && N > 0 && nedge >= 1 && ngroups >= 0,
               GWGS_EINVAL);
    if (ngroups == 0) return GWGS_OK;
    hipStream_t s = (hipStream_t)stream;
    if (d)
        hipLaunchKernelGGL(chisq_pts_k<true>, dim3(ngroups), dim3(NT), 0, s, (const float2*)q, ldq, (const float2*)d,
                           (unsigned long long)N, grow, gstart, gcount, ptime, edges, nedge, out, (float2*)zout);
    else
        hipLaunchKernelGGL(chisq_pts_k<false>, dim3(ngroups), dim3(NT), 0, s, (const float2*)q, ldq, nullptr,
                           (unsigned long long)N, grow, gstart, gcount, ptime, edges, nedge, out, (float2*)zout);
    return gwgs_status();
}

GWGS_API int gwgs_chisq_peaks_per_group(void) { return P; }

// Edges for n requested rows: rows index h (ldh complex per row); ips: 1/psd (float32, >= max k1 entries);
// k1s, nbs per request; out (n, nedge) int32.
GWGS_API int gwgs_edges_fused(const void* h, long ldh, const int* rows, const float* ips, const int* k1s,
                              const int* nbs, int nedge, int kmin, int n, int* out, void* stream)
{
    GWGS_CHECK(h && rows && ips && k1s && nbs && out && nedge >= 2 && n >= 0, GWGS_EINVAL);
    if (n == 0) return GWGS_OK;
    hipLaunchKernelGGL(edges_fused_k, dim3(n), dim3(NT), 0, (hipStream_t)stream, (const float2*)h, ldh, rows, ips,
                       k1s, nbs, nedge, kmin, out);
    return gwgs_status();
}
