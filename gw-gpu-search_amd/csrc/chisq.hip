// Exact complex SNR numerator and power chi-squared bin sums at peaks (the reference's chisq_kernel.py:14-121), and
// equal-power chi-squared bin edges (chisq_edges.py:90-143), redesigned for wave64.
#include "common.h"

namespace {

constexpr int NT = 256;
constexpr int P = 7;       // peak samples per group (one template row's; a peak and its +-3 re-centring samples)
constexpr int R = 32;      // phasor steps per lane run: the phasor is reseeded exactly at each run start

// exp(2 pi i ((k t) mod N) / N) from the exact integer phase m = (k t) mod N; mask = N - 1 when N is a power of two
// up to 2^24 (then the modulo is a mask and 2 m / N is exact in float), else 0 (double division); float sincospi.
__device__ __forceinline__ float2 phasor(const unsigned long long k, const unsigned long long t,
                                         const unsigned long long N, const unsigned long long mask)
{
    const unsigned long long kt = k * t;
    float x;
    if (mask) {
        x = (float)(kt & mask) * (2.f / (float)N);          // both factors exact: N a power of two
    } else {
        x = (float)(2.0 * (double)(kt % N) / (double)N);
    }
    float s, c;
    sincospif(x, &s, &c);
    return make_float2(c, s);
}

// One wave per (group of <= P peaks of one row, chi-squared bin j), four per workgroup: lanes take the bin's bins
// [e_j, e_j+1) interleaved (lane l: ka + l, ka + l + 64, ... for runs of 64 R bins: coalesced loads), each peak's
// phasor reseeded from the exact integer phase at the lane's first bin of a run and advanced by exp(2 pi i 64 t / N)
// per step; a wave butterfly sum gives Z_j of every peak, written to zp[(g nbin + j) P + p]; chisq_finish_k sums
// over j. Wave i = g nbin + j: a group's bins run on neighbouring waves (they share the row in L2). FUSE: the summand
// is conj(h_k) d_k formed here.
template <bool FUSE, bool MULTI = false>
__global__ void __launch_bounds__(NT) chisq_bins_k(const float2* __restrict__ q, const long ldq,
                                                   const float2* __restrict__ dv, const unsigned long long N,
                                                   const int* __restrict__ grow, const int* __restrict__ gstart,
                                                   const int* __restrict__ gcount, const long long* __restrict__ ptime,
                                                   const int* __restrict__ edges, const int nedge, const int ngroups,
                                                   float2* __restrict__ zp, const float2* const* __restrict__ dtab,
                                                   const int* __restrict__ gseg)
{
    const int nbin = nedge - 1;
    const int lane = threadIdx.x & 63;
    // wave index, uniform over the wave by construction (stated, so that the row, its edges and the buffer
    // descriptors below live in scalar registers)
    const long wi = (long)blockIdx.x * (NT / 64) + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (wi >= (long)ngroups * nbin) return;   // whole waves
    const int g = (int)(wi / nbin), j = (int)(wi % nbin);
    const int row = grow[g], p0 = gstart[g], np_ = gcount[g];
    const int* e = edges + (size_t)g * nedge;
    const int k0 = e[j], k1 = e[j + 1];
    float2* zo = zp + (size_t)wi * P;
    if (k0 >= k1) {                           // empty bin: the finish pass skips it
        if (lane < P) zo[lane] = make_float2(0.f, 0.f);
        return;
    }
    const unsigned long long mask = (N & (N - 1)) == 0 && N <= (1ull << 24) ? N - 1 : 0;
    const float2* qr = q + (size_t)row * ldq;
    if (MULTI) dv = dtab[gseg[g]];            // the group's segment (its data row)
    unsigned long long t[P];
    float2 rot[P];
#pragma unroll
    for (int p = 0; p < P; p++) {
        t[p] = p < np_ ? (unsigned long long)ptime[p0 + p] % N : 0ull;
        rot[p] = phasor(64ull, t[p], N, mask);
    }
    float2 acc[P];
#pragma unroll
    for (int p = 0; p < P; p++) acc[p] = make_float2(0.f, 0.f);
    // q (and d) through buffers that end at k1: a lane past the bin's end reads zeros and adds exact zeros, so every
    // lane of a run takes the same number of steps, rounded up to a multiple of 4 (full runs have R = 32), and the
    // loads carry no branch. Four registers rotate by unrolling, so each step's loads are issued three steps ahead
    // (a register copy would wait for the load). Byte offsets 8 k < 2^31.
    const __amdgpu_buffer_rsrc_t rq = buffer_of(qr, (long)k1 * 8), rd = buffer_of(FUSE ? dv : qr, (long)k1 * 8);
    constexpr int RW = 64 * R;                // bins per run
    static_assert(R % 4 == 0, "runs are whole groups of 4 steps");
    auto step = [&](const float2 qv, const float2 dk, float2 (&w)[P]) {
        float2 v = qv;
        if (FUSE) v = make_float2(qv.x * dk.x + qv.y * dk.y, qv.x * dk.y - qv.y * dk.x);
#pragma unroll
        for (int p = 0; p < P; p++) {
            acc[p].x = fmaf(v.x, w[p].x, fmaf(-v.y, w[p].y, acc[p].x));
            acc[p].y = fmaf(v.x, w[p].y, fmaf(v.y, w[p].x, acc[p].y));
            const float wx = fmaf(w[p].x, rot[p].x, -w[p].y * rot[p].y);
            w[p].y = fmaf(w[p].x, rot[p].y, w[p].y * rot[p].x);
            w[p].x = wx;
        }
    };
    for (int ra = k0; ra < k1; ra += RW) {
        const int ka = ra + lane;
        const int ns = ((min(ra + RW, k1) - ra + 255) >> 8) << 2;   // steps of this run, the same for every lane
        float2 w[P];
#pragma unroll
        for (int p = 0; p < P; p++) w[p] = p < np_ ? phasor((unsigned long long)ka, t[p], N, mask) : make_float2(0.f, 0.f);
        const unsigned o0 = (unsigned)ka * 8u;
        float2 qa = load_f2(rq, o0), qb = load_f2(rq, o0 + 512u), qc = load_f2(rq, o0 + 1024u), qd = qa;
        float2 da = qa, db = qb, dc = qc, dd = qa;
        if (FUSE) { da = load_f2(rd, o0); db = load_f2(rd, o0 + 512u); dc = load_f2(rd, o0 + 1024u); }
        for (int s = 0; s < ns; s += 4) {
            const unsigned o = o0 + 512u * (unsigned)s;
            qd = load_f2(rq, o + 1536u);
            if (FUSE) dd = load_f2(rd, o + 1536u);
            step(qa, da, w);
            qa = load_f2(rq, o + 2048u);
            if (FUSE) da = load_f2(rd, o + 2048u);
            step(qb, db, w);
            qb = load_f2(rq, o + 2560u);
            if (FUSE) db = load_f2(rd, o + 2560u);
            step(qc, dc, w);
            qc = load_f2(rq, o + 3072u);
            if (FUSE) dc = load_f2(rd, o + 3072u);
            step(qd, dd, w);
        }
    }
#pragma unroll
    for (int p = 0; p < P; p++) {
        const float ax = wave_sum_f32(acc[p].x), ay = wave_sum_f32(acc[p].y);
        if (lane == p) zo[p] = make_float2(ax, ay);
    }
}

// One thread per (group, peak): z = sum_j Z_j and pw = sum_j |Z_j|^2 over the non-empty bins in increasing j, in
// double (the fixed order makes the result deterministic).
__global__ void __launch_bounds__(NT) chisq_finish_k(const float2* __restrict__ zp, const int* __restrict__ gstart,
                                                     const int* __restrict__ gcount, const int* __restrict__ edges,
                                                     const int nedge, const int ngroups, float* __restrict__ out,
                                                     float2* __restrict__ zout)
{
    const int i = blockIdx.x * NT + threadIdx.x;
    const int g = i / P, p = i % P;
    if (g >= ngroups || p >= gcount[g]) return;
    const int nbin = nedge - 1;
    const int* e = edges + (size_t)g * nedge;
    double tot = 0.0, zx = 0.0, zy = 0.0;
    for (int j = 0; j < nbin; j++) {
        if (e[j] >= e[j + 1]) continue;
        const float2 z = zp[((size_t)g * nbin + j) * P + p];
        tot += (double)z.x * z.x + (double)z.y * z.y;
        zx += z.x; zy += z.y;
    }
    out[gstart[g] + p] = (float)tot;
    if (zout) zout[gstart[g] + p] = make_float2((float)zx, (float)zy);
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

// One block per requested row (multi_filter.py:264-282; chisq_edges.py:211-245): edge j = first k with the float64
// cumulative sum of |h|^2 IP over [0, k1) above j / nb * total, clipped to [kmin, k1]; columns j >= nb and the last
// hold k1. Pass 1: chunk sums, each chunk read by one wave in coalesced 64-bin strides; a block scan of the chunk
// sums. Pass 2: each target's chunk by binary search, then one wave walks it in 64-bin strides (loaded eight strides
// at a time) with a wave prefix sum and takes the first lane above the target (the next chunk if rounding left it
// just short).
constexpr int EMAXCH = 4096;

template <bool MULTI = false>
__global__ void __launch_bounds__(NT) edges_fused_k(const float2* __restrict__ h, const long ldh,
                                                    const int* __restrict__ rows, const float* __restrict__ ips,
                                                    const int* __restrict__ k1s, const int* __restrict__ nbs,
                                                    const int nedge, const int kmin, int* __restrict__ out,
                                                    const float* const* __restrict__ iptab,
                                                    const int* __restrict__ rseg)
{
    __shared__ double ex[EMAXCH + 1];
    __shared__ double sh[NT / 64];
    const int g = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    if (MULTI) ips = iptab[rseg[g]];          // the request's segment (its 1/psd row)
    const float2* __restrict__ hr = h + (size_t)rows[g] * ldh;
    const long k1 = k1s[g];
    const int nb = nbs[g];
    int* eo = out + (size_t)g * nedge;
    for (int i = tid; i < nedge; i += NT) eo[i] = (int)k1;
    // h and 1/psd of the row as buffers ending at k1 (byte offsets below 2^31): bins past k1 read zeros and add
    // exact zeros, so the loops below have the same trip count on every lane and carry no branch around their loads
    const __amdgpu_buffer_rsrc_t rh = buffer_of(hr, k1 * 8), ri = buffer_of(ips, k1 * 4);
    auto pw_at = [&](const long k) {
        const float2 v = load_f2(rh, (unsigned)k * 8u);
        return ((double)v.x * v.x + (double)v.y * v.y) * (double)load_f1(ri, (unsigned)k * 4u);
    };
    // chunk length: a multiple of 64, at least 2048, and at most EMAXCH chunks per row
    const long ech = max(2048L, ((k1 + EMAXCH - 1) / EMAXCH + 63) / 64 * 64);
    const int nch = (int)((k1 + ech - 1) / ech);
    for (int c = wave; c < nch; c += NT / 64) {
        const long c0 = (long)c * ech + lane;
        double s = 0.0;
#pragma unroll 8
        for (int i = 0; i < (int)(ech >> 6); i++) s += pw_at(c0 + 64 * i);
        s = wave_sum_f64(s);
        if (lane == 0) ex[c] = s;
    }
    __syncthreads();
    // exclusive scan of the chunk sums in place (each thread a contiguous run, then a block scan of the runs)
    const int per = (nch + NT - 1) / NT, a = min(tid * per, nch), b = min(a + per, nch);
    double loc = 0.0;
    for (int c = a; c < b; c++) loc += ex[c];
    double tot;
    double run = block_exscan_f64(loc, sh, tot);
    __syncthreads();
    for (int c = a; c < b; c++) { const double x = ex[c]; ex[c] = run; run += x; }
    if (tid == 0) ex[nch] = tot;
    __syncthreads();
    if (nch == 0) return;
    for (int j = wave; j < nb; j += NT / 64) {
        const double target = ((double)j / (double)nb) * tot;
        // last chunk whose start lies at or below the target
        int lo = 0, hi = nch - 1;
        while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (ex[mid] <= target) lo = mid; else hi = mid - 1; }
        long k = k1;
        double acc = ex[lo];
        // the walk in groups of WG strides of 64 bins: a group's loads are issued together, then scanned in order
        constexpr int WG = 8;
        bool found = false;
        for (long base = (long)lo * ech; base < k1 && !found; base += 64 * WG) {
            float2 hv[WG];
            float iv[WG];
#pragma unroll
            for (int s = 0; s < WG; s++) {
                const unsigned o = (unsigned)(base + 64 * s + lane);
                hv[s] = load_f2(rh, o * 8u);
                iv[s] = load_f1(ri, o * 4u);
            }
#pragma unroll
            for (int s = 0; s < WG; s++) {
                const long kk = base + 64 * s + lane;
                double v = ((double)hv[s].x * hv[s].x + (double)hv[s].y * hv[s].y) * (double)iv[s];
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const double y = __shfl_up(v, o, 64);
                    if (lane >= o) v += y;
                }
                const unsigned long long m = __ballot(acc + v > target && kk < k1);
                if (m) { k = base + 64 * s + __ffsll((long long)m) - 1; found = true; break; }
                acc += __shfl(v, 63, 64);
            }
        }
        if (lane == 0) eo[j] = (int)min(max(k, (long)kmin), k1);
    }
}

// Hill climbing on the exact |z|, z(t) = sum over k0 <= k < k1 of conj(h_k) d_k exp(2 pi i k t / N), for peaks whose
// re-centring range ended on its edge: one workgroup per peak r, from start[r]: z at the 2 W + 1 samples c - W ..
// c + W clipped to [lo, hi - 1]; while the largest |z| (first sample on ties) lies on the range's edge inside the
// bounds, the range re-centres on it; out[r] = the range's largest sample at the end (after at most max_iter
// ranges). k0, k1: the first and last entries of the peak's chi-squared edge row (the z the bin sums add up to).
// Lanes take bins k0 + tid + NT s, each point's phasor reseeded from the exact integer phase every CR steps.
constexpr int CW = 3;            // largest half width
constexpr int CP = 2 * CW + 1;
constexpr int CR = 32;
__global__ void __launch_bounds__(NT) climb_k(const float2* __restrict__ h, const long ldh,
                                              const float2* const* __restrict__ dtab, const int* __restrict__ rows,
                                              const int* __restrict__ segs, const int* __restrict__ edges,
                                              const int nedge, const int* __restrict__ eidx,
                                              const long long* __restrict__ start, const long long* __restrict__ lo,
                                              const long long* __restrict__ hi, const unsigned long long N,
                                              const int W, const int max_iter, const int n,
                                              long long* __restrict__ out)
{
    __shared__ float2 red[NT / 64][CP];
    __shared__ long long sc;
    __shared__ int sdone;
    const int r = blockIdx.x;
    if (r >= n) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int* e = edges + (size_t)eidx[r] * nedge;
    const int k0 = e[0], k1 = e[nedge - 1];
    long long c = start[r];
    const long long L = lo[r], Hb = hi[r] - 1;
    if (k1 <= k0) {                               // no bins: z = 0, nothing to climb
        if (tid == 0) out[r] = c;
        return;
    }
    const unsigned long long mask = (N & (N - 1)) == 0 && N <= (1ull << 24) ? N - 1 : 0;
    const float2* hr = h + (size_t)rows[r] * ldh;
    const float2* dv = dtab[segs[r]];
    const __amdgpu_buffer_rsrc_t rh = buffer_of(hr, (long)k1 * 8), rd = buffer_of(dv, (long)k1 * 8);
    const int np_ = 2 * W + 1;
    for (int it = 0; it < max_iter; it++) {
        unsigned long long t[CP];
        long long ts[CP];
#pragma unroll
        for (int p = 0; p < CP; p++) {
            ts[p] = min(max(c - W + min(p, np_ - 1), L), Hb);
            t[p] = (unsigned long long)ts[p] % N;
        }
        float2 acc[CP], rot[CP];
#pragma unroll
        for (int p = 0; p < CP; p++) { acc[p] = make_float2(0.f, 0.f); rot[p] = phasor(NT, t[p], N, mask); }
        for (int ra = k0; ra < k1; ra += NT * CR) {
            const int ka = ra + tid;
            float2 w[CP];
#pragma unroll
            for (int p = 0; p < CP; p++) w[p] = phasor((unsigned long long)ka, t[p], N, mask);
            const int ns = (min(ra + NT * CR, k1) - ra + NT - 1) / NT;   // the same for every lane
            for (int st = 0; st < ns; st++) {
                const int kk = ka + NT * st;
                // past k1 the buffers read zeros: exact zero terms
                const float2 qv = load_f2(rh, (unsigned)kk * 8u), dk = load_f2(rd, (unsigned)kk * 8u);
                const float2 v = make_float2(qv.x * dk.x + qv.y * dk.y, qv.x * dk.y - qv.y * dk.x);
#pragma unroll
                for (int p = 0; p < CP; p++) {
                    acc[p].x = fmaf(v.x, w[p].x, fmaf(-v.y, w[p].y, acc[p].x));
                    acc[p].y = fmaf(v.x, w[p].y, fmaf(v.y, w[p].x, acc[p].y));
                    const float wx = fmaf(w[p].x, rot[p].x, -w[p].y * rot[p].y);
                    w[p].y = fmaf(w[p].x, rot[p].y, w[p].y * rot[p].x);
                    w[p].x = wx;
                }
            }
        }
#pragma unroll
        for (int p = 0; p < CP; p++) {
            const float ax = wave_sum_f32(acc[p].x), ay = wave_sum_f32(acc[p].y);
            if (lane == 0) red[wave][p] = make_float2(ax, ay);
        }
        __syncthreads();
        if (tid == 0) {
            int j = 0;
            double bv = -1.0;
            for (int p = 0; p < np_; p++) {
                double zx = 0.0, zy = 0.0;
                for (int w2 = 0; w2 < NT / 64; w2++) { zx += red[w2][p].x; zy += red[w2][p].y; }
                const double v2 = zx * zx + zy * zy;
                if (v2 > bv) { bv = v2; j = p; }
            }
            const bool edge = (j == 0 && ts[0] > L) || (j == np_ - 1 && ts[np_ - 1] < Hb);
            sc = ts[j];
            sdone = !edge;
        }
        __syncthreads();
        c = sc;
        if (sdone) break;
        __syncthreads();                          // every thread has read sc before thread 0 writes it again
    }
    if (tid == 0) out[r] = c;
}

// ---- the regenerated exact stage (the reference's --exact-stage regen, exact_stage.py:472-647) --------------------
// Templates given by coarse nodes (templates.hip spa_nodes_k): on the fine bins of [kmin, kend), bin q = k - kmin lies
// in interval node[q] with d[q] = f - x_j and t[q] = d / (x_j+1 - x_j); h_k = (A_j (1 - t) + A_j+1 t) exp(i p_j(d)).

// a^2 p0 + 2 a b p1 + b^2 p2 in a fixed association without contraction, so an interval's weight from its totals
// equals its in-interval prefix at its last bin bit for bit
__device__ __forceinline__ double quad3(const double a, const double b, const double* m)
{
    return __dadd_rn(__dadd_rn(__dmul_rn(__dmul_rn(a, a), m[0]), __dmul_rn(__dmul_rn(__dmul_rn(2.0, a), b), m[1])),
                     __dmul_rn(__dmul_rn(b, b), m[2]));
}

// Chi-squared edges of the interpolated templates from the segment's moments of 1/psd: Mi (J, 3) the interval totals
// of ((1-t)^2, t(1-t), t^2) IP and Pk (K, 3) their in-interval inclusive prefixes at each bin, so that the weight
// sum of |h|^2 IP over interval j is quad3(A_j, A_j+1, Mi[j]) and up to bin k of it quad3(A_j, A_j+1, Pk[k]). One
// block per requested row: each thread sums a contiguous run of intervals, a block scan of the runs, then per
// target j / nb * total the run (binary search), the interval (a walk from the run's start) and the bin (binary
// search on the prefixes): edge = first bin whose running sum exceeds the target. Columns >= nb hold the row's end
// k1s[u] (<= kend; the row is zero from there on).
__global__ void __launch_bounds__(NT) edges_moments_k(const double* __restrict__ amp, const int nx,
                                                      const double* const* __restrict__ Mi_tab,
                                                      const double* const* __restrict__ Pk_tab,
                                                      const int* __restrict__ rseg, const int* __restrict__ kb,
                                                      const int* __restrict__ rows, const int* __restrict__ nbs,
                                                      const int* __restrict__ k1s, const int nedge, const int kmin,
                                                      int* __restrict__ out)
{
    __shared__ double pre[NT + 1];
    __shared__ double sh[NT / 64];
    const int u = blockIdx.x, tid = threadIdx.x, J = nx - 1;
    const double* __restrict__ Mi = Mi_tab[rseg[u]];     // the request's segment
    const double* __restrict__ Pk = Pk_tab[rseg[u]];
    const double* A = amp + (size_t)rows[u] * nx;
    const int nb = nbs[u], kend = k1s[u];
    const int per = (J + NT - 1) / NT, j0 = min(tid * per, J), j1 = min(j0 + per, J);
    double s = 0.0;
    for (int j = j0; j < j1; j++) s = __dadd_rn(s, quad3(A[j], A[j + 1], Mi + 3 * (size_t)j));
    double tot;
    const double ex = block_exscan_f64(s, sh, tot);
    pre[tid] = ex;
    if (tid == 0) pre[NT] = tot;
    __syncthreads();
    int* e = out + (size_t)u * nedge;
    for (int i = tid; i < nedge; i += NT) {
        if (i >= nb) { e[i] = kend; continue; }
        const double T = ((double)i / (double)nb) * tot;
        // the last run whose start lies at or below T
        int lo = 0, hi = NT - 1;
        while (lo < hi) { const int m = (lo + hi + 1) >> 1; if (pre[m] <= T) lo = m; else hi = m - 1; }
        // the walk from the run's start (past its end if the scan's association left the run just short)
        const int qa = min(lo * per, J);
        double r = pre[lo];
        int j = 0;
        bool hit = false;
        for (int jj = qa; jj < J; jj++) {
            const double w = quad3(A[jj], A[jj + 1], Mi + 3 * (size_t)jj);
            if (__dadd_rn(r, w) > T) { j = jj; hit = true; break; }
            r = __dadd_rn(r, w);
        }
        if (!hit) { e[i] = kend; continue; }
        const double a = A[j], b = A[j + 1];
        int klo = kb[j] - kmin, khi = kb[j + 1] - kmin - 1;
        while (klo < khi) {
            const int m = (klo + khi) >> 1;
            if (__dadd_rn(r, quad3(a, b, Pk + 3 * (size_t)m)) > T) khi = m; else klo = m + 1;
        }
        e[i] = min(max(klo + kmin, kmin), kend);
    }
}

// chisq_bins_k with the template regenerated on the fly (no reads of H): per bin, the interval's float32 cubic
// (c3, c2, c1, c0 with c0 reduced to [-pi, pi)) at d, reduced by 2 pi, __sincosf, the amplitude lerp; then
// conj(h) D and the P peak phasors as in chisq_bins_k. coef32 (rows, J) float4, amp2 (rows, J) float2.
template <bool MULTI>
__global__ void __launch_bounds__(NT) regen_bins_k(const float4* __restrict__ coef32, const float2* __restrict__ amp2,
                                                   const int J, const int* __restrict__ node,
                                                   const float* __restrict__ dl, const float* __restrict__ tl,
                                                   const int kmin, const float2* __restrict__ dv,
                                                   const unsigned long long N, const int* __restrict__ grow,
                                                   const int* __restrict__ gstart, const int* __restrict__ gcount,
                                                   const long long* __restrict__ ptime, const int* __restrict__ edges,
                                                   const int nedge, const int ngroups, float2* __restrict__ zp,
                                                   const float2* const* __restrict__ dtab,
                                                   const int* __restrict__ gseg)
{
    constexpr float TWO_PI = 6.283185307179586f, INV_TWO_PI = 0.15915494309189535f;
    const int nbin = nedge - 1;
    const int lane = threadIdx.x & 63;
    const long wi = (long)blockIdx.x * (NT / 64) + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (wi >= (long)ngroups * nbin) return;
    const int g = (int)(wi / nbin), j = (int)(wi % nbin);
    const int row = grow[g], p0 = gstart[g], np_ = gcount[g];
    const int* e = edges + (size_t)g * nedge;
    const int k0 = e[j], k1 = e[j + 1];
    float2* zo = zp + (size_t)wi * P;
    if (k0 >= k1) {
        if (lane < P) zo[lane] = make_float2(0.f, 0.f);
        return;
    }
    if (MULTI) dv = dtab[gseg[g]];
    const unsigned long long mask = (N & (N - 1)) == 0 && N <= (1ull << 24) ? N - 1 : 0;
    const float4* cr = coef32 + (size_t)row * J;
    const float2* ar = amp2 + (size_t)row * J;
    unsigned long long t[P];
    float2 rot[P];
#pragma unroll
    for (int p = 0; p < P; p++) {
        t[p] = p < np_ ? (unsigned long long)ptime[p0 + p] % N : 0ull;
        rot[p] = phasor(64ull, t[p], N, mask);
    }
    float2 acc[P];
#pragma unroll
    for (int p = 0; p < P; p++) acc[p] = make_float2(0.f, 0.f);
    constexpr int RW = 64 * R;
    for (int ra = k0; ra < k1; ra += RW) {
        const int ka = ra + lane;
        const int ns = (min(ra + RW, k1) - ra + 63) >> 6;
        float2 w[P];
#pragma unroll
        for (int p = 0; p < P; p++) w[p] = p < np_ ? phasor((unsigned long long)ka, t[p], N, mask) : make_float2(0.f, 0.f);
        for (int st = 0; st < ns; st++) {
            const int k = ka + 64 * st;
            float2 v = make_float2(0.f, 0.f);
            if (k < k1) {
                const int q = k - kmin, jn = node[q];
                const float4 c = cr[jn];
                const float2 am = ar[jn];
                const float dd = dl[q];
                float x = fmaf(fmaf(fmaf(c.x, dd, c.y), dd, c.z), dd, c.w);
                x = fmaf(-TWO_PI, rintf(x * INV_TWO_PI), x);
                const float a = fmaf(am.y - am.x, tl[q], am.x);
                float sn, cs;
                __sincosf(x, &sn, &cs);
                const float hr = a * cs, hi = a * sn;
                const float2 dk = dv[k];
                v = make_float2(fmaf(hr, dk.x, hi * dk.y), fmaf(hr, dk.y, -hi * dk.x));   // conj(h) D
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
        if (lane == p) zo[p] = make_float2(ax, ay);
    }
}

} // namespace

// Groups: for g in [0, ngroups): row grow[g] of q (or h in fused mode), peaks gstart[g] .. gstart[g] + gcount[g]
// (gcount <= P, which gwgs_chisq_peaks_per_group returns) of ptime, bin edges edges[g, 0:nedge]. N: the time length
// of the phase factor (any positive value; the phase k t mod N is exact in 64-bit integers). d: data (fused mode) or
// null. zout may be null. work: device buffer of gwgs_chisq_points_workspace(ngroups, nedge) bytes (the per-bin
// sums).
GWGS_API long gwgs_chisq_points_workspace(int ngroups, int nedge)
{
    return ngroups > 0 && nedge >= 2 ? (long)ngroups * (nedge - 1) * P * (long)sizeof(float2) : 0;
}

GWGS_API int gwgs_chisq_points(const void* q, long ldq, const void* d, long long N, const int* grow, const int* gstart,
                               const int* gcount, const long long* ptime, const int* edges, int nedge, int ngroups,
                               float* out, void* zout, void* work, void* stream)
{
    GWGS_CHECK(q && grow && gstart && gcount && ptime && edges && out && N > 0 && nedge >= 1 && ngroups >= 0,
               GWGS_EINVAL);
    if (ngroups == 0) return GWGS_OK;
    hipStream_t s = (hipStream_t)stream;
    if (nedge == 1) {                        // no bins: every sum is zero
        hipLaunchKernelGGL(chisq_finish_k, dim3((unsigned)(((long)ngroups * P + NT - 1) / NT)), dim3(NT), 0, s,
                           (const float2*)nullptr, gstart, gcount, edges, nedge, ngroups, out, (float2*)zout);
        return gwgs_status();
    }
    GWGS_CHECK(work != nullptr, GWGS_EINVAL);
    GWGS_CHECK((long)ngroups * (nedge - 1) < (1L << 33), GWGS_ESIZE);
    const dim3 grid((unsigned)(((long)ngroups * (nedge - 1) + NT / 64 - 1) / (NT / 64)));
    if (d)
        hipLaunchKernelGGL(chisq_bins_k<true>, grid, dim3(NT), 0, s, (const float2*)q, ldq, (const float2*)d,
                           (unsigned long long)N, grow, gstart, gcount, ptime, edges, nedge, ngroups, (float2*)work,
                           nullptr, nullptr);
    else
        hipLaunchKernelGGL(chisq_bins_k<false>, grid, dim3(NT), 0, s, (const float2*)q, ldq, nullptr,
                           (unsigned long long)N, grow, gstart, gcount, ptime, edges, nedge, ngroups, (float2*)work,
                           nullptr, nullptr);
    hipLaunchKernelGGL(chisq_finish_k, dim3((unsigned)(((long)ngroups * P + NT - 1) / NT)), dim3(NT), 0, s,
                       (const float2*)work, gstart, gcount, edges, nedge, ngroups, out, (float2*)zout);
    return gwgs_status();
}

// gwgs_chisq_points in fused mode over groups of several segments: group g reads its data row from
// d_tab[gseg[g]] (device table of complex64 pointers, every row >= the largest edge long); everything else as
// gwgs_chisq_points, whose workspace size applies.
GWGS_API int gwgs_chisq_points_multi(const void* h, long ldh, const void* const* d_tab, const int* gseg, long long N,
                                     const int* grow, const int* gstart, const int* gcount, const long long* ptime,
                                     const int* edges, int nedge, int ngroups, float* out, void* zout, void* work,
                                     void* stream)
{
    GWGS_CHECK(h && d_tab && gseg && grow && gstart && gcount && ptime && edges && out && N > 0 && nedge >= 1 &&
               ngroups >= 0, GWGS_EINVAL);
    if (ngroups == 0) return GWGS_OK;
    hipStream_t s = (hipStream_t)stream;
    if (nedge == 1) {
        hipLaunchKernelGGL(chisq_finish_k, dim3((unsigned)(((long)ngroups * P + NT - 1) / NT)), dim3(NT), 0, s,
                           (const float2*)nullptr, gstart, gcount, edges, nedge, ngroups, out, (float2*)zout);
        return gwgs_status();
    }
    GWGS_CHECK(work != nullptr, GWGS_EINVAL);
    GWGS_CHECK((long)ngroups * (nedge - 1) < (1L << 33), GWGS_ESIZE);
    const dim3 grid((unsigned)(((long)ngroups * (nedge - 1) + NT / 64 - 1) / (NT / 64)));
    hipLaunchKernelGGL((chisq_bins_k<true, true>), grid, dim3(NT), 0, s, (const float2*)h, ldh, nullptr,
                       (unsigned long long)N, grow, gstart, gcount, ptime, edges, nedge, ngroups, (float2*)work,
                       (const float2* const*)d_tab, gseg);
    hipLaunchKernelGGL(chisq_finish_k, dim3((unsigned)(((long)ngroups * P + NT - 1) / NT)), dim3(NT), 0, s,
                       (const float2*)work, gstart, gcount, edges, nedge, ngroups, out, (float2*)zout);
    return gwgs_status();
}

GWGS_API int gwgs_chisq_peaks_per_group(void) { return P; }

// Hill climbing of n peaks on the exact |z| (see climb_k): rows, segs (int32) the peak's template row of h and data
// row d_tab[segs[r]]; edges (rows of nedge int32) with eidx[r] the peak's edge row; start, lo, hi (int64); W in
// [1, 3] the range half width; out (int64) the final samples.
GWGS_API int gwgs_climb_points(const void* h, long ldh, const void* const* d_tab, long long N, const int* rows,
                               const int* segs, const int* edges, int nedge, const int* eidx, const long long* start,
                               const long long* lo, const long long* hi, int W, int max_iter, int n, long long* out,
                               void* stream)
{
    GWGS_CHECK(h && d_tab && rows && segs && edges && eidx && start && lo && hi && out && N > 0 && nedge >= 2 &&
               W >= 1 && W <= CW && max_iter >= 1 && n >= 0, GWGS_EINVAL);
    if (n == 0) return GWGS_OK;
    hipLaunchKernelGGL(climb_k, dim3((unsigned)n), dim3(NT), 0, (hipStream_t)stream, (const float2*)h, ldh,
                       (const float2* const*)d_tab, rows, segs, edges, nedge, eidx, start, lo, hi,
                       (unsigned long long)N, W, max_iter, n, out);
    return gwgs_status();
}

// Edges for n requested rows: rows index h (ldh complex per row); ips: 1/psd (float32, >= max k1 entries);
// k1s, nbs per request; out (n, nedge) int32.
GWGS_API int gwgs_edges_fused(const void* h, long ldh, const int* rows, const float* ips, const int* k1s,
                              const int* nbs, int nedge, int kmin, int n, int* out, void* stream)
{
    GWGS_CHECK(h && rows && ips && k1s && nbs && out && nedge >= 2 && n >= 0, GWGS_EINVAL);
    if (n == 0) return GWGS_OK;
    hipLaunchKernelGGL(edges_fused_k<false>, dim3(n), dim3(NT), 0, (hipStream_t)stream, (const float2*)h, ldh, rows,
                       ips, k1s, nbs, nedge, kmin, out, nullptr, nullptr);
    return gwgs_status();
}

// gwgs_edges_fused over requests of several segments: request g reads 1/psd from ip_tab[rseg[g]] (device table of
// float32 pointers, every row >= its k1 long).
GWGS_API int gwgs_edges_fused_multi(const void* h, long ldh, const int* rows, const float* const* ip_tab,
                                    const int* rseg, const int* k1s, const int* nbs, int nedge, int kmin, int n,
                                    int* out, void* stream)
{
    GWGS_CHECK(h && rows && ip_tab && rseg && k1s && nbs && out && nedge >= 2 && n >= 0, GWGS_EINVAL);
    if (n == 0) return GWGS_OK;
    hipLaunchKernelGGL(edges_fused_k<true>, dim3(n), dim3(NT), 0, (hipStream_t)stream, (const float2*)h, ldh, rows,
                       nullptr, k1s, nbs, nedge, kmin, out, (const float* const*)ip_tab, rseg);
    return gwgs_status();
}

// Chi-squared edges of n (segment, row) requests of interpolated templates from the segments' moments (see
// edges_moments_k): amp (rows, nx) float64 node amplitudes; per segment s Mi_tab[s] (nx - 1, 3) and Pk_tab[s]
// (kend - kmin, 3) float64 (device tables of pointers), rseg (n) each request's segment; kb (nx) int32 the first bin
// of each interval (kb[nx - 1] = kend); k1s (n) each row's end (<= kend, zero beyond); out (n, nedge) int32.
GWGS_API int gwgs_edges_moments(const double* amp, int nx, const void* const* Mi_tab, const void* const* Pk_tab,
                                const int* rseg, const int* kb, const int* rows, const int* nbs, const int* k1s,
                                int nedge, int kmin, int n, int* out, void* stream)
{
    GWGS_CHECK(amp && Mi_tab && Pk_tab && rseg && kb && rows && nbs && k1s && out && nx >= 2 && nedge >= 1 &&
               kmin >= 0 && n >= 0, GWGS_EINVAL);
    if (n == 0) return GWGS_OK;
    hipLaunchKernelGGL(edges_moments_k, dim3(n), dim3(NT), 0, (hipStream_t)stream, amp, nx,
                       (const double* const*)Mi_tab, (const double* const*)Pk_tab, rseg, kb, rows, nbs, k1s, nedge,
                       kmin, out);
    return gwgs_status();
}

// gwgs_chisq_points_multi for templates regenerated from coarse nodes instead of read from H: coef32 (rows, nx - 1)
// float4 (c3, c2, c1, c0 reduced), amp2 (rows, nx - 1) float2 (A_j, A_j+1); node, d, t over the kend - kmin fine bins
// (int32, float32, float32); every edge within [kmin, kend]. Workspace as gwgs_chisq_points_workspace.
GWGS_API int gwgs_regen_points_multi(const void* coef32, const void* amp2, int nx, const int* node, const float* d,
                                     const float* t, int kmin, const void* const* d_tab, const int* gseg, long long N,
                                     const int* grow, const int* gstart, const int* gcount, const long long* ptime,
                                     const int* edges, int nedge, int ngroups, float* out, void* zout, void* work,
                                     void* stream)
{
    GWGS_CHECK(coef32 && amp2 && node && d && t && d_tab && gseg && grow && gstart && gcount && ptime && edges && out &&
               N > 0 && nx >= 2 && nedge >= 1 && ngroups >= 0, GWGS_EINVAL);
    if (ngroups == 0) return GWGS_OK;
    hipStream_t s = (hipStream_t)stream;
    if (nedge == 1) {
        hipLaunchKernelGGL(chisq_finish_k, dim3((unsigned)(((long)ngroups * P + NT - 1) / NT)), dim3(NT), 0, s,
                           (const float2*)nullptr, gstart, gcount, edges, nedge, ngroups, out, (float2*)zout);
        return gwgs_status();
    }
    GWGS_CHECK(work != nullptr, GWGS_EINVAL);
    GWGS_CHECK((long)ngroups * (nedge - 1) < (1L << 33), GWGS_ESIZE);
    const dim3 grid((unsigned)(((long)ngroups * (nedge - 1) + NT / 64 - 1) / (NT / 64)));
    hipLaunchKernelGGL(regen_bins_k<true>, grid, dim3(NT), 0, s, (const float4*)coef32, (const float2*)amp2, nx - 1,
                       node, d, t, kmin, nullptr, (unsigned long long)N, grow, gstart, gcount, ptime, edges, nedge,
                       ngroups, (float2*)work, (const float2* const*)d_tab, gseg);
    hipLaunchKernelGGL(chisq_finish_k, dim3((unsigned)(((long)ngroups * P + NT - 1) / NT)), dim3(NT), 0, s,
                       (const float2*)work, gstart, gcount, edges, nedge, ngroups, out, (float2*)zout);
    return gwgs_status();
}
