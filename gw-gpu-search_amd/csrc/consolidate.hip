// Pre-consolidation of the window-maximum candidates (the reference's pre_keep, fp_filter.py:665-695), applied to
// the GEMM's window maxima vw of one batch before the candidate list is compacted: a candidate vw[g] > cut that
// the trigger consolidation would drop anyway is overwritten with -1 (below any cut), so the ordinary ordered
// compaction that follows never sees it.
//
// A candidate g = row * nwsum + column of segment si (columns [col0, col0 + ncol) of the batch) has apx =
// sqrt(vw[g]) (float32, correctly rounded) and end time t = gps + (sbase + iw[g]) / sr (float64, sbase = the
// segment's start sample plus its first analysed sample). Exact-band candidates (apx >= exact_cut) are kept.
// A rebuilt one is dropped unless apx > thr and (apx >= nthr or exempt, exempt = apx >= exempt_snr), and unless
// |t - g_t| > reach_g for every gate of its dataset. With an interval, a surviving non-exempt candidate is ranked in
// its keep-loudest window wt = trunc(rint((t - gps) sr) / wdiv) (per segment: the part of the reference is one tile
// and segment): first dropped if its dataset has a floor and !(apx > floor[wt]) (windows past the floor's end pass),
// then only the `num` loudest of its window are kept, by apx descending and, on ties, by g ascending (row-major
// order, the reference's stable sort by position). Exempt survivors are kept.
//
// Passes: count (candidates per window, atomics), exclusive scan of the window counts, scatter (every ranked
// candidate's 64-bit key into its window's bucket; the drops of the cut, the gates and the floor are written here),
// rank (one workgroup per window: up to 1,024 keys, each key's rank counted directly from LDS; more, the num-th
// largest key by num max-reductions; keys below it are dropped). The key
// is (float bits of apx) << 32 | ~g: apx > 0, so larger keys are louder, and the earlier g wins a tie. Each entry is
// classified twice (count and scatter) from the same vw[g] and iw[g], which no pass changes before the scatter
// writes its own entry.
#include "common.h"

namespace {

constexpr int NT = 256;
constexpr int SEGI = 10;         // int64 fields per segment in segi
constexpr int LDS_KEYS = 4096;   // a window's keys are staged in LDS when it has at most this many

// segi row: sbase, wt0, col0, ncol, wbase, nwin, gate0, ngate, fl0, nfl
struct Seg {
    long sbase, wt0;
    int col0, ncol, wbase, nwin, gate0, ngate, fl0, nfl;
};

__device__ __forceinline__ Seg load_seg(const long* __restrict__ segi, const int si)
{
    const long* s = segi + (long)si * SEGI;
    Seg r;
    r.sbase = s[0]; r.wt0 = s[1];
    r.col0 = (int)s[2]; r.ncol = (int)s[3]; r.wbase = (int)s[4]; r.nwin = (int)s[5];
    r.gate0 = (int)s[6]; r.ngate = (int)s[7]; r.fl0 = (int)s[8]; r.nfl = (int)s[9];
    return r;
}

struct Params {
    int nwsum, nseg, num, flags;            // flags: 1 exempt threshold set, 2 keep-loudest interval set
    float exact_cut, thr, nthr, exempt_snr;
    double sr, wdiv;
};

constexpr int KEEP = -1, DROP = -2;

// KEEP, DROP, or the batch window (>= 0) in which the candidate is ranked
__device__ __forceinline__ int classify(const float v, const int g, const int* __restrict__ iw,
                                        const long* __restrict__ segi, const double* __restrict__ segd,
                                        const double* __restrict__ gates, const double* __restrict__ floors,
                                        const Params& p)
{
    const float apx = sqrtf(v);
    if (apx >= p.exact_cut) return KEEP;
    if (!(apx > p.thr)) return DROP;
    const bool exempt = (p.flags & 1) && apx >= p.exempt_snr;
    if (!(apx >= p.nthr) && !exempt) return DROP;
    const int c = g % p.nwsum;
    int si = 0;
    while (si + 1 < p.nseg && c >= (int)segi[(long)(si + 1) * SEGI + 2]) si++;
    const Seg s = load_seg(segi, si);
    const double gps = segd[si];
    const double t = gps + (double)(s.sbase + iw[g]) / p.sr;
    for (int j = 0; j < s.ngate; j++)
        if (!(fabs(t - gates[2 * (s.gate0 + j)]) > gates[2 * (s.gate0 + j) + 1])) return DROP;
    if (exempt || !(p.flags & 2)) return KEEP;
    const long wt = (long)(rint((t - gps) * p.sr) / p.wdiv);
    const long w = wt - s.wt0;
    if (w < 0 || w >= s.nwin) return KEEP;     // outside the segment's window range (not reached): kept
    if (s.nfl > 0) {
        const double bar = wt < s.nfl ? floors[s.fl0 + wt] : -INFINITY;
        if (!((double)apx > bar)) return DROP;
    }
    return s.wbase + (int)w;
}

__global__ void __launch_bounds__(NT) zero_k(int* __restrict__ a, const int n)
{
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i < n) a[i] = 0;
}

// PASS 0: count ranked candidates per window. PASS 1: write drops, scatter ranked keys.
template <int PASS>
__global__ void __launch_bounds__(NT) prekeep_pass_k(float* __restrict__ vw, const int* __restrict__ iw,
                                                     const long total, const float cut, const long* __restrict__ segi,
                                                     const double* __restrict__ segd, const double* __restrict__ gates,
                                                     const double* __restrict__ floors, const Params p,
                                                     int* __restrict__ counts, int* __restrict__ pos,
                                                     unsigned long long* __restrict__ bucket)
{
    const long g = (long)blockIdx.x * NT + threadIdx.x;
    if (g >= total) return;
    const float v = vw[g];
    if (!(v > cut)) return;
    const int k = classify(v, (int)g, iw, segi, segd, gates, floors, p);
    if (PASS == 0) {
        if (k >= 0) atomicAdd(&counts[k], 1);
    } else if (k == DROP) {
        vw[g] = -1.0f;
    } else if (k >= 0) {
        const int at = atomicAdd(&pos[k], 1);
        bucket[at] = ((unsigned long long)__float_as_uint(sqrtf(v)) << 32) | (unsigned)(~(unsigned)g);
    }
}

// exclusive scan of counts[0:n] into off[0:n+1] and pos[0:n] (one block; each thread owns a contiguous run)
__global__ void __launch_bounds__(1024) scan_k(const int* __restrict__ counts, const int n, int* __restrict__ off,
                                               int* __restrict__ pos)
{
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int a = min(t * per, n), b = min(a + per, n);
    int s = 0;
    for (int i = a; i < b; i++) s += counts[i];
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = t ? part[t - 1] : 0;
    for (int i = a; i < b; i++) { off[i] = run; pos[i] = run; run += counts[i]; }
    if (t == 1023) off[n] = part[1023];
}

__device__ __forceinline__ unsigned long long shfl_xor_u64(const unsigned long long x, const int o)
{
    const unsigned lo = __shfl_xor((unsigned)x, o, 64), hi = __shfl_xor((unsigned)(x >> 32), o, 64);
    return ((unsigned long long)hi << 32) | lo;
}

// one workgroup per window: keep the num largest keys, drop the rest
__global__ void __launch_bounds__(NT) prekeep_rank_k(float* __restrict__ vw, const int* __restrict__ off,
                                                     const unsigned long long* __restrict__ bucket, const int num)
{
    const int w = blockIdx.x;
    const int a = off[w], n = off[w + 1] - a;
    if (n <= num) return;
    __shared__ unsigned long long keys[LDS_KEYS];
    __shared__ unsigned long long red[NT / 64];
    const bool staged = n <= LDS_KEYS;
    const unsigned long long* B = bucket + a;
    if (staged) {
        for (int i = threadIdx.x; i < n; i += NT) keys[i] = B[i];
        __syncthreads();
    }
    if (staged && n <= 4 * NT) {
        // few keys: each key's rank is the number of larger keys (keys are distinct), counted from LDS
        for (int i = threadIdx.x; i < n; i += NT) {
            const unsigned long long k = keys[i];
            int rank = 0;
            for (int j = 0; j < n; j++) rank += keys[j] > k;
            if (rank >= num) vw[(unsigned)~(unsigned)k] = -1.0f;
        }
        return;
    }
    const unsigned long long* K = staged ? keys : B;
    unsigned long long lim = ~0ull;             // after pass r: the r-th largest key
    for (int r = 0; r < num; r++) {
        unsigned long long best = 0;
        for (int i = threadIdx.x; i < n; i += NT) {
            const unsigned long long k = K[i];
            if (k < lim && k > best) best = k;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned long long x = shfl_xor_u64(best, o);
            best = x > best ? x : best;
        }
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = best;
        __syncthreads();
        lim = red[0];
#pragma unroll
        for (int j = 1; j < NT / 64; j++) lim = red[j] > lim ? red[j] : lim;
        __syncthreads();
    }
    for (int i = threadIdx.x; i < n; i += NT) {
        const unsigned long long k = K[i];
        if (k < lim) vw[(unsigned)~(unsigned)k] = -1.0f;
    }
}

}  // namespace

// Workspace bytes of gwgs_pre_keep for `total` window maxima and `nwin` keep-loudest windows in all.
GWGS_API long gwgs_pre_keep_workspace(long total, long nwin)
{
    return 8 * total + 4 * (3 * nwin + 1);
}

GWGS_API int gwgs_pre_keep(float* vw, const int* iw, int n, int nwsum, int nseg, const long* segi, const double* segd,
                           const double* gates, const double* floors, double sr, double wdiv, int num, float cut,
                           float exact_cut, float thr, float nthr, float exempt_snr, int flags, int nwin, void* work,
                           long work_bytes, void* stream)
{
    const long total = (long)n * nwsum;
    GWGS_CHECK(n >= 0 && nwsum >= 0 && nseg >= 1 && nwin >= 0 && (flags & ~3) == 0, GWGS_EINVAL);
    GWGS_CHECK(total < (1L << 31), GWGS_ESIZE);
    if (total == 0) return GWGS_OK;
    GWGS_CHECK(vw && iw && segi && segd && sr > 0, GWGS_EINVAL);
    const bool ranked = (flags & 2) != 0;
    GWGS_CHECK(!ranked || (num >= 1 && wdiv > 0 && work && work_bytes >= gwgs_pre_keep_workspace(total, nwin)),
               GWGS_EINVAL);
    Params p;
    p.nwsum = nwsum; p.nseg = nseg; p.num = num; p.flags = flags;
    p.exact_cut = exact_cut; p.thr = thr; p.nthr = nthr; p.exempt_snr = exempt_snr;
    p.sr = sr; p.wdiv = wdiv;
    hipStream_t s = (hipStream_t)stream;
    unsigned long long* bucket = (unsigned long long*)work;
    int* counts = ranked ? (int*)(bucket + total) : nullptr;
    int* off = ranked ? counts + nwin : nullptr;
    int* pos = ranked ? off + nwin + 1 : nullptr;
    const dim3 grid((unsigned)((total + NT - 1) / NT));
    if (ranked && nwin > 0) {
        hipLaunchKernelGGL(zero_k, dim3((nwin + NT - 1) / NT), dim3(NT), 0, s, counts, nwin);
        hipLaunchKernelGGL(prekeep_pass_k<0>, grid, dim3(NT), 0, s, vw, iw, total, cut, segi, segd, gates, floors, p,
                           counts, pos, bucket);
        hipLaunchKernelGGL(scan_k, dim3(1), dim3(1024), 0, s, counts, nwin, off, pos);
    }
    hipLaunchKernelGGL(prekeep_pass_k<1>, grid, dim3(NT), 0, s, vw, iw, total, cut, segi, segd, gates, floors, p,
                       counts, pos, bucket);
    if (ranked && nwin > 0)
        hipLaunchKernelGGL(prekeep_rank_k, dim3(nwin), dim3(NT), 0, s, vw, off, bucket, num);
    return gwgs_status();
}
