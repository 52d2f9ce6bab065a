// Time-slide coincidences over two detectors' trigger lists (the reference's coinc_rank.find_pairs,
// coinc_rank.py:100-138): all (H index, L index, slide k) with the same template and |tH - (tL + k step) mod P| <=
// window for |k| <= K. The L triggers arrive sorted by key = tid * big + t with copies at t - P and t + P (so the
// wrap is a plain range); one thread per H trigger binary-searches its key range, counts (pass 1) and writes
// (pass 2, after an exclusive scan of the counts) the pairs in (H index, L position) order.
#include "common.h"

namespace {

constexpr int NT = 256;

__device__ __forceinline__ long lower(const double* a, long n, const double x)
{
    long lo = 0, hi = n;
    while (lo < hi) { const long m = (lo + hi) >> 1; if (a[m] < x) lo = m + 1; else hi = m; }
    return lo;
}

__device__ __forceinline__ long upper(const double* a, long n, const double x)
{
    long lo = 0, hi = n;
    while (lo < hi) { const long m = (lo + hi) >> 1; if (a[m] <= x) lo = m + 1; else hi = m; }
    return lo;
}

template <bool WRITE>
__global__ void __launch_bounds__(NT) slide_pairs_k(const double* __restrict__ hkey, const double* __restrict__ hx,
                                                    const int nh, const double* __restrict__ lkey,
                                                    const double* __restrict__ lt, const int* __restrict__ li,
                                                    const long nl, const double reach, const double step,
                                                    const int K, const double window, long long* __restrict__ counts,
                                                    int* __restrict__ oh, int* __restrict__ ol, int* __restrict__ ok_,
                                                    double* __restrict__ odt)
{
    const int h = blockIdx.x * NT + threadIdx.x;
    if (h >= nh) return;
    const long a = lower(lkey, nl, hkey[h] - reach), b = upper(lkey, nl, hkey[h] + reach);
    long pos = WRITE ? counts[h] : 0;
    long c = 0;
    for (long j = a; j < b; j++) {
        const double dt = hx[h] - lt[j];                  // tH - tL (tL unwrapped)
        const double kk = rint(dt / step);
        if (fabs(dt - kk * step) <= window && fabs(kk) <= (double)K) {
            if (WRITE) {
                oh[pos] = h; ol[pos] = li[j]; ok_[pos] = (int)kk; odt[pos] = kk * step - dt;
                pos++;
            }
            c++;
        }
    }
    if (!WRITE) counts[h] = c;
}

__global__ void __launch_bounds__(1024) exscan_k(long long* __restrict__ c, const int n, long long* __restrict__ total)
{
    __shared__ long long part[1024];
    const int t = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int a = min(t * per, n), b = min(a + per, n);
    long long s = 0;
    for (int i = a; i < b; i++) s += c[i];
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const long long v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    long long run = t ? part[t - 1] : 0;
    for (int i = a; i < b; i++) { const long long x = c[i]; c[i] = run; run += x; }
    if (t == 1023) *total = part[1023];
}

} // namespace

// Pass 1: counts[h] (nh + 1 int64 workspace; counts become offsets) and *total (device int64) of pairs.
GWGS_API int gwgs_slide_count(const double* hkey, const double* hx, int nh, const double* lkey, const double* lt,
                              const int* li, long nl, double reach, double step, int K, double window,
                              long long* counts, long long* total, void* stream)
{
    GWGS_CHECK(hkey && hx && lkey && lt && li && counts && total && nh >= 0 && nl >= 0 && step > 0, GWGS_EINVAL);
    hipStream_t s = (hipStream_t)stream;
    if (nh > 0)
        hipLaunchKernelGGL(slide_pairs_k<false>, dim3((nh + NT - 1) / NT), dim3(NT), 0, s, hkey, hx, nh, lkey, lt, li,
                           nl, reach, step, K, window, counts, nullptr, nullptr, nullptr, nullptr);
    hipLaunchKernelGGL(exscan_k, dim3(1), dim3(1024), 0, s, counts, nh, total);
    return gwgs_status();
}

// Pass 2: the pairs at the offsets from pass 1: H index, L index (into the unsorted L list), slide k, and
// dt = k step - (tH - tL).
GWGS_API int gwgs_slide_write(const double* hkey, const double* hx, int nh, const double* lkey, const double* lt,
                              const int* li, long nl, double reach, double step, int K, double window,
                              const long long* offsets, int* oh, int* ol, int* ok, double* odt, void* stream)
{
    GWGS_CHECK(hkey && hx && lkey && lt && li && offsets && oh && ol && ok && odt && nh >= 0, GWGS_EINVAL);
    if (nh == 0) return GWGS_OK;
    hipLaunchKernelGGL(slide_pairs_k<true>, dim3((nh + NT - 1) / NT), dim3(NT), 0, (hipStream_t)stream, hkey, hx, nh,
                       lkey, lt, li, nl, reach, step, K, window, (long long*)offsets, oh, ol, ok, odt);
    return gwgs_status();
}
