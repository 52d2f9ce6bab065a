// Peak finding and candidate selection (the reference's peak_kernel.py:8-47 and fp_filter.py:745-764), wave64.
#include "common.h"

#include <mutex>

namespace {

constexpr int NT = 256;

// One block per (window, row): max |z|^2 over z[row, a0 + w cw : a0 + (w+1) cw] and its index in the window.
// Each thread scans a strided subset in increasing index (strict >, first index), then the 64 lanes of each wave and
// the 4 waves merge with the first-index rule (peak_kernel.py:15-37).
__global__ void __launch_bounds__(NT) win_peak_k(const float2* __restrict__ z, const long ldz, const int a0,
                                                 const int cw, const int nwin, float* __restrict__ pmax,
                                                 int* __restrict__ pidx)
{
#pragma clang fp contract(off)
    const int w = blockIdx.x, row = blockIdx.y;
    const float2* zr = z + (size_t)row * ldz + a0 + (size_t)w * cw;
    float best = -1.f;
    int bi = 0x7fffffff;
    for (int i = threadIdx.x; i < cw; i += NT) {
        const float2 v = zr[i];
        const float p = __fadd_rn(__fmul_rn(v.x, v.x), __fmul_rn(v.y, v.y));   // no contraction: numpy's float32 |z|^2
        if (p > best) { best = p; bi = i; }
    }
    wave_argmax_first(best, bi);
    __shared__ float sb[NT / 64];
    __shared__ int si[NT / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) { sb[wave] = best; si[wave] = bi; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < NT / 64; k++) take_max_first(best, bi, sb[k], si[k]);
        if (bi == 0x7fffffff) bi = 0;
        pmax[(size_t)row * nwin + w] = best;
        pidx[(size_t)row * nwin + w] = bi;
    }
}

// Window max over the GEMM's per-block maxima: vw[r, w] = max_j V[r, w per + j], iw = I at that j, first j on ties
// (fp_filter.py:745-746 with the tie rule of peak_kernel.py:21-33). One thread per (row, window).
__global__ void block_window_max_k(const float* __restrict__ V, const int* __restrict__ I, const long ldv, const int n,
                                   const int nw, const int per, float* __restrict__ vw, int* __restrict__ iw)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)n * nw) return;
    const int r = (int)(g / nw), w = (int)(g % nw);
    const float* v = V + (size_t)r * ldv + (size_t)w * per;
    const int* ii = I + (size_t)r * ldv + (size_t)w * per;
    float b = v[0];
    int j = 0;
    for (int k = 1; k < per; k++)
        if (v[k] > b) { b = v[k]; j = k; }
    vw[g] = b;
    iw[g] = ii[j];
}

// Ordered compaction of {g : vw[g] > cut} over a row-major (n, nw) array, the order of numpy/torch nonzero.
// Pass 1: per chunk of CH elements the count. Pass 2 (one block): exclusive scan of the chunk counts. Pass 3: each
// chunk writes its hits at its offset, in order, through per-wave ballots.
constexpr int CH = 4096;

__global__ void __launch_bounds__(NT) count_above_k(const float* __restrict__ vw, const long total, const float cut,
                                                    int* __restrict__ counts)
{
    const long c0 = (long)blockIdx.x * CH;
    int cnt = 0;
    for (int i = threadIdx.x; i < CH; i += NT) {
        const long g = c0 + i;
        if (g < total && vw[g] > cut) cnt++;
    }
    __shared__ int s[NT / 64];
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = s[0] + s[1] + s[2] + s[3];
}

__global__ void __launch_bounds__(1024) scan_counts_k(int* __restrict__ counts, const int nc, int* __restrict__ total)
{
    // exclusive scan in place, one block: each thread owns a contiguous run
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (nc + 1023) / 1024;
    const int a = min(t * per, nc), b = min(a + per, nc);
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
    for (int i = a; i < b; i++) { const int c = counts[i]; counts[i] = run; run += c; }
    if (t == 1023) *total = part[1023];
}

// out[pos] = g for each vw[g] > cut in increasing g; with iw, also out_i[pos] = iw[g] and out_v[pos] = vw[g]
__global__ void __launch_bounds__(NT) write_above_k(const float* __restrict__ vw, const long total, const float cut,
                                                    const int* __restrict__ offs, int* __restrict__ out, const int cap,
                                                    const int* __restrict__ iw, int* __restrict__ out_i,
                                                    float* __restrict__ out_v)
{
    const long c0 = (long)blockIdx.x * CH;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __shared__ int wcount[NT / 64];
    int base = offs[blockIdx.x];
    for (int i0 = 0; i0 < CH; i0 += NT) {
        const long g = c0 + i0 + threadIdx.x;
        const bool hit = g < total && vw[g] > cut;
        const unsigned long long m = __ballot(hit);
        if (lane == 0) wcount[wave] = __popcll(m);
        __syncthreads();
        int before = 0;
        for (int w = 0; w < wave; w++) before += wcount[w];
        const int pos = base + before + __popcll(m & ((1ull << lane) - 1ull));
        if (hit && pos < cap) {
            out[pos] = (int)g;
            if (iw) { out_i[pos] = iw[g]; out_v[pos] = vw[g]; }
        }
        base += wcount[0] + wcount[1] + wcount[2] + wcount[3];
        __syncthreads();
    }
}

// Rebuilt complex SNR numerator at candidates (fp_filter.py:759-764): zr = Ar[tb] . Zt[t], zi = Ai[tb] . Zt[t] in
// fp32 from the GEMM's own operands, times sqrt(pw_scale[sidx]) (the GEMM's row scale of that template and segment).
// One wave per candidate.
template <typename T>
__device__ __forceinline__ float to_f32(const T x);
template <>
__device__ __forceinline__ float to_f32<float>(const float x) { return x; }
template <>
__device__ __forceinline__ float to_f32<uint16_t>(const uint16_t x) { return __uint_as_float((uint32_t)x << 16); }
template <>
__device__ __forceinline__ float to_f32<uint8_t>(const uint8_t x) { return fp8_to_f32(x); }
template <>
__device__ __forceinline__ float to_f32<f16s>(const f16s x) { return f16_to_f32(x.v); }

template <typename T>
__global__ void __launch_bounds__(NT) rebuilt_z_k(const T* __restrict__ Zt, const long ldz, const T* __restrict__ Ar,
                                                  const T* __restrict__ Ai, const long lda, const int K,
                                                  const int* __restrict__ rows, const int* __restrict__ tcol,
                                                  const int* __restrict__ sidx, const int n,
                                                  const float* __restrict__ pw_scale, float2* __restrict__ out)
{
    const int c = blockIdx.x * (NT / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (c >= n) return;
    const T* a = Ar + (size_t)rows[c] * lda;
    const T* b = Ai + (size_t)rows[c] * lda;
    const T* zz = Zt + (size_t)tcol[c] * ldz;
    float sr = 0.f, si = 0.f;
    for (int k = lane; k < K; k += 64) {
        const float zv = to_f32(zz[k]);
        sr += to_f32(a[k]) * zv;
        si += to_f32(b[k]) * zv;
    }
    sr = wave_sum_f32(sr);
    si = wave_sum_f32(si);
    const float sc = pw_scale ? sqrtf(pw_scale[sidx ? sidx[c] : rows[c]]) : 1.f;
    if (lane == 0) out[c] = make_float2(sr * sc, si * sc);
}

// The alternates of every exact-band candidate of a compacted list, from its window's local-maximum list (LM
// entries per window as gwgs_gemm_peaks_multi_lm writes them, larger |z|^2 first, then the earlier sample): the
// entries farther than R samples from the candidate's window maximum samp[c] whose |z|^2 is at least ratio val[c],
// in list order, up to na of them; out[c na + a] = the entry's offset in its window (sample mod cw), -1 past the
// last. Candidates with val[c] < exact_cut get none. One thread per candidate.
__global__ void __launch_bounds__(NT) window_alternates_k(const float* __restrict__ VL, const int* __restrict__ IL,
                                                          const long ldl, const int lm, const int* __restrict__ idx,
                                                          const int* __restrict__ samp,
                                                          const float* __restrict__ val,
                                                          const int* __restrict__ count, const int cap,
                                                          const int nwsum, const int cw, const float exact_cut,
                                                          const float ratio, const int R, const int na,
                                                          int* __restrict__ out)
{
    const int n = min(*count, cap);
    for (int c = blockIdx.x * NT + threadIdx.x; c < n; c += gridDim.x * NT) {
        int a = 0;
        const float v0 = val[c];
        if (v0 >= exact_cut) {
            const long g = idx[c];
            const size_t base = (size_t)(g / nwsum) * ldl + (size_t)(g % nwsum) * lm;
            const int t0 = samp[c];
            for (int k = 0; k < lm && a < na; k++) {
                const float v = VL[base + k];
                const int t = IL[base + k];
                if (t >= 0 && v >= ratio * v0 && abs(t - t0) > R) out[(size_t)c * na + a++] = t % cw;
            }
        }
        for (; a < na; a++) out[(size_t)c * na + a] = -1;
    }
}

} // namespace

GWGS_API int gwgs_window_peaks(const void* z, long ldz, int rows, int a0, int cw, int nwin, float* pmax, int* pidx,
                               void* stream)
{
    GWGS_CHECK(z && pmax && pidx && rows >= 0 && a0 >= 0 && cw > 0 && nwin >= 0, GWGS_EINVAL);
    GWGS_CHECK(ldz >= (long)a0 + (long)nwin * cw, GWGS_EINVAL);
    if (rows == 0 || nwin == 0) return GWGS_OK;
    GWGS_CHECK(rows < 65536, GWGS_ESIZE);
    hipLaunchKernelGGL(win_peak_k, dim3(nwin, rows), dim3(NT), 0, (hipStream_t)stream, (const float2*)z, ldz, a0, cw,
                       nwin, pmax, pidx);
    return gwgs_status();
}

GWGS_API int gwgs_block_window_max(const float* V, const int* I, long ldv, int n, int nw, int per, float* vw, int* iw,
                                   void* stream)
{
    GWGS_CHECK(V && I && vw && iw && n >= 0 && nw >= 0 && per > 0 && ldv >= (long)nw * per, GWGS_EINVAL);
    const long tot = (long)n * nw;
    if (tot == 0) return GWGS_OK;
    hipLaunchKernelGGL(block_window_max_k, dim3((unsigned)((tot + NT - 1) / NT)), dim3(NT), 0, (hipStream_t)stream,
                       V, I, ldv, n, nw, per, vw, iw);
    return gwgs_status();
}

// Workspace for gwgs_select_above: ceil(total / 4096) + 1 ints.
GWGS_API long gwgs_select_workspace(long total) { return (total + CH - 1) / CH + 1; }

// Row-major indices g of vw[g] > cut (strict), in increasing g, into out[0:cap]; the count goes to *count_dev (device).
GWGS_API int gwgs_select_above(const float* vw, long total, float cut, int* out, int cap, int* work, int* count_dev,
                               void* stream)
{
    GWGS_CHECK((vw || total == 0) && (out || cap == 0) && work && count_dev && total >= 0 && cap >= 0,
               GWGS_EINVAL);
    hipStream_t s = (hipStream_t)stream;
    const long nc = (total + CH - 1) / CH;
    GWGS_CHECK(nc < (1L << 24), GWGS_ESIZE);
    if (nc) hipLaunchKernelGGL(count_above_k, dim3((unsigned)nc), dim3(NT), 0, s, vw, total, cut, work);
    hipLaunchKernelGGL(scan_counts_k, dim3(1), dim3(1024), 0, s, work, (int)nc, count_dev);   // nc = 0: count 0
    if (nc)
        hipLaunchKernelGGL(write_above_k, dim3((unsigned)nc), dim3(NT), 0, s, vw, total, cut, work, out, cap, nullptr,
                           nullptr, nullptr);
    return gwgs_status();
}

// As gwgs_select_above, and for each selected g also its window sample iw[g] into out_iw and its value vw[g] into
// out_vw (the candidate list of fp_filter.py:747-757 in one pass, no host round trip).
GWGS_API int gwgs_select_peaks(const float* vw, const int* iw, long total, float cut, int* out, int* out_iw,
                               float* out_vw, int cap, int* work, int* count_dev, void* stream)
{
    GWGS_CHECK((vw || total == 0) && (iw || total == 0) && ((out && out_iw && out_vw) || cap == 0) && work &&
               count_dev && total >= 0 && cap >= 0, GWGS_EINVAL);
    hipStream_t s = (hipStream_t)stream;
    const long nc = (total + CH - 1) / CH;
    GWGS_CHECK(nc < (1L << 24), GWGS_ESIZE);
    if (nc) hipLaunchKernelGGL(count_above_k, dim3((unsigned)nc), dim3(NT), 0, s, vw, total, cut, work);
    hipLaunchKernelGGL(scan_counts_k, dim3(1), dim3(1024), 0, s, work, (int)nc, count_dev);
    if (nc)
        hipLaunchKernelGGL(write_above_k, dim3((unsigned)nc), dim3(NT), 0, s, vw, total, cut, work, out, cap, iw,
                           out_iw, out_vw);
    return gwgs_status();
}

// The candidate compaction without a caller-provided workspace, for callers that cannot allocate (code under
// graph capture, serving loops). Its per-block counts live in a file-scope device array, one copy per GPU that the
// library runs on, held until the process ends; the host takes each GPU's address of it once (hipGetSymbolAddress
// with that GPU current) and passes that address to the kernels, never the array's name. Every block count is
// written before it is read, so nothing is cleared between calls; calls that overlap on one GPU share the array
// and race.
constexpr long NOALLOC_COUNTS = 1L << 16;          // 256 KiB: total <= NOALLOC_COUNTS * CH entries (2.7e8)
static __device__ __align__(256) int scratch[NOALLOC_COUNTS];
constexpr int NOALLOC_MAX_DEVICES = 64;
static int* scratch_dev[NOALLOC_MAX_DEVICES];
static std::mutex scratch_mu;

static int scratch_for_current_device(int** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    GWGS_CHECK(dev >= 0 && dev < NOALLOC_MAX_DEVICES, GWGS_ESIZE);
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_dev[dev]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        scratch_dev[dev] = (int*)ptr;
    }
    *out = scratch_dev[dev];
    return GWGS_OK;
}

GWGS_API int gwgs_select_peaks_noalloc(const float* vw, const int* iw, long total, float cut, int* out, int* out_iw,
                                       float* out_vw, int cap, int* count_dev, void* stream)
{
    GWGS_CHECK((vw || total == 0) && (iw || total == 0) && ((out && out_iw && out_vw) || cap == 0) && count_dev &&
               total >= 0 && cap >= 0, GWGS_EINVAL);
    const long nc = (total + CH - 1) / CH;
    GWGS_CHECK(nc <= NOALLOC_COUNTS, GWGS_ESIZE);
    int* work = nullptr;
    const int rc = scratch_for_current_device(&work);
    if (rc != GWGS_OK) return rc;
    hipStream_t s = (hipStream_t)stream;
    if (nc) hipLaunchKernelGGL(count_above_k, dim3((unsigned)nc), dim3(NT), 0, s, vw, total, cut, work);
    hipLaunchKernelGGL(scan_counts_k, dim3(1), dim3(1024), 0, s, work, (int)nc, count_dev);
    if (nc)
        hipLaunchKernelGGL(write_above_k, dim3((unsigned)nc), dim3(NT), 0, s, vw, total, cut, work, out, cap, iw,
                           out_iw, out_vw);
    return gwgs_status();
}

GWGS_API int gwgs_rebuilt_z(int dtype, const void* Zt, long ldz, const void* Ar, const void* Ai, long lda, int K,
                            const int* rows, const int* tcol, const int* sidx, int n, const float* pw_scale, void* out,
                            void* stream)
{
    GWGS_CHECK(Zt && Ar && Ai && rows && tcol && out && K > 0 && n >= 0, GWGS_EINVAL);
    if (n == 0) return GWGS_OK;
    const dim3 grid((n + NT / 64 - 1) / (NT / 64));
    hipStream_t s = (hipStream_t)stream;
    if (dtype == 3)
        hipLaunchKernelGGL(rebuilt_z_k<f16s>, grid, dim3(NT), 0, s, (const f16s*)Zt, ldz, (const f16s*)Ar,
                           (const f16s*)Ai, lda, K, rows, tcol, sidx, n, pw_scale, (float2*)out);
    else if (dtype == 2)
        hipLaunchKernelGGL(rebuilt_z_k<uint8_t>, grid, dim3(NT), 0, s, (const uint8_t*)Zt, ldz, (const uint8_t*)Ar,
                           (const uint8_t*)Ai, lda, K, rows, tcol, sidx, n, pw_scale, (float2*)out);
    else if (dtype == 1)
        hipLaunchKernelGGL(rebuilt_z_k<uint16_t>, grid, dim3(NT), 0, s, (const uint16_t*)Zt, ldz, (const uint16_t*)Ar,
                           (const uint16_t*)Ai, lda, K, rows, tcol, sidx, n, pw_scale, (float2*)out);
    else if (dtype == 0)
        hipLaunchKernelGGL(rebuilt_z_k<float>, grid, dim3(NT), 0, s, (const float*)Zt, ldz, (const float*)Ar,
                           (const float*)Ai, lda, K, rows, tcol, sidx, n, pw_scale, (float2*)out);
    else
        return GWGS_EUNSUP;
    return gwgs_status();
}

// The alternates of a compacted candidate list (see window_alternates_k); cap: the list's capacity (entries past
// it are not read), out (cap, na) int32.
GWGS_API int gwgs_window_alternates(const float* VL, const int* IL, long ldl, int lm, const int* idx, const int* samp,
                                    const float* val, const int* count, int cap, int nwsum, int cw, float exact_cut,
                                    float ratio, int R, int na, int* out, void* stream)
{
    GWGS_CHECK(VL && IL && idx && samp && val && count && out && lm > 0 && cap >= 0 && nwsum > 0 && cw > 0 &&
               ldl >= (long)lm * nwsum && R >= 0 && na > 0, GWGS_EINVAL);
    if (cap == 0) return GWGS_OK;
    const int nb = (int)std::min<long>(((long)cap + NT - 1) / NT, 1024);
    hipLaunchKernelGGL(window_alternates_k, dim3(nb), dim3(NT), 0, (hipStream_t)stream, VL, IL, ldl, lm, idx, samp,
                       val, count, cap, nwsum, cw, exact_cut, ratio, R, na, out);
    return gwgs_status();
}
