// Fused SNR-reconstruction GEMM with |z|^2 and max/argmax epilogue (the reference's fp_gemm.gemm_absmax,
// fp_gemm.py:14-44), written for CDNA4 MFMA.
//
// Formulation. For one template tile and one data segment,
//     z_re[m, t] = sum_k Ar[m, k] Zt[t, k],   z_im[m, t] = sum_k Ai[m, k] Zt[t, k],   pw = z_re^2 + z_im^2,
// with Zt the basis series stored time-major (row t holds the K real basis values, interleaved Re/Im per complex
// basis vector). The MFMA computes C = Zt_tile (time x k) * A_tile^T (k x templates), so each lane owns one
// template column and 16 time rows of every 32 x 32 output tile: the max over time is an in-register scan in
// increasing t (strict >, so the first index wins), and the two half-waves that share a column are merged with one
// lane swap. Nothing of the M x Tn SNR array is written.
//
// Work decomposition. A launch covers one or more segments whose Zt rows lie in one buffer. It is cut into items:
// a run of 128-sample tiles of one segment. Workgroup = 4 waves = 128 templates x one item; the grid is
// (template blocks) x (items), ordered so that the template blocks of one item are dispatched to one XCD and share
// its L2 copy of the Zt tiles. Output modes:
//   mode 0: V, I per 128-sample block: the max |z|^2 and its sample (samples >= the segment's Tn masked to -1),
//           exactly fused_peaks' contract;
//   mode 1: per cluster window of `per` blocks the max and its sample (fp_filter.py:745-746 fused in), first
//           index on ties at both levels.
// Samples are relative to the segment's first Zt row; segment s writes columns col0(s) + its block / window index.
// An optional row scale rowscale[seg M + m] multiplies the reported |z|^2 (template norms per segment, fp8 scales).
//
// Kernels: register-resident template operands for K <= 256 (bf16) / 128 (fp32), Zt tiles double buffered through
// LDS; and a K-loop kernel (both operands staged through LDS in K chunks) for any K and for fp8 e4m3 operands.
#include "common.h"
#include <algorithm>

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

// The 32x32x16 MFMA of a 2-byte operand type: bf16 for uint16_t, fp16 for f16s; operands are passed as the 8
// elements' bits (a bf16x8 value) in either case.
template <typename T>
__device__ __forceinline__ f32x16 mfma16(const bf16x8 a, const bf16x8 b, const f32x16 c, int, int, int)
{
    if constexpr (__is_same(T, f16s))
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c, 0,
                                                      0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
typedef int i32x8 __attribute__((ext_vector_type(8)));

// One item of a launch: tiles [t_first, t_end) of one segment (tile = 128 samples).
struct GemmItem {
    long long zrow0;   // the segment's first row in Zt
    int t_first;
    int t_end;
    int col0;          // the segment's first output column
    int seg;           // row-scale set
    int Tn;            // the segment's samples (mode 0 masks t >= Tn)
    int pad;
};

namespace {

constexpr int BT = 128;          // time samples per tile (the reference's BN)
constexpr int WAVES = 4;
constexpr int NT = 64 * WAVES;
constexpr int BMT = 32 * WAVES;  // templates per block

__device__ __forceinline__ i32x8 join8(const uint4 a, const uint4 b)
{
    i32x8 r;
    r[0] = (int)a.x; r[1] = (int)a.y; r[2] = (int)a.z; r[3] = (int)a.w;
    r[4] = (int)b.x; r[5] = (int)b.y; r[6] = (int)b.z; r[7] = (int)b.w;
    return r;
}

// The imaginary part without a second template operand: with Ar[m, 2j] = Re As, Ar[m, 2j+1] = -Im As,
//     z_im[m, t] = sum_j Im As Re z_j + Re As Im z_j = sum_k Ar[m, k] Zr[t, k],  Zr[t, 2j] = Im z_j, Zr[t, 2j+1] = -Re z_j,
// so each (Re, Im) pair of a Zt fragment is swapped and its new second element negated in registers: Ai is never
// loaded, halving the template traffic of the two real GEMMs.
__device__ __forceinline__ uint32_t rot_bf16x2(const uint32_t w)
{
    return __builtin_amdgcn_alignbit(w, w, 16) ^ 0x80000000u;     // (re, im) -> (im, -re)
}
__device__ __forceinline__ uint4 rot_bf16(const uint4 v)
{
    return make_uint4(rot_bf16x2(v.x), rot_bf16x2(v.y), rot_bf16x2(v.z), rot_bf16x2(v.w));
}
__device__ __forceinline__ uint32_t rot_fp8x4(const uint32_t w)
{
    return (((w >> 8) & 0x00ff00ffu) | ((w << 8) & 0xff00ff00u)) ^ 0x80008000u;   // bytes (re, im) -> (im, -re)
}
__device__ __forceinline__ uint4 rot_fp8(const uint4 v)
{
    return make_uint4(rot_fp8x4(v.x), rot_fp8x4(v.y), rot_fp8x4(v.z), rot_fp8x4(v.w));
}

// The same identity on the template side: Ai[m, 2j] = -Ar[m, 2j+1], Ai[m, 2j+1] = Ar[m, 2j], i.e. each (lo, hi) pair
// of an Ar fragment becomes (-hi, lo). A wave rotates its own template fragment once per k step and uses it for all
// of its time sub-tiles, a quarter of the rotations of the Zt-side form.
__device__ __forceinline__ uint32_t arot_bf16x2(const uint32_t w)
{
    return __builtin_amdgcn_alignbit(w, w, 16) ^ 0x00008000u;     // (lo, hi) -> (-hi, lo)
}
__device__ __forceinline__ uint4 arot_bf16(const uint4 v)
{
    return make_uint4(arot_bf16x2(v.x), arot_bf16x2(v.y), arot_bf16x2(v.z), arot_bf16x2(v.w));
}
__device__ __forceinline__ uint32_t arot_fp8x4(const uint32_t w)
{
    // v_perm_b32 byte select (1, 0, 3, 2): swap the bytes of each 16-bit pair, then negate the new low byte
    return __builtin_amdgcn_perm(w, w, 0x02030001u) ^ 0x00800080u;
}
__device__ __forceinline__ uint4 arot_fp8(const uint4 v)
{
    return make_uint4(arot_fp8x4(v.x), arot_fp8x4(v.y), arot_fp8x4(v.z), arot_fp8x4(v.w));
}

struct Job {
    const GemmItem* items;   // null: one segment at row 0, items derived from tpb, ntile, Tn
    const int* nitems_dev;   // non-null: the item count in device memory (at most nitems, the grid's size)
    int nitems, nmb, tpb, ntile, Tn, M, per;
    const float* rowscale;
    float* vb;               // mode 2: each window's LM largest separated local maxima of |z|^2 (row scaled, -1
    int* ib;                 // for none) at vb[m ldb + col LM + k], their samples (-1 for none) in ib; maxima
    long ldb;                // closer than sep + 1 samples count once; those below gate x the window's maximum
    int sep;                 // need not be listed, nor those below gate x floor (row scaled)
    float gate, floor;
};

// logical block -> (template block, item); false for the grid's padding blocks
__device__ __forceinline__ bool block_item(const Job& J, int& mb, GemmItem& it)
{
    // the hardware places workgroup b on XCD b % 8: consecutive logical blocks (the template blocks of one item)
    // land on one XCD. With the item count in device memory the grid is sized for the maximum, and the mapping
    // spreads the blocks of the actual count over all eight XCDs (the rest return).
    const int b = blockIdx.x;
    int n8 = gridDim.x >> 3, tot = J.nitems * J.nmb;
    if (J.nitems_dev) {
        tot = min(*J.nitems_dev, J.nitems) * J.nmb;
        n8 = (tot + 7) >> 3;
        if ((b >> 3) >= n8) return false;
    }
    const int L = (b & 7) * n8 + (b >> 3);
    if (L >= tot) return false;
    mb = L % J.nmb;
    const int ix = L / J.nmb;
    if (J.items) {
        it = J.items[ix];
    } else {
        it.zrow0 = 0; it.t_first = ix * J.tpb; it.t_end = min(it.t_first + J.tpb, J.ntile);
        it.col0 = 0; it.seg = 0; it.Tn = J.Tn; it.pad = 0;
    }
    return true;
}

__device__ __forceinline__ void scan_tile(const f32x16 (&cr)[4], const f32x16 (&ci)[4], const int tb0, const int half,
                                          const int Tn, float& best, int& bi)
{
#pragma unroll
    for (int ts = 0; ts < 4; ts++) {
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int t = tb0 + ts * 32 + 8 * (r >> 2) + 4 * half + (r & 3);
            const float re = cr[ts][r], im = ci[ts][r];
            const float p = __builtin_fmaf(re, re, im * im);     // one rounding order in every epilogue
            if (t < Tn && p > best) { best = p; bi = t; }
        }
    }
}

__device__ __forceinline__ void emit(float best, int bi, const int half, const int m, const int M, const float sc,
                                     float* __restrict__ V, int* __restrict__ I, const long ldv, const int col_out)
{
    const float ob = __shfl_xor(best, 32, 64);
    const int oi = __shfl_xor(bi, 32, 64);
    take_max_first(best, bi, ob, oi);
    if (half == 0 && m < M) {
        V[(size_t)m * ldv + col_out] = best * sc;
        I[(size_t)m * ldv + col_out] = bi;
    }
}

// after tile tt (segment-relative) of item `it`: emit per block (mode 0) or at each window end (mode 1)
template <int MODE>
__device__ __forceinline__ void tile_done(const Job& J, const GemmItem& it, const int tt, const int half, const int m,
                                          float& best, int& bi, float* __restrict__ V, int* __restrict__ I,
                                          const long ldv)
{
    const bool out = MODE == 0 || (tt + 1) % J.per == 0;   // modes 1 and 2: at each window end
    if (!out) return;
    const float sc = (J.rowscale && m < J.M) ? J.rowscale[(size_t)it.seg * J.M + m] : 1.f;
    emit(best, bi, half, m, J.M, sc, V, I, ldv, it.col0 + (MODE == 0 ? tt : (tt + 1) / J.per - 1));
    best = -1.f;
    bi = (tt + 1) * BT;
}

// Mode 2 also keeps, per (template, window), the LM largest separated local maxima of |z|^2 (larger value first,
// then smaller sample): sample t of the window is a local maximum when p[t] >= p[t - 1] and p[t] > p[t + 1], the
// window's edges counting as open, and of local maxima closer than sep + 1 samples only the larger is kept (the
// earlier on ties), so that the ripples rounding leaves on a broad peak's top do not crowd the list. A lane holds
// the samples tb0 + 32 ts + 8 q + 4 half + j (run u = 4 ts + q, j = 0..3) of its template, so a run's neighbours
// lie in the other half's lane: half 0's run u sits between half 1's runs u - 1 and u, half 1's run u between half
// 0's runs u and u + 1. Half 0 carries half 1's last sample of the previous run as its left neighbour; half 1 holds
// its run's last sample until half 0's first sample of the next run (the next tile's, across a tile edge) arrives.
// A lane meets its samples in increasing order, so only its last accepted local maximum can lie within sep of a
// new one; the two halves' lists merge at the window's end with the full separation test.
constexpr int LM = 4;

struct LocalMax {
    float v[LM];
    int i[LM];
    float carry;   // half 0: the left neighbour of its next run (-1 at the window's left edge)
    float pv;      // half 1: its last sample, waiting for its right neighbour
    int pt;
    bool pok;      // ... that sample is >= its left neighbour
    float lv;      // the lane's last accepted local maximum
    int lt;
};

__device__ __forceinline__ void lm_reset(LocalMax& L)
{
#pragma unroll
    for (int k = 0; k < LM; k++) { L.v[k] = -1.f; L.i[k] = 0x7fffffff; }
    L.carry = -1.f; L.pv = -1.f; L.pt = 0; L.pok = false; L.lv = -1.f; L.lt = -(1 << 30);
}

__device__ __forceinline__ bool beats(const float x, const int t, const float v, const int i)
{
    return x > v || (x == v && t < i);
}

// (x, t) into the sorted list, the smallest entry dropping out
__device__ __forceinline__ void lm_insert(LocalMax& L, float x, int t)
{
#pragma unroll
    for (int k = 0; k < LM; k++) {
        const bool take = beats(x, t, L.v[k], L.i[k]);
        const float ov = L.v[k];
        const int oi = L.i[k];
        if (take) { L.v[k] = x; L.i[k] = t; x = ov; t = oi; }
    }
}

// a local maximum (x, t) met in increasing t: within sep of the last accepted one it replaces it if larger (in the
// list, moving up; or entering it), else it is dropped; farther, it is a new one
__device__ __forceinline__ void lm_try(LocalMax& L, const bool lm, const float x, const int t, const int sep)
{
    if (!lm) return;
    if (t - L.lt <= sep) {
        if (x > L.lv) {
            bool found = false;
#pragma unroll
            for (int k = 0; k < LM; k++)
                if (L.i[k] == L.lt) { L.v[k] = x; L.i[k] = t; found = true; }
            if (found) {
#pragma unroll
                for (int k = LM - 1; k > 0; k--) {
                    if (beats(L.v[k], L.i[k], L.v[k - 1], L.i[k - 1])) {
                        const float tv = L.v[k]; L.v[k] = L.v[k - 1]; L.v[k - 1] = tv;
                        const int ti = L.i[k]; L.i[k] = L.i[k - 1]; L.i[k - 1] = ti;
                    }
                }
            } else if (beats(x, t, L.v[LM - 1], L.i[LM - 1])) {
                lm_insert(L, x, t);
            }
            L.lv = x; L.lt = t;
        }
        return;
    }
    L.lv = x; L.lt = t;
    if (beats(x, t, L.v[LM - 1], L.i[LM - 1])) lm_insert(L, x, t);
}

__device__ __forceinline__ void cswap(LocalMax& L, const int a, const int b)
{
    if (beats(L.v[b], L.i[b], L.v[a], L.i[a])) {
        const float tv = L.v[a]; L.v[a] = L.v[b]; L.v[b] = tv;
        const int ti = L.i[a]; L.i[a] = L.i[b]; L.i[b] = ti;
    }
}

// (x, t) from the other half into the list, in any order: dropped if an entry within sep beats it, else it
// removes the entries within sep (which it beats) and enters
__device__ __forceinline__ void lm_merge_one(LocalMax& L, const float x, const int t, const int sep)
{
    bool dom = false;
#pragma unroll
    for (int k = 0; k < LM; k++) dom |= abs(L.i[k] - t) <= sep && !beats(x, t, L.v[k], L.i[k]);
    if (dom) return;
#pragma unroll
    for (int k = 0; k < LM; k++)
        if (abs(L.i[k] - t) <= sep) { L.v[k] = -1.f; L.i[k] = 0x7fffffff; }
    static_assert(LM == 4, "the sorting network below is for four entries");
    cswap(L, 0, 1); cswap(L, 2, 3); cswap(L, 0, 2); cswap(L, 1, 3); cswap(L, 1, 2);
    if (beats(x, t, L.v[LM - 1], L.i[LM - 1])) lm_insert(L, x, t);
}

// The tile's local maxima at or above thr into the lists; p: |z|^2 in place of cr (-1 past the segment's end, which
// also makes the last valid sample's right edge open). thr is uniform over a template's two lanes, so a run in
// which no lane of the wave holds a sample at or above its thr is skipped whole: its samples are all below thr, so
// the next run's first sample, if at or above thr, is >= its left neighbour (carry -1 says so), and the held sample
// cannot qualify.
__device__ __forceinline__ void lm_tile(LocalMax& L, const f32x16 (&p)[4], const int tb0, const int half,
                                        const int sep, const float thr)
{
#pragma unroll
    for (int u = 0; u < 16; u++) {
        const int ts = u >> 2, q = u & 3;
        const float x0 = p[ts][4 * q], x1 = p[ts][4 * q + 1], x2 = p[ts][4 * q + 2], x3 = p[ts][4 * q + 3];
        const int t0 = tb0 + 32 * ts + 8 * q + 4 * half;
        const bool hot = fmaxf(fmaxf(x0, x1), fmaxf(x2, x3)) >= thr || (half && L.pok && L.pv >= thr);
        if (__ballot(hot) == 0) {
            L.carry = -1.f;
            L.pok = false;
            continue;
        }
        const float r0 = __shfl_xor(x0, 32, 64), r3 = __shfl_xor(x3, 32, 64);
        const float prev = half ? r3 : L.carry;
        lm_try(L, half && L.pok && L.pv > r0 && L.pv >= thr, L.pv, L.pt, sep);
        lm_try(L, x0 >= thr && x0 >= prev && x0 > x1, x0, t0, sep);
        lm_try(L, x1 >= thr && x1 >= x0 && x1 > x2, x1, t0 + 1, sep);
        lm_try(L, x2 >= thr && x2 >= x1 && x2 > x3, x2, t0 + 2, sep);
        lm_try(L, !half && x3 >= thr && x3 >= x2 && x3 > r0, x3, t0 + 3, sep);
        L.carry = r3;
        L.pv = x3; L.pt = t0 + 3; L.pok = half && x3 >= x2 && x3 >= 0.f;
    }
}

// window end: half 1's held sample has an open right edge; the halves' lists merged; half 0 writes them row scaled
__device__ __forceinline__ void lm_emit(LocalMax& L, const int half, const int m, const int M, const float sc,
                                        const int sep, float* __restrict__ VL, int* __restrict__ IL, const long ldl,
                                        const int col)
{
    lm_try(L, half && L.pok, L.pv, L.pt, sep);
    float ov[LM];
    int oi[LM];
#pragma unroll
    for (int k = 0; k < LM; k++) { ov[k] = __shfl_xor(L.v[k], 32, 64); oi[k] = __shfl_xor(L.i[k], 32, 64); }
#pragma unroll
    for (int k = 0; k < LM; k++)
        if (ov[k] >= 0.f) lm_merge_one(L, ov[k], oi[k], sep);
    if (half == 0 && m < M) {
#pragma unroll
        for (int k = 0; k < LM; k++) {
            VL[(size_t)m * ldl + (size_t)col * LM + k] = L.v[k] >= 0.f ? L.v[k] * sc : -1.f;
            IL[(size_t)m * ldl + (size_t)col * LM + k] = L.v[k] >= 0.f ? L.i[k] : -1;
        }
    }
}

// The local-maximum state parked in LDS between tiles (one slot per thread, fields strided by NT so that a wave's
// accesses are conflict-free), so that it is not held in registers through the K loop.
struct LocalMaxPark {
    float v[LM][NT];
    int i[LM][NT];
    float carry[NT], pv[NT], lv[NT];
    int pt[NT], lt[NT], pok[NT];
};

template <bool ON>
struct ParkSlot {
    LocalMaxPark p;
};
template <>
struct ParkSlot<false> {
    int unused;
};

__device__ __forceinline__ void lm_park(LocalMaxPark& P, const LocalMax& L, const int tid)
{
#pragma unroll
    for (int k = 0; k < LM; k++) { P.v[k][tid] = L.v[k]; P.i[k][tid] = L.i[k]; }
    P.carry[tid] = L.carry; P.pv[tid] = L.pv; P.lv[tid] = L.lv;
    P.pt[tid] = L.pt; P.lt[tid] = L.lt; P.pok[tid] = L.pok;
}

__device__ __forceinline__ void lm_unpark(const LocalMaxPark& P, LocalMax& L, const int tid)
{
#pragma unroll
    for (int k = 0; k < LM; k++) { L.v[k] = P.v[k][tid]; L.i[k] = P.i[k][tid]; }
    L.carry = P.carry[tid]; L.pv = P.pv[tid]; L.lv = P.lv[tid];
    L.pt = P.pt[tid]; L.lt = P.lt[tid]; L.pok = P.pok[tid] != 0;
}

// The epilogue of time tile tt: |z|^2 into the running (best, bi) (first index on ties), in mode 2 also into the
// local-maximum lists; then the window (or block) outputs.
template <int MODE>
__device__ __forceinline__ void epilogue(const Job& J, const GemmItem& it, const int tt, const int half, const int m,
                                         f32x16 (&cr)[4], const f32x16 (&ci)[4], float& best, int& bi, LocalMax& L,
                                         float* __restrict__ V, int* __restrict__ I, const long ldv)
{
    if constexpr (MODE == 2) {
        if (tt % J.per == 0) lm_reset(L);
        float tmax = -1.f;
#pragma unroll
        for (int ts = 0; ts < 4; ts++) {
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int t = tt * BT + ts * 32 + 8 * (r >> 2) + 4 * half + (r & 3);
                const float re = cr[ts][r], im = ci[ts][r];
                const float pp = __builtin_fmaf(re, re, im * im);  // as scan_tile: the same window maxima
                if (t < it.Tn && pp > best) { best = pp; bi = t; }
                cr[ts][r] = t < it.Tn ? pp : -1.f;
                tmax = fmaxf(tmax, cr[ts][r]);
            }
        }
        // local maxima below gate x the window's running maximum (so far, this tile included, over both halves)
        // are below gate x its final maximum, and those below gate x floor (row scaled) below gate x any window
        // maximum that reaches the floor: they need not be listed
        const float sc = (J.rowscale && m < J.M) ? J.rowscale[(size_t)it.seg * J.M + m] : 1.f;
        const float fl = sc > 0.f ? J.floor / sc : 0.f;
        const float thr = J.gate * fmaxf(fmaxf(best, __shfl_xor(best, 32, 64)), fl);
        // a tile with no sample at or above thr in any lane of the wave (nor a held one) is skipped whole, as a run
        if (__ballot(tmax >= thr || (half && L.pok && L.pv >= thr)) != 0) {
            lm_tile(L, cr, tt * BT, half, J.sep, thr);
        } else {
            L.carry = -1.f;
            L.pok = false;
        }
        if ((tt + 1) % J.per == 0) {
            lm_emit(L, half, m, J.M, sc, J.sep, J.vb, J.ib, J.ldb, it.col0 + (tt + 1) / J.per - 1);
        }
    } else {
        scan_tile(cr, ci, tt * BT, half, it.Tn, best, bi);
    }
    tile_done<MODE>(J, it, tt, half, m, best, bi, V, I, ldv);
}

template <int KP, int MODE>
__global__ void __launch_bounds__(NT, 2)
gemm_peaks_bf16_k(const uint16_t* __restrict__ Zt, const long ldz, const uint16_t* __restrict__ Ar,
                  const uint16_t* __restrict__ Ai, const long lda, const Job J, float* __restrict__ V,
                  int* __restrict__ I, const long ldv)
{
    constexpr int KS = KP / 16;          // MFMA k-steps (v_mfma_f32_32x32x16_bf16)
    constexpr int LDP = KP + 8;          // LDS row pitch in elements: 16 rows of one ds_read_b128 phase hit 64 banks
    constexpr int CPR = KP / 8;          // 16-byte chunks per Zt row
    constexpr int CPT = (BT * CPR) / NT; // chunks per thread per tile
    static_assert(KP % 32 == 0 && (BT * CPR) % NT == 0, "K must be a multiple of 32");
    __shared__ __attribute__((aligned(16))) uint16_t lds[2][BT * LDP];

    int mb;
    GemmItem it;
    if (!block_item(J, mb, it)) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col = lane & 31, half = lane >> 5;
    const int m = mb * BMT + wave * 32 + col;
    const int M = J.M;

    uint4 ar[KS], ai[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
        if (m < M) {
            ar[ks] = *(const uint4*)(Ar + (size_t)m * lda + ks * 16 + half * 8);
            ai[ks] = *(const uint4*)(Ai + (size_t)m * lda + ks * 16 + half * 8);
        } else {
            ar[ks] = make_uint4(0, 0, 0, 0);
            ai[ks] = make_uint4(0, 0, 0, 0);
        }
    }

    uint4 zr[CPT];
    auto gload = [&](const int tt) {
        const uint16_t* src = Zt + ((size_t)it.zrow0 + (size_t)tt * BT) * ldz;
#pragma unroll
        for (int j = 0; j < CPT; j++) {
            const int c = tid + j * NT, row = c / CPR, cc = c % CPR;
            zr[j] = *(const uint4*)(src + (size_t)row * ldz + cc * 8);
        }
    };
    auto sstore = [&](const int buf) {
#pragma unroll
        for (int j = 0; j < CPT; j++) {
            const int c = tid + j * NT, row = c / CPR, cc = c % CPR;
            *(uint4*)(&lds[buf][row * LDP + cc * 8]) = zr[j];
        }
    };

    float best = -1.f;
    LocalMax L;
    int bi = it.t_first * BT;
    gload(it.t_first);
    sstore(0);
    __syncthreads();
    for (int tt = it.t_first; tt < it.t_end; tt++) {
        const int buf = (tt - it.t_first) & 1;
        if (tt + 1 < it.t_end) gload(tt + 1);
        f32x16 cr[4], ci[4];
#pragma unroll
        for (int ts = 0; ts < 4; ts++) {
#pragma unroll
            for (int r = 0; r < 16; r++) { cr[ts][r] = 0.f; ci[ts][r] = 0.f; }
        }
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const bf16x8 a_r = __builtin_bit_cast(bf16x8, ar[ks]);
            const bf16x8 a_i = __builtin_bit_cast(bf16x8, ai[ks]);
#pragma unroll
            for (int ts = 0; ts < 4; ts++) {
                const uint4 zv = *(const uint4*)(&lds[buf][(ts * 32 + col) * LDP + ks * 16 + half * 8]);
                const bf16x8 z = __builtin_bit_cast(bf16x8, zv);
                cr[ts] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(z, a_r, cr[ts], 0, 0, 0);
                ci[ts] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(z, a_i, ci[ts], 0, 0, 0);
            }
        }
        epilogue<MODE>(J, it, tt, half, m, cr, ci, best, bi, L, V, I, ldv);
        if (tt + 1 < it.t_end) sstore(buf ^ 1);
        __syncthreads();
    }
}

template <int KP, int MODE>
__global__ void __launch_bounds__(NT, KP <= 32 ? 2 : 1)
gemm_peaks_f32_k(const float* __restrict__ Zt, const long ldz, const float* __restrict__ Ar,
                 const float* __restrict__ Ai, const long lda, const Job J, float* __restrict__ V,
                 int* __restrict__ I, const long ldv)
{
    constexpr int KS = KP / 2;           // MFMA k-steps (v_mfma_f32_32x32x2_f32)
    constexpr int LDP = KP + 2;          // rows 0..31 on even banks, their k+1 partners on odd banks
    constexpr int CPR = KP / 4;          // 16-byte chunks per row
    constexpr int CPT = (BT * CPR) / NT;
    static_assert(KP % 32 == 0 && (BT * CPR) % NT == 0, "K must be a multiple of 32");
    __shared__ __attribute__((aligned(16))) float lds[BT * LDP];

    int mb;
    GemmItem it;
    if (!block_item(J, mb, it)) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col = lane & 31, half = lane >> 5;
    const int m = mb * BMT + wave * 32 + col;
    const int M = J.M;

    float ar[KS], ai[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
        ar[ks] = m < M ? Ar[(size_t)m * lda + 2 * ks + half] : 0.f;
        ai[ks] = m < M ? Ai[(size_t)m * lda + 2 * ks + half] : 0.f;
    }
    float4 zr[CPT];
    auto gload = [&](const int tt) {
        const float* src = Zt + ((size_t)it.zrow0 + (size_t)tt * BT) * ldz;
#pragma unroll
        for (int j = 0; j < CPT; j++) {
            const int c = tid + j * NT, row = c / CPR, cc = c % CPR;
            zr[j] = *(const float4*)(src + (size_t)row * ldz + cc * 4);
        }
    };
    auto sstore = [&]() {
#pragma unroll
        for (int j = 0; j < CPT; j++) {
            const int c = tid + j * NT, row = c / CPR, cc = c % CPR;
            float* d = &lds[row * LDP + cc * 4];        // row pitch is not 16-byte aligned: scalar stores
            d[0] = zr[j].x; d[1] = zr[j].y; d[2] = zr[j].z; d[3] = zr[j].w;
        }
    };

    float best = -1.f;
    LocalMax L;
    int bi = it.t_first * BT;
    gload(it.t_first);
    for (int tt = it.t_first; tt < it.t_end; tt++) {
        __syncthreads();                 // every wave is done reading the previous tile
        sstore();
        __syncthreads();
        if (tt + 1 < it.t_end) gload(tt + 1);
        f32x16 cr[4], ci[4];
#pragma unroll
        for (int ts = 0; ts < 4; ts++) {
#pragma unroll
            for (int r = 0; r < 16; r++) { cr[ts][r] = 0.f; ci[ts][r] = 0.f; }
        }
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
#pragma unroll
            for (int ts = 0; ts < 4; ts++) {
                const float z = lds[(ts * 32 + col) * LDP + 2 * ks + half];
                cr[ts] = __builtin_amdgcn_mfma_f32_32x32x2f32(z, ar[ks], cr[ts], 0, 0, 0);
                ci[ts] = __builtin_amdgcn_mfma_f32_32x32x2f32(z, ai[ks], ci[ts], 0, 0, 0);
            }
        }
        epilogue<MODE>(J, it, tt, half, m, cr, ci, best, bi, L, V, I, ldv);
    }
}

// General-K variant: the block's output tile (128 time samples x 128 templates) accumulates over K in chunks of
// KC staged through LDS for both operands (Zt rows, and the re/im template rows), with the next chunk's global loads
// issued before the current chunk's MFMAs. The loop runs over (time tile, k chunk) pairs, so the prefetch crosses
// time tiles, and the epilogue after each tile's last chunk is the same in-register scan as above.
template <typename T> struct KloopTraits;
template <> struct KloopTraits<uint16_t> { static constexpr int KC = 64, EPV = 8, PAD = 8; };
template <> struct KloopTraits<f16s> { static constexpr int KC = 64, EPV = 8, PAD = 8; };
template <> struct KloopTraits<float> { static constexpr int KC = 32, EPV = 4, PAD = 4; };
template <> struct KloopTraits<uint8_t> { static constexpr int KC = 128, EPV = 16, PAD = 16; };

template <typename T, int MODE, bool ROT_>
__global__ void __launch_bounds__(NT, MODE == 2 ? 1 : 2)
gemm_peaks_kloop_k(const T* __restrict__ Zt, const long ldz, const T* __restrict__ Ar, const T* __restrict__ Ai,
                   const long lda, const int K, const Job J, float* __restrict__ V, int* __restrict__ I,
                   const long ldv)
{
    using TR = KloopTraits<T>;
    constexpr int KC = TR::KC, EPV = TR::EPV, PITCH = KC + TR::PAD;
    constexpr int CPR = KC / EPV;                  // 16-byte chunks per row of a k chunk
    constexpr int CPT = (BT * CPR) / NT;           // per thread, per operand tile of 128 rows
    static_assert((BT * CPR) % NT == 0, "tile shape");
    constexpr bool ROT = ROT_ && sizeof(T) <= 2;   // imaginary part from rotated Zt fragments (no Ai)
    __shared__ __attribute__((aligned(16))) T zs[BT * PITCH];
    __shared__ __attribute__((aligned(16))) T as_r[BMT * PITCH];
    __shared__ __attribute__((aligned(16))) T as_i[ROT ? 16 : BMT * PITCH];

    int mb;
    GemmItem it;
    if (!block_item(J, mb, it)) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col = lane & 31, half = lane >> 5;
    const int m0 = mb * BMT;
    const int m = m0 + wave * 32 + col;
    const int M = J.M;
    const int nkc = (K + KC - 1) / KC;
    const int nsteps = (it.t_end - it.t_first) * nkc;

    uint4 rz[CPT], rr[CPT], ri[CPT];
    auto gload = [&](const int step) {
        const int tt = it.t_first + step / nkc, kc = (step % nkc) * KC;
#pragma unroll
        for (int j = 0; j < CPT; j++) {
            const int c = tid + j * NT, row = c / CPR, cc = kc + (c % CPR) * EPV;
            const bool kin = cc < K;
            rz[j] = kin ? *(const uint4*)(Zt + ((size_t)it.zrow0 + (size_t)tt * BT + row) * ldz + cc)
                        : make_uint4(0, 0, 0, 0);
            const bool ain = kin && (m0 + row) < M;
            rr[j] = ain ? *(const uint4*)(Ar + (size_t)(m0 + row) * lda + cc) : make_uint4(0, 0, 0, 0);
            if constexpr (!ROT)
                ri[j] = ain ? *(const uint4*)(Ai + (size_t)(m0 + row) * lda + cc) : make_uint4(0, 0, 0, 0);
        }
    };
    auto sstore = [&]() {
#pragma unroll
        for (int j = 0; j < CPT; j++) {
            const int c = tid + j * NT, row = c / CPR, cc = (c % CPR) * EPV;
            *(uint4*)(&zs[row * PITCH + cc]) = rz[j];
            *(uint4*)(&as_r[row * PITCH + cc]) = rr[j];
            if constexpr (!ROT) *(uint4*)(&as_i[row * PITCH + cc]) = ri[j];
        }
    };

    f32x16 cr[4], ci[4];
    float best = -1.f;
    LocalMax L;
    int bi = it.t_first * BT;
    gload(0);
    for (int step = 0; step < nsteps; step++) {
        const int kcix = step % nkc, tt = it.t_first + step / nkc;
        if (kcix == 0) {
#pragma unroll
            for (int ts = 0; ts < 4; ts++) {
#pragma unroll
                for (int r = 0; r < 16; r++) { cr[ts][r] = 0.f; ci[ts][r] = 0.f; }
            }
        }
        __syncthreads();
        sstore();
        __syncthreads();
        if (step + 1 < nsteps) gload(step + 1);
        if constexpr (sizeof(T) == 1) {
            // v_mfma_scale_f32_32x32x64_f8f6f4 with unit (E8M0 127) block scales: fp8 e4m3 operands, 32 bytes per
            // lane; lanes 0-31 and 32-63 hold the two 32-byte halves of the 64-wide k step for their row/column
#pragma unroll
            for (int ks = 0; ks < KC / 64; ks++) {
                const T* pa = &as_r[(wave * 32 + col) * PITCH + ks * 64 + half * 32];
                const i32x8 a_r = join8(*(const uint4*)pa, *(const uint4*)(pa + 16));
                i32x8 a_i = a_r;
                if constexpr (!ROT) {
                    const T* pi = &as_i[(wave * 32 + col) * PITCH + ks * 64 + half * 32];
                    a_i = join8(*(const uint4*)pi, *(const uint4*)(pi + 16));
                }
#pragma unroll
                for (int ts = 0; ts < 4; ts++) {
                    const T* pz = &zs[(ts * 32 + col) * PITCH + ks * 64 + half * 32];
                    const uint4 z0 = *(const uint4*)pz, z1 = *(const uint4*)(pz + 16);
                    const i32x8 z = join8(z0, z1);
                    cr[ts] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(z, a_r, cr[ts], 0, 0, 0, 127, 0, 127);
                    if constexpr (ROT)
                        ci[ts] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(join8(rot_fp8(z0), rot_fp8(z1)), a_r,
                                                                                ci[ts], 0, 0, 0, 127, 0, 127);
                    else
                        ci[ts] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(z, a_i, ci[ts], 0, 0, 0, 127, 0, 127);
                }
            }
        } else if constexpr (sizeof(T) == 2) {
#pragma unroll
            for (int ks = 0; ks < KC / 16; ks++) {
                const bf16x8 a_r = __builtin_bit_cast(bf16x8, *(const uint4*)(&as_r[(wave * 32 + col) * PITCH + ks * 16 + half * 8]));
                bf16x8 a_i = a_r;
                if constexpr (!ROT)
                    a_i = __builtin_bit_cast(bf16x8, *(const uint4*)(&as_i[(wave * 32 + col) * PITCH + ks * 16 + half * 8]));
#pragma unroll
                for (int ts = 0; ts < 4; ts++) {
                    const uint4 zv = *(const uint4*)(&zs[(ts * 32 + col) * PITCH + ks * 16 + half * 8]);
                    const bf16x8 z = __builtin_bit_cast(bf16x8, zv);
                    cr[ts] = mfma16<T>(z, a_r, cr[ts], 0, 0, 0);
                    if constexpr (ROT)
                        ci[ts] = mfma16<T>(__builtin_bit_cast(bf16x8, rot_bf16(zv)), a_r,
                                                                         ci[ts], 0, 0, 0);
                    else
                        ci[ts] = mfma16<T>(z, a_i, ci[ts], 0, 0, 0);
                }
            }
        } else {
#pragma unroll
            for (int ks = 0; ks < KC / 2; ks++) {
                const float a_r = as_r[(wave * 32 + col) * PITCH + 2 * ks + half];
                const float a_i = as_i[(wave * 32 + col) * PITCH + 2 * ks + half];
#pragma unroll
                for (int ts = 0; ts < 4; ts++) {
                    const float z = zs[(ts * 32 + col) * PITCH + 2 * ks + half];
                    cr[ts] = __builtin_amdgcn_mfma_f32_32x32x2f32(z, a_r, cr[ts], 0, 0, 0);
                    ci[ts] = __builtin_amdgcn_mfma_f32_32x32x2f32(z, a_i, ci[ts], 0, 0, 0);
                }
            }
        }
        if (kcix == nkc - 1) {
            epilogue<MODE>(J, it, tt, half, m, cr, ci, best, bi, L, V, I, ldv);
        }
    }
}

// bf16 / fp8 K loop with only the Zt chunk staged through LDS (all 4 waves read it) and each wave's Ar fragments
// loaded straight into registers in the MFMA operand layout (no other wave reads them), one step ahead (DB) or at
// the start of the step; the imaginary part comes from rotated Zt fragments (complex-stacked operands only).
template <typename T, int MODE, bool DB>
__global__ void __launch_bounds__(NT, 2)
gemm_peaks_kloop_ra_k(const T* __restrict__ Zt, const long ldz, const T* __restrict__ Ar, const T* __restrict__ Ai,
                      const long lda, const int K, const Job J, float* __restrict__ V, int* __restrict__ I,
                      const long ldv)
{
    static_assert(sizeof(T) <= 2, "bf16 or fp8 operands");
    using TR = KloopTraits<T>;
    constexpr int KC = TR::KC, EPV = TR::EPV, PITCH = KC + TR::PAD;
    constexpr int CPR = KC / EPV;
    constexpr int CPT = (BT * CPR) / NT;
    constexpr int NA = sizeof(T) == 2 ? KC / 16 : 2 * (KC / 64);   // 16-byte fragments per operand per chunk
    static_assert((BT * CPR) % NT == 0, "tile shape");
    __shared__ __attribute__((aligned(16))) T zs[BT * PITCH];

    int mb;
    GemmItem it;
    if (!block_item(J, mb, it)) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col = lane & 31, half = lane >> 5;
    const int m0 = mb * BMT;
    const int m = m0 + wave * 32 + col;
    const int M = J.M;
    const int nkc = (K + KC - 1) / KC;
    const int nsteps = (it.t_end - it.t_first) * nkc;
    const T* __restrict__ arow = Ar + (size_t)(m < M ? m : 0) * lda;

    uint4 rz[CPT];
    auto gload = [&](const int step) {
        const int tt = it.t_first + step / nkc, kc = (step % nkc) * KC;
#pragma unroll
        for (int j = 0; j < CPT; j++) {
            const int c = tid + j * NT, row = c / CPR, cc = kc + (c % CPR) * EPV;
            rz[j] = cc < K ? *(const uint4*)(Zt + ((size_t)it.zrow0 + (size_t)tt * BT + row) * ldz + cc)
                           : make_uint4(0, 0, 0, 0);
        }
    };
    // fragment q of the chunk: bf16 k-step q (8 elements at q 16 + half 8); fp8 k-step q / 2, 16-byte part q % 2
    // (32 bytes at (q / 2) 64 + half 32); K is a multiple of the fragment width, so a fragment is all in or all out
    auto aload = [&](const int step, uint4 (&dr)[NA]) {
        const int kc = (step % nkc) * KC;
#pragma unroll
        for (int q = 0; q < NA; q++) {
            const int k = sizeof(T) == 2 ? kc + q * 16 + half * 8 : kc + (q >> 1) * 64 + half * 32 + (q & 1) * 16;
            dr[q] = m < M && k < K ? *(const uint4*)(arow + k) : make_uint4(0, 0, 0, 0);
        }
    };
    auto sstore = [&]() {
#pragma unroll
        for (int j = 0; j < CPT; j++) {
            const int c = tid + j * NT, row = c / CPR, cc = (c % CPR) * EPV;
            *(uint4*)(&zs[row * PITCH + cc]) = rz[j];
        }
    };

    f32x16 cr[4], ci[4];
    uint4 fr[NA], nr[NA];
    float best = -1.f;
    LocalMax L;
    int bi = it.t_first * BT;
    gload(0);
    if (DB) aload(0, nr);
    for (int step = 0; step < nsteps; step++) {
        const int kcix = step % nkc, tt = it.t_first + step / nkc;
        if (kcix == 0) {
#pragma unroll
            for (int ts = 0; ts < 4; ts++) {
#pragma unroll
                for (int r = 0; r < 16; r++) { cr[ts][r] = 0.f; ci[ts][r] = 0.f; }
            }
        }
        if (!DB) aload(step, fr);
        __syncthreads();
        sstore();
        __syncthreads();
        if (DB) {
#pragma unroll
            for (int q = 0; q < NA; q++) fr[q] = nr[q];
        }
        if (step + 1 < nsteps) {
            gload(step + 1);
            if (DB) aload(step + 1, nr);
        }
        if constexpr (sizeof(T) == 1) {
#pragma unroll
            for (int ks = 0; ks < KC / 64; ks++) {
                const i32x8 a_r = join8(fr[2 * ks], fr[2 * ks + 1]);
#pragma unroll
                for (int ts = 0; ts < 4; ts++) {
                    const T* pz = &zs[(ts * 32 + col) * PITCH + ks * 64 + half * 32];
                    const uint4 z0 = *(const uint4*)pz, z1 = *(const uint4*)(pz + 16);
                    cr[ts] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(join8(z0, z1), a_r, cr[ts], 0, 0, 0, 127,
                                                                            0, 127);
                    ci[ts] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(join8(rot_fp8(z0), rot_fp8(z1)), a_r,
                                                                            ci[ts], 0, 0, 0, 127, 0, 127);
                }
            }
        } else {
#pragma unroll
            for (int ks = 0; ks < KC / 16; ks++) {
                const bf16x8 a_r = __builtin_bit_cast(bf16x8, fr[ks]);
#pragma unroll
                for (int ts = 0; ts < 4; ts++) {
                    const uint4 zv = *(const uint4*)(&zs[(ts * 32 + col) * PITCH + ks * 16 + half * 8]);
                    cr[ts] = mfma16<T>(__builtin_bit_cast(bf16x8, zv), a_r, cr[ts], 0, 0,
                                                                     0);
                    ci[ts] = mfma16<T>(__builtin_bit_cast(bf16x8, rot_bf16(zv)), a_r,
                                                                     ci[ts], 0, 0, 0);
                }
            }
        }
        if (kcix == nkc - 1) {
            epilogue<MODE>(J, it, tt, half, m, cr, ci, best, bi, L, V, I, ldv);
        }
    }
}

// Complex-stacked K loop with two LDS stages: the chunk after next is fetched into registers while the current
// chunk's MFMAs run and is written to the idle stage after them, so a chunk costs one barrier and its global loads
// have a whole chunk of MFMAs to land. Only Zt and Ar are staged (73,728 bytes of LDS per workgroup, two workgroups
// per CU); each wave forms its imaginary-part template fragment from its Ar fragment in registers (arot_*).
template <typename T, int MODE>
__global__ void __launch_bounds__(NT, 2)
gemm_peaks_ring_k(const T* __restrict__ Zt, const long ldz, const T* __restrict__ Ar, const long lda, const int K,
                  const Job J, float* __restrict__ V, int* __restrict__ I, const long ldv)
{
    static_assert(sizeof(T) <= 2, "bf16 or fp8 operands");
    using TR = KloopTraits<T>;
    constexpr int KC = TR::KC, EPV = TR::EPV, PITCH = KC + TR::PAD;
    constexpr int CPR = KC / EPV;
    constexpr int CPT = (BT * CPR) / NT;
    constexpr int STAGE = BT * PITCH;              // elements per operand per stage (BT == BMT rows each)
    static_assert((BT * CPR) % NT == 0 && BT == BMT, "tile shape");
    __shared__ __attribute__((aligned(16))) T zs[2 * STAGE];
    __shared__ __attribute__((aligned(16))) T as[2 * STAGE];

    int mb;
    GemmItem it;
    if (!block_item(J, mb, it)) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col = lane & 31, half = lane >> 5;
    const int m0 = mb * BMT;
    const int m = m0 + wave * 32 + col;
    const int M = J.M;
    const int nkc = (K + KC - 1) / KC;
    const int nsteps = (it.t_end - it.t_first) * nkc;
    if (nsteps <= 0) return;                       // uniform over the workgroup

    uint4 rz[CPT], rr[CPT];
    auto gload = [&](const int step) {
        const int tt = it.t_first + step / nkc, kc = (step % nkc) * KC;
#pragma unroll
        for (int j = 0; j < CPT; j++) {
            const int c = tid + j * NT, row = c / CPR, cc = kc + (c % CPR) * EPV;
            const bool kin = cc < K;
            rz[j] = kin ? *(const uint4*)(Zt + ((size_t)it.zrow0 + (size_t)tt * BT + row) * ldz + cc)
                        : make_uint4(0, 0, 0, 0);
            rr[j] = kin && (m0 + row) < M ? *(const uint4*)(Ar + (size_t)(m0 + row) * lda + cc)
                                          : make_uint4(0, 0, 0, 0);
        }
    };
    auto sstore = [&](const int stage) {
        T* z = zs + stage * STAGE;
        T* a = as + stage * STAGE;
#pragma unroll
        for (int j = 0; j < CPT; j++) {
            const int c = tid + j * NT, row = c / CPR, cc = (c % CPR) * EPV;
            *(uint4*)(&z[row * PITCH + cc]) = rz[j];
            *(uint4*)(&a[row * PITCH + cc]) = rr[j];
        }
    };

    f32x16 cr[4], ci[4];
    float best = -1.f;
    LocalMax L;
    int bi = it.t_first * BT;
    gload(0);
    sstore(0);
    if (nsteps > 1) gload(1);
    __syncthreads();
    for (int step = 0; step < nsteps; step++) {
        const int kcix = step % nkc, tt = it.t_first + step / nkc;
        const T* z = zs + (step & 1) * STAGE;
        const T* a = as + (step & 1) * STAGE;
        if (kcix == 0) {
#pragma unroll
            for (int ts = 0; ts < 4; ts++) {
#pragma unroll
                for (int r = 0; r < 16; r++) { cr[ts][r] = 0.f; ci[ts][r] = 0.f; }
            }
        }
        if constexpr (sizeof(T) == 1) {
#pragma unroll 1
            for (int ks = 0; ks < KC / 64; ks++) {
                const T* pa = &a[(wave * 32 + col) * PITCH + ks * 64 + half * 32];
                const uint4 a0 = *(const uint4*)pa, a1 = *(const uint4*)(pa + 16);
                const i32x8 a_r = join8(a0, a1), a_i = join8(arot_fp8(a0), arot_fp8(a1));
#pragma unroll
                for (int ts = 0; ts < 4; ts++) {
                    const T* pz = &z[(ts * 32 + col) * PITCH + ks * 64 + half * 32];
                    const i32x8 zf = join8(*(const uint4*)pz, *(const uint4*)(pz + 16));
                    cr[ts] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(zf, a_r, cr[ts], 0, 0, 0, 127, 0, 127);
                    ci[ts] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(zf, a_i, ci[ts], 0, 0, 0, 127, 0, 127);
                }
            }
        } else {
#pragma unroll
            for (int ks = 0; ks < KC / 16; ks++) {
                const uint4 av = *(const uint4*)(&a[(wave * 32 + col) * PITCH + ks * 16 + half * 8]);
                const bf16x8 a_r = __builtin_bit_cast(bf16x8, av), a_i = __builtin_bit_cast(bf16x8, arot_bf16(av));
#pragma unroll
                for (int ts = 0; ts < 4; ts++) {
                    const bf16x8 zf = __builtin_bit_cast(bf16x8, *(const uint4*)(&z[(ts * 32 + col) * PITCH + ks * 16 + half * 8]));
                    cr[ts] = mfma16<T>(zf, a_r, cr[ts], 0, 0, 0);
                    ci[ts] = mfma16<T>(zf, a_i, ci[ts], 0, 0, 0);
                }
            }
        }
        if (step + 1 < nsteps) {
            sstore((step + 1) & 1);
            if (step + 2 < nsteps) gload(step + 2);
        }
        if (kcix == nkc - 1) {
            epilogue<MODE>(J, it, tt, half, m, cr, ci, best, bi, L, V, I, ldv);
        }
        __syncthreads();
    }
}

// Complex-stacked K loop fed by LDS DMA (global_load_lds_dwordx4): the hardware writes each 16-byte piece straight
// from L2 into LDS, so no staging registers are held and S stages of 64-byte rows (KC = 32 bf16 / 64 fp8 per row)
// are in flight at once: while stage s is multiplied, stages s+1 .. s+S-1 load. Each wave issues 4 DMA instructions
// per stage (2 for its 32 Zt rows, 2 for its 32 Ar rows; one instruction = 64 lanes x 16 bytes = 16 rows) and waits
// for them with an explicit vmcnt before the stage barrier. The 16-byte slots of row r are stored XOR-swizzled by
// ((r >> 1) ^ (r >> 2)) & 3, which keeps the MFMA fragment reads (32 consecutive rows, one slot) free of LDS bank
// conflicts; the DMA side chooses which global piece each lane loads, so the swizzle costs nothing there.
// The DMA instructions are inline assembly: the compiler does not count them in its vmcnt bookkeeping, so the loop
// holds no compiler-issued vector memory load (the row scale is read before it) and every wait on the stages is
// written out here. The kernel ends only after its last DMA has landed (an exiting workgroup's LDS may be reassigned).
// The list mode (PARK) keeps the local-maximum state in LDS between tiles and runs two stages: 46 KiB of LDS and 168
// registers, so that three workgroups share a CU instead of two (the compiler's few spills fall in the epilogue;
// vector memory returns in order, so the waits above only grow stricter around them). The fp8 window-maximum
// launches are held to the same 168 registers (5 spilled; 5-6% faster than at two workgroups per CU).
constexpr int DMA_STAGES = 3;

__device__ __forceinline__ int dma_swizzle(const int r) { return ((r >> 1) ^ (r >> 2)) & 3; }

__device__ __forceinline__ void dma_load16(const uint32_t lds_addr, const uint32_t voff, const void* sbase)
{
    // M0 holds the wave's LDS destination (lane i writes M0 + 16 i); the global address is sbase + voff
    asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, %2" :: "s"(lds_addr), "v"(voff),
                 "s"(sbase) : "memory");
}

template <typename T, int MODE, bool PARK = false>
__global__ void __launch_bounds__(NT, (PARK || sizeof(T) == 1) ? 3 : 2)
gemm_peaks_dma_k(const T* __restrict__ Zt, const long ldz, const T* __restrict__ Ar, const long lda, const int K,
                 const Job J, float* __restrict__ V, int* __restrict__ I, const long ldv)
{
    static_assert(sizeof(T) <= 2, "bf16 or fp8 operands");
    static_assert(!PARK || MODE == 2, "parking is for the list mode");
    // parked lists: two stages and the lists' 13 KiB of LDS, so that three workgroups fit a CU at 168 registers
    constexpr int S = PARK ? 2 : DMA_STAGES;
    __shared__ ParkSlot<PARK> lpark;
    constexpr int ROWB = 64;                       // bytes per row per stage
    constexpr int KC = ROWB / (int)sizeof(T);      // k per stage
    constexpr int OPB = BT * ROWB, STB = 2 * OPB;  // operand and stage bytes (Zt rows, then Ar rows)
    __shared__ __attribute__((aligned(1024))) uint8_t lds[S * STB];

    int mb;
    GemmItem it;
    if (!block_item(J, mb, it)) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col = lane & 31, half = lane >> 5;
    const int m0 = mb * BMT;
    const int m = m0 + wave * 32 + col;
    const int M = J.M;
    const int nkc = K / KC;
    const int ntl = it.t_end - it.t_first;
    if (ntl <= 0 || nkc <= 0) return;              // uniform over the workgroup, before any DMA
    const float sc = (J.rowscale && m < M) ? J.rowscale[(size_t)it.seg * M + m] : 1.f;

    // DMA lane constants: instruction i of this wave fills rows (2 wave + i) 16 .. + 15, lane -> (row, slot)
    uint32_t zoff[2], aoff[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const int R = (wave * 2 + i) * 16 + (lane >> 2);
        const int c = (lane & 3) ^ dma_swizzle(R);
        zoff[i] = (uint32_t)((size_t)R * ldz * sizeof(T) + c * 16);
        aoff[i] = (uint32_t)((size_t)(min(m0 + R, M - 1) - m0) * lda * sizeof(T) + c * 16);   // rows >= M repeat M-1
    }
    const uint32_t lds0 = (uint32_t)(uintptr_t)lds + __builtin_amdgcn_readfirstlane((uint32_t)(wave * 2 * 1024));
    // issue cursor: the next chunk to load (its tile's first Zt row, its k chunk) and its LDS stage
    const uint8_t* zrow = (const uint8_t*)(Zt + ((size_t)it.zrow0 + (size_t)it.t_first * BT) * ldz);
    const uint8_t* const arow = (const uint8_t*)(Ar + (size_t)m0 * lda);
    const size_t tile_bytes = (size_t)BT * ldz * sizeof(T);
    int ikc = 0, ist = 0;
    auto issue = [&]() {
        const uint8_t* zb = zrow + ikc * ROWB;
        const uint8_t* ab = arow + ikc * ROWB;
        const uint32_t st = lds0 + (uint32_t)(ist * STB);
        dma_load16(st, zoff[0], zb);
        dma_load16(st + 1024, zoff[1], zb);
        dma_load16(st + OPB, aoff[0], ab);
        dma_load16(st + OPB + 1024, aoff[1], ab);
        if (++ikc == nkc) { ikc = 0; zrow += tile_bytes; }
        if (++ist == S) ist = 0;
    };

    const int fs = dma_swizzle(col);               // rows ts 32 + col and wave 32 + col share col's swizzle
    f32x16 cr[4], ci[4];
    float best = -1.f;
    LocalMax L;
    int bi = it.t_first * BT;
    const int nsteps = ntl * nkc;
#pragma unroll
    for (int p = 0; p < S - 1; p++)
        if (p < nsteps) issue();
    int left = nsteps;                             // chunks not yet multiplied
    int cst = 0;                                   // the LDS stage multiplied next
    for (int tl = 0; tl < ntl; tl++) {
        const int tt = it.t_first + tl;
#pragma unroll
        for (int ts = 0; ts < 4; ts++) {
#pragma unroll
            for (int r = 0; r < 16; r++) { cr[ts][r] = 0.f; ci[ts][r] = 0.f; }
        }
        for (int kc = 0; kc < nkc; kc++) {
            // this stage's DMA done (at most the S - 2 later stages, 4 instructions each, still in flight), then
            // every wave's part of it, and every wave done reading the stage the next issue overwrites
            if (left >= S - 1) __builtin_amdgcn_s_waitcnt(0x0F70 | (4 * (S - 2)));   // vmcnt(4 (S - 2))
            else __builtin_amdgcn_s_waitcnt(0x0F70);                                   // vmcnt(0)
            asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
            if (left >= S) issue();
            const uint8_t* z = lds + cst * STB;
            const uint8_t* a = z + OPB;
            if constexpr (sizeof(T) == 1) {
                // one 64-wide k step: lanes 0-31 / 32-63 hold bytes 0-31 / 32-63 (slots 2 half, 2 half + 1)
                const int s0 = ((2 * half) ^ fs) * 16, s1 = ((2 * half + 1) ^ fs) * 16;
                const uint8_t* pa = a + (wave * 32 + col) * ROWB;
                const uint4 a0 = *(const uint4*)(pa + s0), a1 = *(const uint4*)(pa + s1);
                const i32x8 a_r = join8(a0, a1), a_i = join8(arot_fp8(a0), arot_fp8(a1));
#pragma unroll
                for (int ts = 0; ts < 4; ts++) {
                    const uint8_t* pz = z + (ts * 32 + col) * ROWB;
                    const i32x8 zf = join8(*(const uint4*)(pz + s0), *(const uint4*)(pz + s1));
                    cr[ts] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(zf, a_r, cr[ts], 0, 0, 0, 127, 0, 127);
                    ci[ts] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(zf, a_i, ci[ts], 0, 0, 0, 127, 0, 127);
                }
            } else {
#pragma unroll
                for (int ks = 0; ks < KC / 16; ks++) {
                    const int so = ((2 * ks + half) ^ fs) * 16;
                    const uint4 av = *(const uint4*)(a + (wave * 32 + col) * ROWB + so);
                    const bf16x8 a_r = __builtin_bit_cast(bf16x8, av), a_i = __builtin_bit_cast(bf16x8, arot_bf16(av));
#pragma unroll
                    for (int ts = 0; ts < 4; ts++) {
                        const bf16x8 zf = __builtin_bit_cast(bf16x8, *(const uint4*)(z + (ts * 32 + col) * ROWB + so));
                        cr[ts] = mfma16<T>(zf, a_r, cr[ts], 0, 0, 0);
                        ci[ts] = mfma16<T>(zf, a_i, ci[ts], 0, 0, 0);
                    }
                }
            }
            if (++cst == S) cst = 0;
            left--;
        }
        if constexpr (PARK) {
            LocalMax Lp;
            lm_unpark(lpark.p, Lp, tid);
            epilogue<MODE>(J, it, tt, half, m, cr, ci, best, bi, Lp, V, I, ldv);
            lm_park(lpark.p, Lp, tid);
        } else {
            epilogue<MODE>(J, it, tt, half, m, cr, ci, best, bi, L, V, I, ldv);
        }
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);            // vmcnt(0): no DMA in flight when the workgroup ends
}

template <typename T>
using resident_fn = void (*)(const T*, long, const T*, const T*, long, const Job, float*, int*, long);

template <int KP>
resident_fn<uint16_t> pick_bf16(int mode) { return mode ? gemm_peaks_bf16_k<KP, 1> : gemm_peaks_bf16_k<KP, 0>; }
template <int KP>
resident_fn<float> pick_f32(int mode) { return mode ? gemm_peaks_f32_k<KP, 1> : gemm_peaks_f32_k<KP, 0>; }

resident_fn<uint16_t> bf16_kernel(int K, int mode)
{
    switch (K) {
    case 32: return pick_bf16<32>(mode);
    case 64: return pick_bf16<64>(mode);
    case 96: return pick_bf16<96>(mode);
    case 128: return pick_bf16<128>(mode);
    case 160: return pick_bf16<160>(mode);
    case 192: return pick_bf16<192>(mode);
    case 224: return pick_bf16<224>(mode);
    case 256: return pick_bf16<256>(mode);
    default: return nullptr;
    }
}

resident_fn<float> f32_kernel(int K, int mode)
{
    switch (K) {
    case 32: return pick_f32<32>(mode);
    case 64: return pick_f32<64>(mode);
    case 96: return pick_f32<96>(mode);
    case 128: return pick_f32<128>(mode);
    default: return nullptr;
    }
}

template <typename T, bool ROT>
void launch_kloop(int mode, dim3 grid, dim3 block, hipStream_t s, const T* Zt, long ldz, const T* Ar, const T* Ai,
                  long lda, int K, const Job& J, float* V, int* I, long ldv)
{
    hipLaunchKernelGGL((mode == 2 ? gemm_peaks_kloop_k<T, 2, ROT> : mode ? gemm_peaks_kloop_k<T, 1, ROT>
                                                                         : gemm_peaks_kloop_k<T, 0, ROT>),
                       grid, block, 0, s, Zt, ldz, Ar, Ai, lda, K, J, V, I, ldv);
}

template <typename T, bool DB>
void launch_ra(int mode, dim3 grid, dim3 block, hipStream_t s, const T* Zt, long ldz, const T* Ar, const T* Ai,
               long lda, int K, const Job& J, float* V, int* I, long ldv)
{
    hipLaunchKernelGGL((mode ? gemm_peaks_kloop_ra_k<T, 1, DB> : gemm_peaks_kloop_ra_k<T, 0, DB>), grid, block, 0, s,
                       Zt, ldz, Ar, Ai, lda, K, J, V, I, ldv);
}

template <typename T>
void launch_ring(int mode, dim3 grid, dim3 block, hipStream_t s, const T* Zt, long ldz, const T* Ar, long lda, int K,
                 const Job& J, float* V, int* I, long ldv)
{
    hipLaunchKernelGGL((mode ? gemm_peaks_ring_k<T, 1> : gemm_peaks_ring_k<T, 0>), grid, block, 0, s, Zt, ldz, Ar,
                       lda, K, J, V, I, ldv);
}

template <typename T>
void launch_dma(int mode, dim3 grid, dim3 block, hipStream_t s, const T* Zt, long ldz, const T* Ar, long lda, int K,
                const Job& J, float* V, int* I, long ldv)
{
    hipLaunchKernelGGL((mode ? gemm_peaks_dma_k<T, 1> : gemm_peaks_dma_k<T, 0>), grid, block, 0, s, Zt, ldz, Ar, lda, K,
                       J, V, I, ldv);
}

// the list mode: the lists' state parked in LDS between tiles (three workgroups per CU instead of two; 7% faster)
template <typename T>
void launch_dma_park(dim3 grid, dim3 block, hipStream_t s, const T* Zt, long ldz, const T* Ar, long lda, int K,
                     const Job& J, float* V, int* I, long ldv)
{
    hipLaunchKernelGGL((gemm_peaks_dma_k<T, 2, true>), grid, block, 0, s, Zt, ldz, Ar, lda, K, J, V, I, ldv);
}

// bf16 / fp8 K loops for complex-stacked operands: variant 0 or 5 the LDS-DMA loop (when K is a multiple of its
// 64-byte chunk, else the ring), 4 the two-stage ring, 2 LDS-staged with rotated Zt, 3 register fragments at the step
// start, 1 register fragments one chunk ahead; without the stacking, LDS-staged with both template operands
template <typename T>
void launch_kloop_any(int variant, bool rot, int mode, dim3 grid, dim3 block, hipStream_t s, const T* Zt, long ldz,
                      const T* Ar, const T* Ai, long lda, int K, const Job& J, float* V, int* I, long ldv)
{
    const bool dma_ok = K % (64 / (int)sizeof(T)) == 0;
    if (mode == 2) {
        // the local-maximum lists: the LDS-DMA loop where K allows, else the LDS-staged loop
        if (rot && dma_ok) launch_dma_park<T>(grid, block, s, Zt, ldz, Ar, lda, K, J, V, I, ldv);
        else if (rot) launch_kloop<T, true>(mode, grid, block, s, Zt, ldz, Ar, Ai, lda, K, J, V, I, ldv);
        else launch_kloop<T, false>(mode, grid, block, s, Zt, ldz, Ar, Ai, lda, K, J, V, I, ldv);
        return;
    }
    if (!rot) launch_kloop<T, false>(mode, grid, block, s, Zt, ldz, Ar, Ai, lda, K, J, V, I, ldv);
    else if ((variant == 0 || variant == 5) && dma_ok) launch_dma<T>(mode, grid, block, s, Zt, ldz, Ar, lda, K, J, V,
                                                                     I, ldv);
    else if (variant == 4 || variant == 5 || variant == 0) launch_ring<T>(mode, grid, block, s, Zt, ldz, Ar, lda, K, J,
                                                                          V, I, ldv);
    else if (variant == 2) launch_kloop<T, true>(mode, grid, block, s, Zt, ldz, Ar, Ai, lda, K, J, V, I, ldv);
    else if (variant == 3) launch_ra<T, false>(mode, grid, block, s, Zt, ldz, Ar, Ai, lda, K, J, V, I, ldv);
    else if (variant == 1) launch_ra<T, true>(mode, grid, block, s, Zt, ldz, Ar, Ai, lda, K, J, V, I, ldv);
    else launch_kloop<T, true>(mode, grid, block, s, Zt, ldz, Ar, Ai, lda, K, J, V, I, ldv);
}

int launch(int dtype, int variant, const void* Zt, long ldz, const void* Ar, const void* Ai, long lda, int K, int mode,
           const Job& J, float* V, int* I, long ldv, hipStream_t s)
{
    const bool rot = (mode & 4) != 0;
    mode &= 3;
    const long nblocks = (long)J.nitems * J.nmb;
    GWGS_CHECK(nblocks < (1L << 30), GWGS_ESIZE);
    if (nblocks == 0) return GWGS_OK;
    const dim3 grid((unsigned)((nblocks + 7) / 8 * 8)), block(NT);
    const bool aligned = ((uintptr_t)Zt & 15) == 0 && ((uintptr_t)Ar & 15) == 0 && ((uintptr_t)Ai & 15) == 0;
    // mode 2 (local-maximum lists) runs on the K loops only
    const bool kloop = variant >= 1 || dtype == 2 || (dtype == 1 ? K > 256 : K > 128) || K % 32 || mode == 2;
    if (dtype == 1) {
        GWGS_CHECK(aligned && ldz % 8 == 0 && lda % 8 == 0 && K % 8 == 0, GWGS_EINVAL);
        if (kloop)
            launch_kloop_any<uint16_t>(variant, rot, mode, grid, block, s, (const uint16_t*)Zt, ldz,
                                       (const uint16_t*)Ar, (const uint16_t*)Ai, lda, K, J, V, I, ldv);
        else
            hipLaunchKernelGGL(bf16_kernel(K, mode), grid, block, 0, s, (const uint16_t*)Zt, ldz, (const uint16_t*)Ar,
                               (const uint16_t*)Ai, lda, J, V, I, ldv);
    } else if (dtype == 0) {
        GWGS_CHECK(aligned && ldz % 4 == 0 && lda % 4 == 0 && K % 4 == 0, GWGS_EINVAL);
        if (kloop)
            launch_kloop<float, false>(mode, grid, block, s, (const float*)Zt, ldz, (const float*)Ar, (const float*)Ai,
                                       lda, K, J, V, I, ldv);
        else
            hipLaunchKernelGGL(f32_kernel(K, mode), grid, block, 0, s, (const float*)Zt, ldz, (const float*)Ar,
                               (const float*)Ai, lda, J, V, I, ldv);
    } else if (dtype == 2) {
        GWGS_CHECK(aligned && ldz % 16 == 0 && lda % 16 == 0 && K % 16 == 0, GWGS_EINVAL);
        launch_kloop_any<uint8_t>(variant, rot, mode, grid, block, s, (const uint8_t*)Zt, ldz, (const uint8_t*)Ar,
                                  (const uint8_t*)Ai, lda, K, J, V, I, ldv);
    } else if (dtype == 3) {
        // fp16 operands (the reference's type): the K loops only, with v_mfma_f32_32x32x16_f16
        GWGS_CHECK(aligned && ldz % 8 == 0 && lda % 8 == 0 && K % 8 == 0, GWGS_EINVAL);
        launch_kloop_any<f16s>(variant, rot, mode, grid, block, s, (const f16s*)Zt, ldz, (const f16s*)Ar,
                               (const f16s*)Ai, lda, K, J, V, I, ldv);
    } else {
        return GWGS_EUNSUP;
    }
    return gwgs_status();
}

} // namespace

// One segment. dtype 0 = fp32, 1 = bf16, 2 = fp8 e4m3 (OCP), 3 = fp16 operands. variant: 0 = pick by K (register-resident
// template operands up to K = 256 bf16 / 128 fp32, else the K loop), 1 = always the K loop (fp8 always).
// rowscale (M floats, may be null) multiplies each template's reported |z|^2.
GWGS_API int gwgs_gemm_peaks(int dtype, int variant, const void* Zt, long ldz, const void* Ar, const void* Ai, long lda,
                             int M, int K, int Tn, int mode, int per, int tiles_per_block, float* V, int* I, long ldv,
                             const float* rowscale, void* stream)
{
    GWGS_CHECK(Zt && Ar && Ai && V && I, GWGS_EINVAL);
    GWGS_CHECK(M > 0 && Tn > 0 && K > 0 && ((mode & 3) == 0 || (mode & 3) == 1) && (mode & ~7) == 0, GWGS_EINVAL);
    GWGS_CHECK(ldz >= K && lda >= K && ldv > 0, GWGS_EINVAL);
    const int ntile = (Tn + BT - 1) / BT;
    if ((mode & 3) == 1) {
        GWGS_CHECK(per > 0 && Tn % (per * BT) == 0, GWGS_EINVAL);
        GWGS_CHECK(ldv >= Tn / (per * BT), GWGS_EINVAL);
    } else {
        per = 1;
        GWGS_CHECK(ldv >= ntile, GWGS_EINVAL);
    }
    int tpb = tiles_per_block > 0 ? tiles_per_block : ((mode & 3) == 1 ? per : 16);
    if ((mode & 3) == 1 && tpb % per) tpb = (tpb / per + 1) * per;
    Job J;
    J.nitems_dev = nullptr;
    J.items = nullptr; J.nmb = (M + BMT - 1) / BMT; J.tpb = tpb; J.ntile = ntile; J.Tn = Tn; J.M = M; J.per = per;
    J.nitems = (ntile + tpb - 1) / tpb; J.rowscale = rowscale; J.vb = nullptr; J.ib = nullptr; J.ldb = 0;
    J.sep = 0; J.gate = 0.f; J.floor = 0.f;
    return launch(dtype, variant, Zt, ldz, Ar, Ai, lda, K, mode, J, V, I, ldv, (hipStream_t)stream);
}

// Several segments in one launch. items: device array of nitems GemmItem (32 bytes each: zrow0 int64, t_first,
// t_end, col0, seg, Tn, pad int32), with t_first a multiple of `per` in mode 1. rowscale: (nseg, M) floats or null.
GWGS_API int gwgs_gemm_peaks_multi(int dtype, int variant, const void* Zt, long ldz, const void* Ar, const void* Ai,
                                   long lda, int M, int K, const void* items, int nitems, int mode, int per, float* V,
                                   int* I, long ldv, const float* rowscale, void* stream)
{
    GWGS_CHECK(Zt && Ar && Ai && V && I && items && nitems >= 0, GWGS_EINVAL);
    GWGS_CHECK(M > 0 && K > 0 && ((mode & 3) == 0 || (mode & 3) == 1) && (mode & ~7) == 0 && ldz >= K && lda >= K &&
               ldv > 0, GWGS_EINVAL);
    GWGS_CHECK(((uintptr_t)items & 7) == 0, GWGS_EINVAL);
    Job J;
    J.nitems_dev = nullptr;
    J.items = (const GemmItem*)items; J.nitems = nitems; J.nmb = (M + BMT - 1) / BMT; J.tpb = 0; J.ntile = 0;
    J.Tn = 0; J.M = M; J.per = (mode & 3) == 1 ? per : 1; J.rowscale = rowscale; J.vb = nullptr; J.ib = nullptr;
    J.ldb = 0; J.sep = 0; J.gate = 0.f; J.floor = 0.f;
    GWGS_CHECK(J.per > 0, GWGS_EINVAL);
    return launch(dtype, variant, Zt, ldz, Ar, Ai, lda, K, mode, J, V, I, ldv, (hipStream_t)stream);
}

// As gwgs_gemm_peaks_multi in mode 1 (windows of `per` tiles), and for every (template, window) also its LM = 4
// largest separated local maxima of |z|^2 (p[t] >= p[t - 1] and p[t] > p[t + 1], open at the window's edges; of
// maxima within sep samples the larger, see LocalMax), larger value first, then smaller sample: VL[m ldl + col LM +
// k] the row-scaled value (-1 where the window has fewer), IL the segment-relative sample (-1 likewise). ldl >= LM x
// the window columns. Local maxima below gate x max(the window's maximum, floor) (0 <= gate <= 1; floor on the
// row-scaled |z|^2) need not be listed; the largest ones at or above it are.
GWGS_API int gwgs_gemm_peaks_multi_lm(int dtype, int variant, const void* Zt, long ldz, const void* Ar, const void* Ai,
                                      long lda, int M, int K, const void* items, int nitems, int stacked, int per,
                                      int sep, float gate, float floor, float* V, int* I, long ldv, float* VL, int* IL,
                                      long ldl, const float* rowscale, void* stream)
{
    GWGS_CHECK(Zt && Ar && Ai && V && I && VL && IL && items && nitems >= 0, GWGS_EINVAL);
    GWGS_CHECK(M > 0 && K > 0 && (stacked == 0 || stacked == 1) && per > 0 && sep >= 0 && sep < per * BT &&
               gate >= 0.f && gate <= 1.f && floor >= 0.f &&
               ldz >= K && lda >= K && ldv > 0 && ldl >= (long)LM * ldv, GWGS_EINVAL);
    GWGS_CHECK(((uintptr_t)items & 7) == 0, GWGS_EINVAL);
    Job J;
    J.nitems_dev = nullptr;
    J.items = (const GemmItem*)items; J.nitems = nitems; J.nmb = (M + BMT - 1) / BMT; J.tpb = 0; J.ntile = 0;
    J.Tn = 0; J.M = M; J.per = per; J.rowscale = rowscale; J.vb = VL; J.ib = IL; J.ldb = ldl; J.sep = sep; J.gate = gate; J.floor = floor;
    return launch(dtype, variant, Zt, ldz, Ar, Ai, lda, K, 2 | (stacked ? 4 : 0), J, V, I, ldv, (hipStream_t)stream);
}

// gwgs_gemm_peaks_multi_lm over an item table whose length is in device memory (*nitems_dev <= max_items), such as
// gwgs_hot_items writes: the grid is sized for max_items and blocks past the count return at once.
GWGS_API int gwgs_gemm_peaks_multi_lm_dyn(int dtype, int variant, const void* Zt, long ldz, const void* Ar,
                                          const void* Ai, long lda, int M, int K, const void* items,
                                          const int* nitems_dev, int max_items, int stacked, int per, int sep,
                                          float gate, float floor, float* V, int* I, long ldv, float* VL, int* IL,
                                          long ldl, const float* rowscale, void* stream)
{
    GWGS_CHECK(Zt && Ar && Ai && V && I && VL && IL && items && nitems_dev && max_items >= 0, GWGS_EINVAL);
    GWGS_CHECK(M > 0 && K > 0 && (stacked == 0 || stacked == 1) && per > 0 && sep >= 0 && sep < per * BT &&
               gate >= 0.f && gate <= 1.f && floor >= 0.f &&
               ldz >= K && lda >= K && ldv > 0 && ldl >= (long)LM * ldv, GWGS_EINVAL);
    GWGS_CHECK(((uintptr_t)items & 7) == 0, GWGS_EINVAL);
    Job J;
    J.nitems_dev = nitems_dev;
    J.items = (const GemmItem*)items; J.nitems = max_items; J.nmb = (M + BMT - 1) / BMT; J.tpb = 0; J.ntile = 0;
    J.Tn = 0; J.M = M; J.per = per; J.rowscale = rowscale; J.vb = VL; J.ib = IL; J.ldb = ldl; J.sep = sep; J.gate = gate; J.floor = floor;
    return launch(dtype, variant, Zt, ldz, Ar, Ai, lda, K, 2 | (stacked ? 4 : 0), J, V, I, ldv, (hipStream_t)stream);
}

namespace {

// One segment of a batch for gwgs_hot_items: its first Zt row, first window column, analysed samples.
struct HotSeg {
    long long zrow0;
    int col0;
    int Tn;
};

__global__ void __launch_bounds__(256) hot_zero_k(int* __restrict__ flags, const int n, int* __restrict__ nitems)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) flags[i] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) *nitems = 0;
}

__global__ void __launch_bounds__(256) hot_mark_k(const int* __restrict__ idx, const float* __restrict__ val,
                                                  const int* __restrict__ count, const int cap, const int nwsum,
                                                  const float cut, int* __restrict__ flags)
{
    const int n = min(*count, cap);
    for (int c = blockIdx.x * 256 + threadIdx.x; c < n; c += gridDim.x * 256)
        if (val[c] >= cut) flags[idx[c] % nwsum] = 1;
}

__global__ void __launch_bounds__(256) hot_items_k(const int* __restrict__ flags, const int nwsum,
                                                   const HotSeg* __restrict__ segs, const int nseg, const int per,
                                                   GemmItem* __restrict__ items, int* __restrict__ nitems)
{
    for (int col = blockIdx.x * 256 + threadIdx.x; col < nwsum; col += gridDim.x * 256) {
        if (!flags[col]) continue;
        int s = 0;
        while (s + 1 < nseg && segs[s + 1].col0 <= col) s++;
        GemmItem it;
        it.zrow0 = segs[s].zrow0;
        it.t_first = (col - segs[s].col0) * per;
        it.t_end = it.t_first + per;
        it.col0 = segs[s].col0;
        it.seg = s;
        it.Tn = segs[s].Tn;
        it.pad = 0;
        items[atomicAdd(nitems, 1)] = it;
    }
}

} // namespace

// The windows holding candidates at or above `cut` as one-window GEMM items (for gwgs_gemm_peaks_multi_lm_dyn): from
// a compacted candidate list (idx: row-major index into (M, nwsum) window maxima; *count entries, at most cap), flag
// each candidate window's column, then write one item per flagged column (its segment from segs: nseg rows of
// {int64 first Zt row, int32 first column, int32 analysed samples}, columns ascending) in any order, and their count
// to *nitems. flags: nwsum int32 of workspace; items: room for nwsum.
GWGS_API int gwgs_hot_items(const int* idx, const float* val, const int* count, int cap, int nwsum, float cut,
                            const void* segs, int nseg, int per, int* flags, void* items, int* nitems, void* stream)
{
    GWGS_CHECK(idx && val && count && segs && flags && items && nitems && cap >= 0 && nwsum > 0 && nseg > 0 &&
               per > 0, GWGS_EINVAL);
    GWGS_CHECK(((uintptr_t)items & 7) == 0 && ((uintptr_t)segs & 7) == 0, GWGS_EINVAL);
    hipStream_t s = (hipStream_t)stream;
    const int nb = std::min((nwsum + 255) / 256, 1024);
    hipLaunchKernelGGL(hot_zero_k, dim3(nb), dim3(256), 0, s, flags, nwsum, nitems);
    if (cap > 0)
        hipLaunchKernelGGL(hot_mark_k, dim3(std::min((cap + 255) / 256, 1024)), dim3(256), 0, s, idx, val, count, cap,
                           nwsum, cut, flags);
    hipLaunchKernelGGL(hot_items_k, dim3(nb), dim3(256), 0, s, flags, nwsum, (const HotSeg*)segs, nseg, per,
                       (GemmItem*)items, nitems);
    return gwgs_status();
}

GWGS_API int gwgs_gemm_lm_count(void) { return LM; }

GWGS_API int gwgs_gemm_item_bytes(void) { return (int)sizeof(GemmItem); }

GWGS_API int gwgs_gemm_max_k(int dtype)
{
    return dtype >= 0 && dtype <= 3 ? (1 << 20) : 0;
}
