// Basis series of one tile against one segment by a fused four-step inverse FFT, written straight into the GEMM's
// time-major operand (the reference's fp_filter.py basis_series :597-613: q = conj(Vh) D / sw zero-padded to the
// segment length, z = ifft(q) * n, then the time-major fp16/fp32 stack the GEMM reads).
//
// n = n1 n2 (powers of two, 256..2048 each). With k = k1 + n1 k2 and t = t2 + n2 t1:
//   z_j[t] = sum_k1 w_n^(k1 t2) w_n1^(k1 t1) sum_k2 x_j[k1 + n1 k2] w_n2^(k2 t2),   w_m = e^(+2 pi i / m),
//   x_j[k] = Vc[j, k] D[k] on [kmin, kend) and 0 elsewhere.
// cols_k: per basis row j and 8 to 32 columns k1, the length-n2 transform over k2 (form of x fused into the load,
// zeros above kend never read), times w_n^(k1 t2), to Y[j, t2, k1] (complex64, r x n). The band [kmin, kend) is
// usually a small part of n, so x[k1 + n1 k2] is zero for all k2 >= kend / n1: the kernel is instantiated for the
// number of leading nonzero inputs per thread (Q / 8, Q / 4, Q / 2, Q) and its first DFT skips the known zeros.
// rows_k: per t2 and 8 to 32 basis rows, the length-n1 transform over k1 of Y[j, t2, :], written as
// Zt[t - a0, 2j] = Re z_j[t] cs_j, Zt[t - a0, 2j + 1] = Im z_j[t] cs_j for t in [a0, a0 + Tn) (fp32, bf16 or fp8), and
// zeros for rows r <= j < rp.
// Each length-L line transform runs on P threads holding Q = L / P values: a Q-point DFT in registers per thread,
// the twiddle w_L^(p m2), an exchange through LDS (real then imaginary parts, so two blocks fit a CU), and
// Q / P P-point DFTs per thread. Twiddles of the in-register DFTs are compile-time constants; the others come from
// float64 sincospi seeds advanced by float64 products (rounded once to float32).
// Inputs are read through buffer resources whose range check returns zeros outside the band (cols_k) or past the
// last basis row (rows_k): every load of a thread is issued unconditionally, back to back, where a per-load branch
// made the compiler wait for each load (or pair of loads) before issuing the next.
// One launch of each pass covers every segment of a batch (blockIdx.z = segment, up to kMaxSeg): the segments share
// the tile's Vc and band and have their own D, intermediate, analysed span and output rows.
// HBM traffic per (tile, segment): 8 r (kend - kmin) bytes read of Vc, 8 r n written and read of Y, 2 Tn rp
// elements of the output type written: two passes over Y instead of the batched FFT's several plus its input and
// transpose passes.
#include "common.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

// e^(+2 pi i k / E) for k < E in float, evaluated by the compiler (Taylor series on the first quadrant in double,
// rotated by whole quadrants)
template <int E>
struct Twiddles {
    float c[E], s[E];
    constexpr Twiddles() : c(), s()
    {
        for (int k = 0; k < E; k++) {
            const int quad = (4 * k) / E;
            const double x = 0.5 * kPi * (double)(4 * k - quad * E) / E;
            double t = x, sn = x, u = 1, cs = 1;
            for (int i = 1; i < 14; i++) {
                t *= -x * x / ((2.0 * i) * (2.0 * i + 1));
                sn += t;
                u *= -x * x / ((2.0 * i - 1) * (2.0 * i));
                cs += u;
            }
            for (int i = 0; i < quad; i++) { const double w = cs; cs = -sn; sn = w; }
            c[k] = (float)cs;
            s[k] = (float)sn;
        }
    }
};
template <int E>
__device__ constexpr Twiddles<E> kTw{};

constexpr int ilog2(int e) { int l = 0; while ((1 << l) < e) l++; return l; }
constexpr int brev(int i, int lg) { int r = 0; for (int b = 0; b < lg; b++) r |= ((i >> b) & 1) << (lg - 1 - b); return r; }

// Zero<E, NZ>::m[s] bit i: before the radix-2 stage of span 2^s, w[i] is a known zero (every input feeding it, in
// bit-reversed order, has index NZ or above)
template <int E, int NZ>
struct Zero {
    unsigned long long m[ilog2(E) + 1];
    constexpr Zero() : m()
    {
        constexpr int LG = ilog2(E);
        for (int st = 0; st <= LG; st++) {
            const int h = 1 << st;
            unsigned long long bits = 0;
            for (int i = 0; i < E; i++) {
                bool z = true;
                for (int x = i & ~(h - 1); x < (i & ~(h - 1)) + h; x++)
                    if (brev(x, LG) < NZ) z = false;
                if (z) bits |= 1ull << i;
            }
            m[st] = bits;
        }
    }
};

// v[m] <- sum_n v[n] e^(+2 pi i n m / E), radix-2 decimation in time, every index a compile-time constant. Inputs
// v[n] with n >= NZ are zero and not read: butterflies with a known-zero operand become copies. A butterfly with a
// general twiddle (c, s) is a +- g (b + i u b) with (g, u) = (c, s / c) when |c| >= |s| and (s, -c / s) times i
// otherwise: six fused multiply-adds instead of four products and four sums.
template <int E, int NZ = E>
__device__ __forceinline__ void dft_reg(float2 (&v)[E])
{
    constexpr int LG = ilog2(E);
    constexpr Zero<E, NZ> zm{};
    float2 w[E];
#pragma unroll
    for (int i = 0; i < NZ; i++) w[brev(i, LG)] = v[i];
#pragma unroll
    for (int st = 0; st < LG; st++)
#pragma unroll
        for (int k = 0; k < (1 << st); k++) {
            const int h = 1 << st;
            const int ti = k * (E / (2 * h));
#pragma unroll
            for (int i = k; i < E; i += 2 * h) {
                const bool za = (zm.m[st] >> i) & 1, zb = (zm.m[st] >> (i + h)) & 1;
                if (za && zb) continue;
                const float2 a = w[i], b0 = w[i + h];
                if (zb) { w[i + h] = a; continue; }
                const bool general = ti != 0 && 4 * ti != E;
                if (general && !za) {
                    const float c = kTw<E>.c[ti], s = kTw<E>.s[ti];
                    float2 bb;
                    float g;
                    if (__builtin_fabsf(c) >= __builtin_fabsf(s)) {
                        const float u = (float)((double)s / (double)c);
                        bb = make_float2(__builtin_fmaf(-u, b0.y, b0.x), __builtin_fmaf(u, b0.x, b0.y));
                        g = c;
                    } else {
                        const float u = (float)((double)c / (double)s);
                        bb = make_float2(__builtin_fmaf(u, b0.x, -b0.y), __builtin_fmaf(u, b0.y, b0.x));
                        g = s;
                    }
                    w[i] = make_float2(__builtin_fmaf(g, bb.x, a.x), __builtin_fmaf(g, bb.y, a.y));
                    w[i + h] = make_float2(__builtin_fmaf(-g, bb.x, a.x), __builtin_fmaf(-g, bb.y, a.y));
                    continue;
                }
                float2 b;
                if (ti == 0) b = b0;
                else if (4 * ti == E) b = make_float2(-b0.y, b0.x);
                else {
                    const float c = kTw<E>.c[ti], s = kTw<E>.s[ti];
                    b = make_float2(b0.x * c - b0.y * s, b0.x * s + b0.y * c);
                }
                if (za) {
                    w[i] = b;
                    w[i + h] = make_float2(-b.x, -b.y);
                } else {
                    w[i] = make_float2(a.x + b.x, a.y + b.y);
                    w[i + h] = make_float2(a.x - b.x, a.y - b.y);
                }
            }
        }
#pragma unroll
    for (int i = 0; i < E; i++) v[i] = w[i];
}

// e^(+2 pi i e / m) for an integer e reduced mod m (m a power of two), in double
__device__ __forceinline__ void unit(const long e, const long m, double& c, double& s)
{
    const long r = e & (m - 1);
    sincospi(2.0 * (double)r / (double)m, &s, &c);
}

__device__ __forceinline__ float2 cmulf(const float2 a, const double c, const double s)
{
    const float fc = (float)c, fs = (float)s;
    return make_float2(a.x * fc - a.y * fs, a.x * fs + a.y * fc);
}

__device__ __forceinline__ void advance(double& c, double& s, const double dc, const double ds)
{
    const double t = c * dc - s * ds;
    s = c * ds + s * dc;
    c = t;
}

// A twiddle e^(2 pi i e / m) advanced by a fixed step: float64 seeds and products, rounded once to float32 where
// applied (Rot<false>); or float32 throughout (Rot<true>, for the forms whose intermediate is fp16: sincospif seeds
// of exact arguments and float products drift by a few 1e-7 over the 64 steps of a line, far below fp16's 2^-11).
template <bool LP>
struct Rot;
template <>
struct Rot<false> {
    double c = 1.0, s = 0.0, dc, ds;
    __device__ __forceinline__ void seed(const long e, const long m) { unit(e, m, c, s); }
    __device__ __forceinline__ void step(const long e, const long m) { unit(e, m, dc, ds); }
    __device__ __forceinline__ void next() { advance(c, s, dc, ds); }
    __device__ __forceinline__ float2 apply(const float2 a) const { return cmulf(a, c, s); }
};
template <>
struct Rot<true> {
    float c = 1.f, s = 0.f, dc, ds;
    static __device__ __forceinline__ void unitf(const long e, const long m, float& co, float& si)
    {
        // (e mod m) < m <= 2^22: 2 (e mod m) / m is exact in float
        sincospif(2.f * (float)(e & (m - 1)) / (float)m, &si, &co);
    }
    __device__ __forceinline__ void seed(const long e, const long m) { unitf(e, m, c, s); }
    __device__ __forceinline__ void step(const long e, const long m) { unitf(e, m, dc, ds); }
    __device__ __forceinline__ void next()
    {
        const float t = c * dc - s * ds;
        s = c * ds + s * dc;
        c = t;
    }
    __device__ __forceinline__ float2 apply(const float2 a) const
    {
        return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
    }
};

// v[m2] *= w_L^(p m2) for 0 < m2 < Q
template <int Q, bool LP>
__device__ __forceinline__ void twiddle_line(float2 (&v)[Q], const int p, const int L)
{
    Rot<LP> w;
    w.step(p, L);
#pragma unroll
    for (int m2 = 1; m2 < Q; m2++) {
        w.next();
        v[m2] = w.apply(v[m2]);
    }
}

// v[m2] *= sc w_L^(p m2) for 0 <= m2 < Q: a real scale carried by the rotation (seeded with magnitude sc)
template <int Q, bool LP>
__device__ __forceinline__ void twiddle_line_scaled(float2 (&v)[Q], const int p, const int L, const float sc)
{
    Rot<LP> w;
    w.c = sc;
    w.step(p, L);
    v[0] = make_float2(v[0].x * sc, v[0].y * sc);
#pragma unroll
    for (int m2 = 1; m2 < Q; m2++) {
        w.next();
        v[m2] = w.apply(v[m2]);
    }
}

// Cache policy of the intermediate's and the operand's accesses: non-temporal (bit 1 of the buffer instructions'
// cache-policy operand), since each is written or read once by a pass and their footprints (gigabytes per launch)
// exceed the caches.
constexpr int kStream = 2;

// LDS row pitch (floats) per m2: congruent to 16 mod 32, so a wave's 4 rows of 16 lanes fall on distinct banks
constexpr int pitch(int base) { return base + (16 - base % 32 + 32) % 32; }

// Storage of the intermediate Y: complex64, or two fp16 (round to nearest even) of the value times a per-row scale
// (1 / rms of the row's Y, so the half range is never approached and its 11-bit precision applies relative to the
// row's own level; used for the bf16 / fp8 operand forms, whose own rounding is coarser). The column pass applies
// the scale through its output rotation, so put() only converts.
typedef _Float16 h16x2 __attribute__((ext_vector_type(2)));
struct YF32 {
    using E = float2;
    static __device__ __forceinline__ E put(const float2 v) { return v; }
    static __device__ __forceinline__ void store(E* p, const float2 v)
    {
        typedef float f32x2_t __attribute__((ext_vector_type(2)));
        f32x2_t w;
        w.x = v.x;
        w.y = v.y;
        __builtin_nontemporal_store(w, (f32x2_t*)p);
    }
    static __device__ __forceinline__ float2 get(const E e) { return e; }
    static __device__ __forceinline__ float2 load(const __amdgpu_buffer_rsrc_t rs, const unsigned off)
    {
        return load_f2(rs, off);
    }
};
struct YF16 {
    using E = h16x2;
    static __device__ __forceinline__ E put(const float2 v)
    {
        E h;
        h.x = (_Float16)v.x;
        h.y = (_Float16)v.y;
        return h;
    }
    static __device__ __forceinline__ void store(E* p, const float2 v)
    {
        __builtin_nontemporal_store(__builtin_bit_cast(uint32_t, put(v)), (uint32_t*)p);
    }
    static __device__ __forceinline__ float2 get(const E e) { return make_float2((float)e.x, (float)e.y); }
    static __device__ __forceinline__ float2 load(const __amdgpu_buffer_rsrc_t rs, const unsigned off)
    {
        return get(__builtin_bit_cast(E, __builtin_amdgcn_raw_buffer_load_b32(rs, (int)off, 0, kStream)));
    }
};

// Per segment of a launch: its data, the intermediate's scale (per basis row, or null), its analysed span and its first
// row in Zt. Passed by value (kernel arguments) and indexed by the wave-uniform blockIdx.z.
constexpr int kMaxSeg = 16;
struct SegArgs {
    const float2* D[kMaxSeg];
    const float* ysc[kMaxSeg];
    long a0[kMaxSeg];
    long Tn[kMaxSeg];
    long zrow[kMaxSeg];
};

// Column transforms: block (k1 block, j, segment); threads C x P, thread (c, p) = (tid % C, tid / C) in step 1;
// inputs k1 + n1 (p + P q) with q >= NZ are at or above kend (zero, not read)
template <int P, int Q, int C, int NZ, typename YS>
__global__ void __launch_bounds__(C * P) cols_k(const float2* __restrict__ Vc, const long ldv, const SegArgs S,
                                                const long kmin, const long kend, const int n1, const int lgn,
                                                typename YS::E* __restrict__ Y)
{
    constexpr int L = P * Q, ROW = pitch(P * C);
    constexpr bool LP = sizeof(typename YS::E) == 4;     // fp16 intermediate: float32 twiddles
    __shared__ float ex[Q * ROW];
    const int tid = threadIdx.x;
    const int c = tid % C, p = tid / C;
    const int j = blockIdx.y, seg = blockIdx.z;
    const long n = 1L << lgn;
    const float2* __restrict__ D = S.D[seg];
    const float* __restrict__ ysc = S.ysc[seg];
    Y += (size_t)seg * gridDim.y * n;
    const int k1 = blockIdx.x * C + c;
    const float2* __restrict__ vr = Vc + (size_t)j * ldv;
    // per-lane 32-bit offset from wave-uniform bases (k < n <= 2^22)
    const unsigned off = (unsigned)(k1 + n1 * p);

    float2 v[Q];
    {
        // x over [kmin, kend) through buffers of that range (8 (kend - kmin) < 2^25 bytes): loads outside it return
        // zeros; k < kmin maps to an offset past the end
        const long span = kend - kmin;
        const __amdgpu_buffer_rsrc_t ra = buffer_of(vr + kmin, span * 8), rd = buffer_of(D + kmin, span * 8);
        float2 a[NZ], d[NZ];
#pragma unroll
        for (int q = 0; q < NZ; q++) {
            const long kk = (long)n1 * P * q + off - kmin;
            const unsigned o = kk < 0 ? 0x80000000u : (unsigned)kk * 8u;
            a[q] = load_f2(ra, o);
            d[q] = load_f2(rd, o);
        }
#pragma unroll
        for (int q = 0; q < NZ; q++)
            v[q] = make_float2(a[q].x * d[q].x - a[q].y * d[q].y, a[q].x * d[q].y + a[q].y * d[q].x);
    }
    dft_reg<Q, NZ>(v);
    twiddle_line<Q, LP>(v, p, L);
    // exchange: U(c, p, m2) at m2 ROW + p C + c; step 2 takes tasks (c, m2 = tid / C + P i)
    const int mb = tid / C;
    float2 u[Q / P][P];
#pragma unroll
    for (int part = 0; part < 2; part++) {
        if (part) __syncthreads();
#pragma unroll
        for (int m2 = 0; m2 < Q; m2++) ex[m2 * ROW + p * C + c] = part ? v[m2].y : v[m2].x;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < Q / P; i++)
#pragma unroll
            for (int pp = 0; pp < P; pp++) {
                const float x = ex[(mb + P * i) * ROW + pp * C + c];
                if (part) u[i][pp].y = x; else u[i][pp].x = x;
            }
    }
    typename YS::E* __restrict__ yb = Y + (size_t)j * n;
    const float ys = ysc ? ysc[j] : 1.f;
#pragma unroll
    for (int i = 0; i < Q / P; i++) {
        const int m2 = mb + P * i;
        dft_reg<P>(u[i]);
        // w_n^(k1 t2), t2 = m2 + Q m1: Y[j, t2, k1] at the uniform base Q m1 n1 plus the lane offset m2 n1 + k1. The
        // fp16 form's scale rides on the rotation (its seed times ys; the unit steps keep that magnitude).
        Rot<LP> w;
        w.seed((long)k1 * m2, n);
        if constexpr (LP) { w.c *= ys; w.s *= ys; }
        w.step((long)k1 * Q, n);
        const unsigned yo = (unsigned)(m2 * n1 + k1);
#pragma unroll
        for (int m1 = 0; m1 < P; m1++) {
            if (m1) w.next();
            YS::store(&(yb + (size_t)Q * m1 * n1)[yo], w.apply(u[i][m1]));
        }
    }
}

// One (Re, Im) pair of the output type as a single buffer store; stores whose offset is at or past the buffer's
// size are dropped by the hardware range check, so rows outside the analysed span need no branch.
template <typename T>
struct Out;
template <>
struct Out<float> {
    static __device__ __forceinline__ void store(const __amdgpu_buffer_rsrc_t rs, const unsigned off, const float re,
                                                 const float im)
    {
        typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
        u32x2 v;
        v.x = __float_as_uint(re);
        v.y = __float_as_uint(im);
        __builtin_amdgcn_raw_buffer_store_b64(v, rs, (int)off, 0, kStream);
    }
};
template <>
struct Out<uint16_t> {
    static __device__ __forceinline__ void store(const __amdgpu_buffer_rsrc_t rs, const unsigned off, const float re,
                                                 const float im)
    {
        __builtin_amdgcn_raw_buffer_store_b32(f32x2_to_bf16x2(re, im), rs, (int)off, 0, kStream);
    }
};
template <>
struct Out<f16s> {
    static __device__ __forceinline__ void store(const __amdgpu_buffer_rsrc_t rs, const unsigned off, const float re,
                                                 const float im)
    {
        __builtin_amdgcn_raw_buffer_store_b32(f32x2_to_f16x2(re, im), rs, (int)off, 0, kStream);
    }
};
template <>
struct Out<uint8_t> {
    static __device__ __forceinline__ void store(const __amdgpu_buffer_rsrc_t rs, const unsigned off, const float re,
                                                 const float im)
    {
        __builtin_amdgcn_raw_buffer_store_b16(f32x2_to_fp8x2(re, im), rs, (int)off, 0, kStream);
    }
};

// Row transforms: block (row block, t2, segment), so that the blocks running together complete whole output rows;
// threads JB x P, thread (p, jr) = (tid % P, tid / P) in step 1. The row scale 1 / yscale (and, for every output
// type but fp8, the column scale) is applied with the step-1 twiddles; fp8's column scale multiplies the final
// values, so that its bytes are the rounding of the float32 form's output times the scale.
template <int P, int Q, int JB, typename T, typename YS>
__global__ void __launch_bounds__(JB * P) rows_k(const typename YS::E* __restrict__ Y, const int r, const int rp,
                                                 const int lgn1, const int lgn, const SegArgs S,
                                                 const float* __restrict__ cs, T* __restrict__ Zt, const long ldz)
{
    constexpr int L = P * Q, ROW = pitch(P * (JB + 1));
    constexpr bool LP = sizeof(typename YS::E) == 4;     // fp16 intermediate: float32 twiddles
    constexpr bool LATE = sizeof(T) == 1;                // fp8: column scale on the final values
    __shared__ float ex[Q * ROW];
    const int tid = threadIdx.x;
    const int p = tid % P, jr = tid / P;
    const long n = 1L << lgn, n1 = 1L << lgn1, n2 = n >> lgn1;
    const long t2 = blockIdx.y;
    const int j0 = blockIdx.x * JB;
    const int seg = blockIdx.z;
    const long a0 = S.a0[seg], Tn = S.Tn[seg];
    const float* __restrict__ ysc = S.ysc[seg];
    Y += (size_t)seg * r * n;
    Zt += (size_t)S.zrow[seg] * ldz;

    float2 v[Q];
    {
        // Y[j0 + jr, t2, p + P q] from a buffer that ends with Y, so rows j >= r read zeros; byte offsets
        // (jr n + p + P q) E < 2^30
        constexpr unsigned ES = sizeof(typename YS::E);
        const __amdgpu_buffer_rsrc_t rs = buffer_of(Y + (size_t)j0 * n + t2 * n1, ((long)(r - j0) * n - t2 * n1) * ES);
        const unsigned yo = (unsigned)(jr * n + p) * ES;
#pragma unroll
        for (int q = 0; q < Q; q++) v[q] = YS::load(rs, yo + P * q * ES);
    }
    dft_reg<Q>(v);
    {
        const int jj = j0 + jr;
        float sc = 1.f;
        if (jj < r) {
            if (ysc) sc = 1.f / ysc[jj];
            if (!LATE && cs) sc *= cs[jj];
        }
        twiddle_line_scaled<Q, LP>(v, p, L, sc);
    }
    // exchange: U(jr, p, m2) at m2 ROW + p (JB + 1) + jr; step 2 takes tasks (jr2 = tid % JB, m2 = tid / JB + P i)
    const int jr2 = tid % JB, mb = tid / JB;
    float2 u[Q / P][P];
#pragma unroll
    for (int part = 0; part < 2; part++) {
        if (part) __syncthreads();
#pragma unroll
        for (int m2 = 0; m2 < Q; m2++) ex[m2 * ROW + p * (JB + 1) + jr] = part ? v[m2].y : v[m2].x;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < Q / P; i++)
#pragma unroll
            for (int pp = 0; pp < P; pp++) {
                const float x = ex[(mb + P * i) * ROW + pp * (JB + 1) + jr2];
                if (part) u[i][pp].y = x; else u[i][pp].x = x;
            }
    }
    const int j = j0 + jr2;
    if (j >= rp) return;
    const float sc = LATE && j < r && cs ? cs[j] : 1.f;
    // Zt as a buffer of Tn rows: row t - a0 = tl + n2 Q m1; rows outside [0, Tn) go to offset Tn ldz (dropped)
    constexpr unsigned SZ = 2 * sizeof(T);
    const unsigned nbytes = (unsigned)(Tn * ldz * sizeof(T));
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(Zt, 0, (int)nbytes, kRsrcWord3);
    const unsigned rowb = (unsigned)(ldz * sizeof(T));
#pragma unroll
    for (int i = 0; i < Q / P; i++) {
        const int m2 = mb + P * i;
        dft_reg<P>(u[i]);
        const int tl = (int)(t2 - a0 + n2 * m2);           // |tl| < n <= 2^22
#pragma unroll
        for (int m1 = 0; m1 < P; m1++) {
            const int t = tl + (int)n2 * Q * m1;
            const unsigned row = (unsigned)t < (unsigned)Tn ? (unsigned)t : (unsigned)Tn;
            const float2 z = u[i][m1];
            if constexpr (LATE) Out<T>::store(rs, row * rowb + (unsigned)j * SZ, z.x * sc, z.y * sc);
            else Out<T>::store(rs, row * rowb + (unsigned)j * SZ, z.x, z.y);
        }
    }
}

// sum_k |Vc[j, k] D[k]|^2 over [kmin, kend) in double (1 / n of sum_t |z_j[t]|^2 by Parseval); one block per row
__global__ void __launch_bounds__(256) vd_power_k(const float2* __restrict__ Vc, const long ldv,
                                                  const float2* __restrict__ D, const long kmin, const long kend,
                                                  double* __restrict__ out)
{
    const int j = blockIdx.x;
    double s = 0.0;
    for (long k = kmin + threadIdx.x; k < kend; k += 256) {
        const float2 a = Vc[(size_t)j * ldv + k], d = D[k];
        const double m = ((double)a.x * a.x + (double)a.y * a.y) * ((double)d.x * d.x + (double)d.y * d.y);
        s += m;
    }
    s = wave_sum_f64(s);
    __shared__ double sh[4];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[j] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

template <int P, int Q, int NZ, typename YS>
void launch_cols_nz(const float2* Vc, long ldv, const SegArgs& S, int nseg, long kmin, long kend, int n1, int lgn,
                    int r, void* Y, hipStream_t s)
{
    // columns per block: 16 (8 for the 2048-point lines), so that a block is 512 threads and two run on a CU (1,024
    // threads need more than the 64 registers that two such blocks would leave each): 128-byte runs per Y row for
    // complex64, 64-byte runs for fp16 pairs
    constexpr int C = Q > 32 ? 8 : 16;
    hipLaunchKernelGGL((cols_k<P, Q, C, NZ, YS>), dim3((unsigned)(n1 / C), r, nseg), dim3(C * P), 0, s, Vc, ldv, S,
                       kmin, kend, n1, lgn, (typename YS::E*)Y);
}

// the thread of column k1 and lane p holds k = k1 + n1 (p + P q): all k >= kend for q >= NZ when n1 P NZ >= kend
template <int P, int Q, typename YS>
void launch_cols(const float2* Vc, long ldv, const SegArgs& S, int nseg, long kmin, long kend, int n1, int lgn, int r,
                 void* Y, hipStream_t s)
{
    const long span = (long)n1 * P;
    if constexpr (Q >= 32) {
        if (kend <= span * (Q / 32)) {
            launch_cols_nz<P, Q, Q / 32, YS>(Vc, ldv, S, nseg, kmin, kend, n1, lgn, r, Y, s);
            return;
        }
    }
    if constexpr (Q >= 16) {
        if (kend <= span * (Q / 16)) {
            launch_cols_nz<P, Q, Q / 16, YS>(Vc, ldv, S, nseg, kmin, kend, n1, lgn, r, Y, s);
            return;
        }
    }
    if (kend <= span * (Q / 8)) launch_cols_nz<P, Q, Q / 8, YS>(Vc, ldv, S, nseg, kmin, kend, n1, lgn, r, Y, s);
    else if (kend <= span * (Q / 4)) launch_cols_nz<P, Q, Q / 4, YS>(Vc, ldv, S, nseg, kmin, kend, n1, lgn, r, Y, s);
    else if (kend <= span * (Q / 2)) launch_cols_nz<P, Q, Q / 2, YS>(Vc, ldv, S, nseg, kmin, kend, n1, lgn, r, Y, s);
    else launch_cols_nz<P, Q, Q, YS>(Vc, ldv, S, nseg, kmin, kend, n1, lgn, r, Y, s);
}

template <typename YS>
int cols_dispatch(const int lgn2, const float2* Vc, long ldv, const SegArgs& S, int nseg, long kmin, long kend,
                  int n1, int lgn, int r, void* Y, hipStream_t s)
{
    switch (lgn2) {
    case 7: launch_cols<8, 16, YS>(Vc, ldv, S, nseg, kmin, kend, n1, lgn, r, Y, s); break;
    case 8: launch_cols<16, 16, YS>(Vc, ldv, S, nseg, kmin, kend, n1, lgn, r, Y, s); break;
    case 9: launch_cols<16, 32, YS>(Vc, ldv, S, nseg, kmin, kend, n1, lgn, r, Y, s); break;
    case 10: launch_cols<32, 32, YS>(Vc, ldv, S, nseg, kmin, kend, n1, lgn, r, Y, s); break;
    case 11: launch_cols<32, 64, YS>(Vc, ldv, S, nseg, kmin, kend, n1, lgn, r, Y, s); break;
    default: return GWGS_EUNSUP;
    }
    return GWGS_OK;
}

template <int P, int Q, typename T, typename YS>
void launch_rows(const void* Y, int r, int rp, int lgn1, int lgn, const SegArgs& S, int nseg, const float* cs, T* Zt,
                 long ldz, hipStream_t s)
{
    // rows per block: 16 float32 pairs or 32 bf16 / fp8 pairs give 128- / 128- / 64-byte runs per output row
    constexpr int JB = Q > 32 ? 8 : sizeof(T) == 4 ? 16 : 32;
    const long n2 = 1L << (lgn - lgn1);
    hipLaunchKernelGGL((rows_k<P, Q, JB, T, YS>), dim3((rp + JB - 1) / JB, (unsigned)n2, nseg), dim3(JB * P), 0, s,
                       (const typename YS::E*)Y, r, rp, lgn1, lgn, S, cs, Zt, ldz);
}

template <typename T, typename YS>
int rows_dispatch(const int lgn1, const void* Y, int r, int rp, int lgn, const SegArgs& S, int nseg, const float* cs,
                  T* Zt, long ldz, hipStream_t s)
{
    switch (lgn1) {
    case 6: launch_rows<8, 8, T, YS>(Y, r, rp, lgn1, lgn, S, nseg, cs, Zt, ldz, s); break;
    case 7: launch_rows<8, 16, T, YS>(Y, r, rp, lgn1, lgn, S, nseg, cs, Zt, ldz, s); break;
    case 8: launch_rows<16, 16, T, YS>(Y, r, rp, lgn1, lgn, S, nseg, cs, Zt, ldz, s); break;
    case 9: launch_rows<16, 32, T, YS>(Y, r, rp, lgn1, lgn, S, nseg, cs, Zt, ldz, s); break;
    case 10: launch_rows<32, 32, T, YS>(Y, r, rp, lgn1, lgn, S, nseg, cs, Zt, ldz, s); break;
    case 11: launch_rows<32, 64, T, YS>(Y, r, rp, lgn1, lgn, S, nseg, cs, Zt, ldz, s); break;
    default: return GWGS_EUNSUP;
    }
    return GWGS_OK;
}

template <typename YS>
int series(int dtype, int lg, const float2* vc, long ldv, int r, int rp, const SegArgs& S, int nseg, long kmin,
           long kend, const float* cs, void* Y, void* Zt, long ldz, hipStream_t s)
{
    const int lgn1 = lg / 2, lgn2 = lg - lg / 2;
    int e = cols_dispatch<YS>(lgn2, vc, ldv, S, nseg, kmin, kend, 1 << lgn1, lg, r, Y, s);
    if (e) return e;
    if (dtype == 0) return rows_dispatch<float, YS>(lgn1, Y, r, rp, lg, S, nseg, cs, (float*)Zt, ldz, s);
    if (dtype == 1) return rows_dispatch<uint16_t, YS>(lgn1, Y, r, rp, lg, S, nseg, cs, (uint16_t*)Zt, ldz, s);
    if (dtype == 3) return rows_dispatch<f16s, YS>(lgn1, Y, r, rp, lg, S, nseg, cs, (f16s*)Zt, ldz, s);
    return rows_dispatch<uint8_t, YS>(lgn1, Y, r, rp, lg, S, nseg, cs, (uint8_t*)Zt, ldz, s);
}

int log2_exact(const long n)
{
    if (n <= 0 || (n & (n - 1))) return -1;
    int l = 0;
    while ((1L << l) < n) l++;
    return l;
}

} // namespace

// Bytes of the intermediate Y for r basis rows and transform length n; 0 when n is not a supported length
// (a power of two from 2^13 to 2^22).
GWGS_API long gwgs_basis_series_workspace(int r, long n)
{
    const int lg = log2_exact(n);
    if (lg < 13 || lg > 22 || r <= 0) return 0;
    return (long)r * n * 8;
}

// The checks and launches of both entry points: segment i of nseg reads D[i], writes Zt rows zrow[i] .. zrow[i] +
// Tn[i] - 1 and uses its own part of work (the intermediate of r rows each).
static int series_call(int dtype, const void* Vc, long ldv, int r, int rp, int nseg, const void* const* D, long kmin,
                       long kend, long n, const long* a0, const long* Tn, const long* zrow, const float* colscale,
                       int ytype, const float* const* yscale, void* work, void* Zt, long ldz, void* stream)
{
    GWGS_CHECK(Vc && D && a0 && Tn && zrow && work && Zt && r > 0 && rp >= r && kmin >= 0 && kmin <= kend &&
               kend <= ldv && kend <= n && ldz >= 2 * (long)rp && ldz % 2 == 0, GWGS_EINVAL);
    if (nseg < 1 || nseg > kMaxSeg) return GWGS_ESIZE;
    const int lg = log2_exact(n);
    if (lg < 13 || lg > 22) return GWGS_EUNSUP;
    if (dtype < 0 || dtype > 3 || ytype < 0 || ytype > 1) return GWGS_EUNSUP;
    // the row kernel addresses a segment's Zt rows as one buffer with 32-bit byte offsets up to (Tn + 1) rows
    const double esz = dtype == 0 ? 4.0 : dtype == 2 ? 1.0 : 2.0;
    SegArgs S = {};
    for (int i = 0; i < nseg; i++) {
        GWGS_CHECK(D[i] && a0[i] >= 0 && Tn[i] >= 0 && a0[i] + Tn[i] <= n && zrow[i] >= 0, GWGS_EINVAL);
        if ((double)(Tn[i] + 1) * (double)ldz * esz >= 4294967295.0) return GWGS_ESIZE;
        S.D[i] = (const float2*)D[i];
        S.ysc[i] = ytype && yscale ? yscale[i] : nullptr;
        S.a0[i] = a0[i];
        S.Tn[i] = Tn[i];
        S.zrow[i] = zrow[i];
    }
    hipStream_t s = (hipStream_t)stream;
    const int e = ytype == 0 ? series<YF32>(dtype, lg, (const float2*)Vc, ldv, r, rp, S, nseg, kmin, kend, colscale,
                                            work, Zt, ldz, s)
                             : series<YF16>(dtype, lg, (const float2*)Vc, ldv, r, rp, S, nseg, kmin, kend, colscale,
                                            work, Zt, ldz, s);
    if (e) return e;
    return gwgs_status();
}

// dtype 0: float32 Zt, 1: bfloat16, 2: fp8 e4m3 (clamped to +-448), 3: fp16. Vc (>= r rows, row pitch ldv >= kend) and D
// (>= kend) complex64; work: gwgs_basis_series_workspace(r, n) bytes; Zt (Tn rows, ldz >= 2 rp elements, even);
// colscale: per basis row, or null. ytype 0: the intermediate in complex64; 1: in fp16 pairs scaled by yscale[j]
// (r floats, e.g. sqrt(n1 / sum_k |x_jk|^2); null = 1), divided out again in the row pass.
GWGS_API int gwgs_basis_series(int dtype, const void* Vc, long ldv, int r, int rp, const void* D, long kmin, long kend,
                               long n, long a0, long Tn, const float* colscale, int ytype, const float* yscale,
                               void* work, void* Zt, long ldz, void* stream)
{
    const long z0 = 0;
    return series_call(dtype, Vc, ldv, r, rp, 1, &D, kmin, kend, n, &a0, &Tn, &z0, colscale, ytype, &yscale, work,
                       Zt, ldz, stream);
}

// nseg (1..16) segments in one launch of each pass: D[i], a0[i], Tn[i], zrow[i] (Zt row of the segment's first
// sample), yscale[i] (host arrays of device addresses and sizes); work: nseg r n x 8 bytes (ytype 0) or x 4 (ytype
// 1), segment i's intermediate at i r n elements; otherwise as gwgs_basis_series.
GWGS_API int gwgs_basis_series_batch(int dtype, const void* Vc, long ldv, int r, int rp, int nseg, const void* const* D,
                                     long kmin, long kend, long n, const long* a0, const long* Tn, const long* zrow,
                                     const float* colscale, int ytype, const float* const* yscale, void* work, void* Zt,
                                     long ldz, void* stream)
{
    return series_call(dtype, Vc, ldv, r, rp, nseg, D, kmin, kend, n, a0, Tn, zrow, colscale, ytype, yscale, work,
                       Zt, ldz, stream);
}

// out[j] = sum over [kmin, kend) of |Vc[j, k] D[k]|^2 in double, j < r.
GWGS_API int gwgs_vd_power(const void* Vc, long ldv, int r, const void* D, long kmin, long kend, double* out,
                           void* stream)
{
    GWGS_CHECK(Vc && D && out && r > 0 && kmin >= 0 && kmin <= kend && kend <= ldv, GWGS_EINVAL);
    hipLaunchKernelGGL(vd_power_k, dim3(r), dim3(256), 0, (hipStream_t)stream, (const float2*)Vc, ldv,
                       (const float2*)D, kmin, kend, out);
    return gwgs_status();
}
