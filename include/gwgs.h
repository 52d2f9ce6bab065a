/* C ABI of libgwgs_hip.so: hand-written gfx950 (MI355X) kernels for gw-gpu-search's filtering hot path.
 *
 * Conventions. Every pointer argument is a device pointer unless noted; `stream` is a hipStream_t (pass torch's
 * torch.cuda.current_stream().cuda_stream, or 0 for the null stream). Leading dimensions are in elements. Every
 * function only enqueues work on `stream` and allocates nothing. Return value: 0 = ok; negative = argument error
 * (GWGS_EINVAL -1, GWGS_ESIZE -2, GWGS_EUNSUP -3), raised by the Python layer as AssertionError like the
 * reference's asserts; positive = a hipError_t, raised as RuntimeError. Complex arrays are interleaved float2 /
 * double2. The functions are re-entrant; calls on different streams may overlap.
 *
 * Each declaration names the reference interface it replaces (files of DivyanshGolyan/gw-gpu-search).
 */
#ifndef GWGS_H
#define GWGS_H

#ifdef __cplusplus
extern "C" {
#endif

#define GWGS_OK 0
#define GWGS_EINVAL -1
#define GWGS_ESIZE -2
#define GWGS_EUNSUP -3

/* ---- fused SNR-reconstruction GEMM (pipeline/fp_gemm.py:14-44 gemm_absmax / fused_peaks; with mode 1 also the
 * cluster-window max of pipeline/fp_filter.py:745-746) ------------------------------------------------------------
 * z_re[m, t] = sum_k Ar[m, k] Zt[t, k], z_im[m, t] = sum_k Ai[m, k] Zt[t, k], pw = z_re^2 + z_im^2.
 * dtype: 0 = float32 operands, 1 = bfloat16 (bits in uint16), 2 = fp8 e4m3 bytes, 3 = float16 (the reference's
 * operand type, v_mfma_f32_32x32x16_f16; K loops only); fp32 accumulation.
 * rowscale (M floats or NULL) multiplies template m's reported |z|^2.
 * variant: 0 = choose by K (template operands held in registers up to K = 256 bf16 / 128 fp32, else the K loop).
 * Zt: (>= ceil(Tn/128)*128 rows, ldz) time-major; Ar, Ai: (M, lda).
 * mode 0: V, I (M, ldv >= ceil(Tn/128)): per 128 samples the max pw (columns >= Tn masked to -1) and its sample.
 * mode 1: V, I (M, ldv >= Tn/(128 per)): per window of `per` blocks the max and its sample; Tn % (128 per) == 0.
 * mode | 4: the operands are complex-stacked (Ai[m, 2j] = -Ar[m, 2j+1], Ai[m, 2j+1] = Ar[m, 2j], as gwgs_stack_a
 * writes them): the bf16 / fp8 K loops then form z_im from Ar and pair-rotated fragments and do not read Ai.
 * variant (K loop, mode | 4): 0 or 5 = operands loaded by LDS DMA through a 3-stage ring (K a multiple of 32 bf16 /
 * 64 fp8, else as 4), 4 = two LDS stages filled from registers, 1 = template fragments in registers one chunk ahead,
 * 2 = both operands staged through LDS, 3 = template fragments in registers at each chunk. Without mode | 4 every
 * K-loop variant stages Ar and Ai through LDS.
 * First index wins ties at both levels. tiles_per_block: 128-sample tiles per workgroup (0 = default). */
int gwgs_gemm_peaks(int dtype, int variant, const void* Zt, long ldz, const void* Ar, const void* Ai, long lda,
                    int M, int K, int Tn, int mode, int per, int tiles_per_block, float* V, int* I, long ldv,
                    const float* rowscale, void* stream);
/* Several segments in one launch (the batched form): items is a device array of nitems work items, each
 * { int64 zrow0; int32 t_first, t_end, col0, seg, Tn, pad } (gwgs_gemm_item_bytes() = 32 bytes): 128-sample tiles
 * [t_first, t_end) of the segment whose Zt rows start at zrow0; its outputs go to columns col0 + block (mode 0) or
 * col0 + window (mode 1, t_first a multiple of per); rowscale (nseg, M) or NULL multiplies |z|^2 of template m in
 * segment seg (the segment's template norm^2, times fp8 row scales). dtype 2 = fp8 e4m3 (OCP) operands. */
int gwgs_gemm_peaks_multi(int dtype, int variant, const void* Zt, long ldz, const void* Ar, const void* Ai, long lda,
                          int M, int K, const void* items, int nitems, int mode, int per, float* V, int* I, long ldv,
                          const float* rowscale, void* stream);
/* No reference counterpart (the reference keeps one maximum per window, fp_gemm.py:21-33): gwgs_gemm_peaks_multi in
 * mode 1 (stacked: 1 for the rotated-operand stacking of mode bit 4), and per (template, window) also its
 * gwgs_gemm_lm_count() = 4 largest separated local maxima of |z|^2 (p[t] >= p[t - 1] and p[t] > p[t + 1], the
 * window's edges open; of local maxima within sep samples of each other only the larger, the earlier on ties),
 * larger value first, then earlier sample: VL[m ldl + col LM + k] row scaled (-1 where there are fewer), IL their
 * segment-relative samples (-1 likewise); ldl >= LM ldv; 0 <= sep < per x 128. Local maxima below gate x
 * max(window maximum, floor) (0 <= gate <= 1, floor >= 0 on the row-scaled |z|^2) need not be listed: the list holds
 * the largest of those at or above it (a window below the floor may list none). The
 * exact stage evaluates a candidate's window maximum together with the loudest other local maxima, so that two
 * operand precisions report the same sample. */
int gwgs_gemm_peaks_multi_lm(int dtype, int variant, const void* Zt, long ldz, const void* Ar, const void* Ai,
                             long lda, int M, int K, const void* items, int nitems, int stacked, int per, int sep,
                             float gate, float floor, float* V, int* I, long ldv, float* VL, int* IL, long ldl,
                             const float* rowscale, void* stream);
/* No reference counterpart (the reference keeps no local-maximum lists): gwgs_gemm_peaks_multi_lm over an item
 * table whose length is the int32 *nitems_dev in device memory, at most max_items (the grid is sized for max_items;
 * blocks past the count return at once), as gwgs_hot_items writes it, so that a second list pass over the windows
 * holding candidates needs no host round trip. */
int gwgs_gemm_peaks_multi_lm_dyn(int dtype, int variant, const void* Zt, long ldz, const void* Ar, const void* Ai,
                                 long lda, int M, int K, const void* items, const int* nitems_dev, int max_items,
                                 int stacked, int per, int sep, float gate, float floor, float* V, int* I, long ldv,
                                 float* VL, int* IL, long ldl, const float* rowscale, void* stream);
/* No reference counterpart: the windows holding a candidate with val >= cut of a compacted list (idx / val / count
 * as gwgs_select_peaks wrote them, capacity cap, idx a row-major index into (M, nwsum) window maxima) as one-window
 * GEMM items (GemmItem rows, any order; their count to *nitems). segs: nseg rows of {int64 first Zt row, int32 first
 * window column, int32 analysed samples}, columns ascending; per: 128-sample tiles per window; flags: nwsum int32
 * of workspace; items: room for nwsum rows (8-byte aligned). */
int gwgs_hot_items(const int* idx, const float* val, const int* count, int cap, int nwsum, float cut,
                   const void* segs, int nseg, int per, int* flags, void* items, int* nitems, void* stream);
int gwgs_gemm_lm_count(void);
int gwgs_gemm_item_bytes(void);
/* largest K the GEMM accepts for dtype (0 if the dtype is unsupported) */
int gwgs_gemm_max_k(int dtype);

/* ---- peaks and candidates --------------------------------------------------------------------------------------
 * pipeline/peak_kernel.py:8-47 win_peak / window_peaks: z (rows, ldz) complex64; per (row, window) the max |z|^2
 * (float32, no fused multiply-add) over z[row, a0 + w cw : a0 + (w+1) cw] and its index in the window. */
int gwgs_window_peaks(const void* z, long ldz, int rows, int a0, int cw, int nwin, float* pmax, int* pidx,
                      void* stream);
/* pipeline/fp_filter.py:745-746: vw[r, w] = max_j V[r, w per + j], iw = I at that j (first j on ties). */
int gwgs_block_window_max(const float* V, const int* I, long ldv, int n, int nw, int per, float* vw, int* iw,
                          void* stream);
/* pipeline/fp_filter.py:747 (torch.nonzero(vw > cut)): row-major indices g with vw[g] > cut, in order, into
 * out[0:cap]; the count to *count_dev (device int). work: gwgs_select_workspace(total) device ints (host call). */
long gwgs_select_workspace(long total);
int gwgs_select_above(const float* vw, long total, float cut, int* out, int cap, int* work, int* count_dev,
                      void* stream);
/* pipeline/fp_filter.py:747-757 in one pass: as gwgs_select_above, plus out_iw[i] = iw[out[i]] (the window's sample)
 * and out_vw[i] = vw[out[i]] for the selected entries, so the candidate list needs no host round trip. */
int gwgs_select_peaks(const float* vw, const int* iw, long total, float cut, int* out, int* out_iw, float* out_vw,
                      int cap, int* work, int* count_dev, void* stream);
/* No counterpart in the reference: gwgs_select_peaks for callers that cannot allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc). The same compaction with its block counts in a static device
 * array of 65,536 ints (256 KiB per GPU, held until the process ends) instead of a caller workspace: total must be
 * at most 65,536 x 4,096 = 268,435,456 entries, else GWGS_ESIZE. Allocates nothing. Calls that overlap on one GPU
 * share the array and race; calls on one stream, or ordered by events, do not. */
int gwgs_select_peaks_noalloc(const float* vw, const int* iw, long total, float cut, int* out, int* out_iw,
                              float* out_vw, int cap, int* count_dev, void* stream);
/* No reference counterpart (the reference evaluates the exact stage at the rebuilt argmax only, fp_filter.py:759-
 * 783): for every candidate of a compacted list (idx / samp / val / count as gwgs_select_peaks wrote them, capacity
 * cap) whose val >= exact_cut, the entries of its window's local-maximum list (VL / IL / ldl / lm as
 * gwgs_gemm_peaks_multi_lm wrote them, nwsum window columns) farther than R samples from samp[c] with |z|^2 >=
 * ratio val[c], in list order, up to na: out[c na + a] = the entry's offset in its window of cw samples, -1 past
 * the last. */
int gwgs_window_alternates(const float* VL, const int* IL, long ldl, int lm, const int* idx, const int* samp,
                           const float* val, const int* count, int cap, int nwsum, int cw, float exact_cut,
                           float ratio, int R, int na, int* out, void* stream);
/* pipeline/fp_filter.py:759-764: rebuilt complex SNR at candidates, out[c] = (Ar[rows[c]] . Zt[tcol[c]],
 * Ai[rows[c]] . Zt[tcol[c]]) sqrt(pw_scale[sidx[c]]) in fp32 (complex64 out); sidx NULL indexes pw_scale by rows[c];
 * pw_scale NULL: no scale. dtype as gwgs_gemm_peaks. */
int gwgs_rebuilt_z(int dtype, const void* Zt, long ldz, const void* Ar, const void* Ai, long lda, int K,
                   const int* rows, const int* tcol, const int* sidx, int n, const float* pw_scale, void* out,
                   void* stream);

/* ---- exact values at peaks (pipeline/chisq_kernel.py:14-121 chisq_pts / chisq_pts_fused / bin_power_sums;
 * the dense exact stage of pipeline/exact_stage.py:196-203) ------------------------------------------------------
 * Group g: row grow[g] of q (or of the templates h when d != NULL: summand conj(h_k) d_k), peaks
 * gstart[g] .. gstart[g] + gcount[g] - 1 (gcount <= gwgs_chisq_peaks_per_group()) at samples ptime[],
 * bin edges edges[g, 0:nedge]. out[p] = sum_j |Z_j|^2, zout[p] = sum_j Z_j (may be NULL), with
 * Z_j = sum_{e_j <= k < e_j+1} v_k exp(2 pi i ((k t) mod N) / N). Any N > 0 (64-bit integer phase). One
 * workgroup per (group, bin); the per-bin sums go to `work`, then a second pass adds them in bin order. */
int gwgs_chisq_points(const void* q, long ldq, const void* d, long long N, const int* grow, const int* gstart,
                      const int* gcount, const long long* ptime, const int* edges, int nedge, int ngroups, float* out,
                      void* zout, void* work, void* stream);
/* bytes of the device workspace gwgs_chisq_points needs (the per-bin sums of every group; 0 when there is none) */
long gwgs_chisq_points_workspace(int ngroups, int nedge);
int gwgs_chisq_peaks_per_group(void);
/* pipeline/exact_stage.py:472-510 (edges_moments, the regen path's chi-squared edges): edges of n (segment, row)
 * requests of interpolated templates without reading H, from each segment's moments of 1/psd over each interval
 * (Mi_tab[s] (nx - 1, 3): sums of ((1-t)^2, t(1-t), t^2) IP) and their in-interval inclusive prefixes at each bin
 * (Pk_tab[s] (kend - kmin, 3)), float64 (device tables of pointers; rseg[u] the request's segment), and the rows'
 * node amplitudes amp (rows, nx) float64: edge j = the first bin whose running sum of |h|^2 IP exceeds j / nb of
 * the total, kb (nx int32) the first bin of each interval (kb[nx - 1] = kend); columns >= nb hold the row's end
 * k1s[u] (<= kend; the row is zero from there). out (n, nedge) int32. */
int gwgs_edges_moments(const double* amp, int nx, const void* const* Mi_tab, const void* const* Pk_tab,
                       const int* rseg, const int* kb, const int* rows, const int* nbs, const int* k1s, int nedge,
                       int kmin, int n, int* out, void* stream);
/* pipeline/exact_stage.py:512-647 (regen_bins + finalize): gwgs_chisq_points_multi with each template row
 * regenerated from its nodes instead of read from H: coef32 (rows, nx - 1) float4 (c3, c2, c1, c0 with c0 reduced
 * to [-pi, pi)), amp2 (rows, nx - 1) float2 (A_j, A_j+1), the grid's node (int32), d and t (float32) per bin of
 * [kmin, kend); h_k = (A_j (1 - t) + A_j+1 t) exp(i p_j(d)) in float32, the phase reduced by 2 pi before the
 * sincos. Every edge within [kmin, kend]. Workspace: gwgs_chisq_points_workspace(ngroups, nedge) bytes. */
int gwgs_regen_points_multi(const void* coef32, const void* amp2, int nx, const int* node, const float* d,
                            const float* t, int kmin, const void* const* d_tab, const int* gseg, long long N,
                            const int* grow, const int* gstart, const int* gcount, const long long* ptime,
                            const int* edges, int nedge, int ngroups, float* out, void* zout, void* work,
                            void* stream);
/* No reference counterpart (the reference evaluates the exact stage at the rebuilt argmax, exact_stage.py:704-731):
 * hill climbing of n peaks on the exact |z| = |sum over e0 <= k < eB of conj(h_k) d_k exp(2 pi i k t / N)|, e0 and eB
 * the first and last entries of edge row eidx[r] (edges: rows of nedge int32): from start[r], z at the 2 W + 1 samples
 * around c clipped to [lo[r], hi[r] - 1]; while the largest |z| (first on ties) lies on the range's edge inside the
 * bounds the range re-centres on it, at most max_iter times; out[r] = the final largest sample. rows, segs int32:
 * the template row of h (ldh complex64 per row) and the data row d_tab[segs[r]]; 1 <= W <= 3. One workgroup per
 * peak. The exact stage uses it for peaks whose re-centring range ended on its edge. */
int gwgs_climb_points(const void* h, long ldh, const void* const* d_tab, long long N, const int* rows, const int* segs,
                      const int* edges, int nedge, const int* eidx, const long long* start, const long long* lo,
                      const long long* hi, int W, int max_iter, int n, long long* out, void* stream);
/* pipeline/exact_stage.py:173-191 ExactStage.evaluate for the peaks of several segments of one tile in one launch
 * (the reference evaluates segment by segment): gwgs_chisq_points in fused mode where group g reads its data row
 * from d_tab[gseg[g]] (device table of complex64 row addresses). Same workspace as gwgs_chisq_points. */
int gwgs_chisq_points_multi(const void* h, long ldh, const void* const* d_tab, const int* gseg, long long N,
                            const int* grow, const int* gstart, const int* gcount, const long long* ptime,
                            const int* edges, int nedge, int ngroups, float* out, void* zout, void* work,
                            void* stream);
/* pipeline/chisq_edges.py:90-143 edges_fused (and edges_blocked, multi_filter.py:264-282 edges_per_template):
 * per request i, row rows[i] of h: edge j = first k with float64 cumsum(|h|^2 ips)[k] > j / nbs[i] * total over
 * [0, k1s[i]), clipped to [kmin, k1s[i]]; columns j >= nbs[i] and the last hold k1s[i]. out (n, nedge) int32. */
int gwgs_edges_fused(const void* h, long ldh, const int* rows, const float* ips, const int* k1s, const int* nbs,
                     int nedge, int kmin, int n, int* out, void* stream);
/* pipeline/chisq_edges.py:90-143 edges_fused over requests of several segments: request i reads 1/psd from
 * ip_tab[rseg[i]] (device table of float32 row addresses); otherwise as gwgs_edges_fused. */
int gwgs_edges_fused_multi(const void* h, long ldh, const int* rows, const float* const* ip_tab, const int* rseg,
                           const int* k1s, const int* nbs, int nedge, int kmin, int n, int* out, void* stream);

/* ---- basis series and GEMM operands (pipeline/fp_filter.py:597-621 basis_series, stacked_A) --------------------
 * q[j, k] = Vc[j, k] D[k] for k < flen, j < rp (q: (rp, ldq) complex64; columns >= flen are not written). */
int gwgs_form_q(const void* Vc, int rp, long flen, const void* D, long ldq, void* q, void* stream);
/* Zt[t, 2j] = Re Zc[j, a0 + t] cs[j], Zt[t, 2j+1] = Im Zc[j, a0 + t] cs[j] for t < Tn, zero for Tn <= t < Tpad;
 * Zc (rp, ldc) complex64, rp % 16 == 0; Zt (Tpad, 2 rp) float32 (dtype 0), bfloat16 (1), fp8 e4m3 bytes (2,
 * saturated to +-448) or float16 (3); cs may be NULL. */
int gwgs_pack_zt(int dtype, const void* Zc, long ldc, int rp, long a0, long Tn, long Tpad, const float* colscale,
                 void* Zt, void* stream);
/* As = A[m, j] rowscale[m] colscale[j] (A (n, r) complex64); Ar[m, 2j] = Re As, Ar[m, 2j+1] = -Im As,
 * Ai[m, 2j] = Im As, Ai[m, 2j+1] = Re As; zero for j >= r, m >= n. Ar, Ai (Mpad, K); dtype 0, 1 or 3. */
int gwgs_stack_a(int dtype, const void* A, int n, int r, const float* rowscale, const float* colscale, int Mpad,
                 int K, void* Ar, void* Ai, void* stream);
/* fp8 operands: per template row the scale ra = 224 / max |As| (As as gwgs_stack_a), Ar/Ai = e4m3(As ra) bytes
 * (n, K), pw_scale[m] = 1 / ra^2 (to multiply |z|^2). */
int gwgs_stack_a_fp8(const void* A, int n, int r, const float* rowscale, const float* colscale, int K, void* Ar,
                     void* Ai, float* pw_scale, void* stream);
/* pipeline/fp_filter.py:615-621 stacked_A with fp16 operands as the reference's (whose single global scale is here a
 * per-template ra = 1024 / max |As|): Ar/Ai = fp16(As ra), pw_scale[m] = 1 / ra^2. */
int gwgs_stack_a_f16(const void* A, int n, int r, const float* rowscale, const float* colscale, int K, void* Ar,
                     void* Ai, float* pw_scale, void* stream);
/* out[j] = sum_{a0 <= t < a0 + Tn} |Zc[j, t]|^2 (double), the rms scale of pipeline/fp_filter.py:609-610 */
int gwgs_basis_power(const void* Zc, long ldc, int rp, long a0, long Tn, double* out, void* stream);

/* Fused basis series (replaces pipeline/fp_filter.py:597-613 basis_series: the zero-padded q = conj(Vh) D / sw,
 * torch.fft.ifft(q) * n and the fp16 stacked Z), as a four-step FFT of two passes for power-of-two n from 2^13 to
 * 2^22: x_j[k] = Vc[j, k] D[k] on [kmin, kend) (0 elsewhere), z_j[t] = sum_k x_j[k] e^(2 pi i k t / n), written as
 * Zt[t - a0, 2j] = Re z_j[t] cs[j], Zt[t - a0, 2j + 1] = Im z_j[t] cs[j] for a0 <= t < a0 + Tn, j < r, and zeros
 * for r <= j < rp. dtype 0 float32, 1 bfloat16 (nearest even), 2 fp8 e4m3 (saturated to +-448), 3 float16
 * (nearest even); colscale may be
 * NULL. Vc: >= r rows of pitch ldv complex64; D complex64; work: gwgs_basis_series_workspace(r, n) bytes; Zt: Tn
 * rows of ldz >= 2 rp elements (ldz even). Returns GWGS_EUNSUP for other n, GWGS_ESIZE when Zt exceeds 4 GiB. */
long gwgs_basis_series_workspace(int r, long n);
/* ytype 0 keeps the intermediate in complex64 (float64 twiddles); 1 in fp16 pairs of the value times yscale[j]
 * (r floats; NULL = 1), divided out in the second pass (half the intermediate's traffic; float32 twiddles). The
 * filter passes ytype 1 for its bf16 / fp8 operands with one yscale value per segment, 1 / (64 rms), and ytype 0
 * for fp32. */
int gwgs_basis_series(int dtype, const void* Vc, long ldv, int r, int rp, const void* D, long kmin, long kend, long n,
                      long a0, long Tn, const float* colscale, int ytype, const float* yscale, void* work, void* Zt,
                      long ldz, void* stream);
/* pipeline/fp_filter.py:597-613 basis_series for every segment of one (tile, batch) in one launch of each pass:
 * nseg (1..16) segments, segment i with data D[i], analysed span a0[i] .. a0[i] + Tn[i] - 1 written to Zt rows
 * zrow[i] .. zrow[i] + Tn[i] - 1, and its own yscale[i] (ytype 1); D and yscale are host arrays of device addresses,
 * a0, Tn and zrow host arrays. work: nseg r n bytes times 8 (ytype 0) or 4 (ytype 1). Otherwise as
 * gwgs_basis_series; GWGS_ESIZE for nseg outside 1..16. */
int gwgs_basis_series_batch(int dtype, const void* Vc, long ldv, int r, int rp, int nseg, const void* const* D,
                            long kmin, long kend, long n, const long* a0, const long* Tn, const long* zrow,
                            const float* colscale, int ytype, const float* const* yscale, void* work, void* Zt,
                            long ldz, void* stream);
/* out[j] = sum over kmin <= k < kend of |Vc[j, k] D[k]|^2 (double), j < r: the transform input's power by rows, from
 * which pipeline/fp_filter.py:609-610's rms scale follows by Parseval without forming q */
int gwgs_vd_power(const void* Vc, long ldv, int r, const void* D, long kmin, long kend, double* out, void* stream);

/* ---- templates (the closed-form chirps standing in for pipeline/imr_interp.py's interpolated waveforms) and
 * sigma^2 (pipeline/fp_filter.py:710-714) ------------------------------------------------------------------------
 * H[i, k] = f^(-7/6) exp(-i psi_i(f)), f = k df, for int(flow/df) <= k < kmax_i = min(int(f_isco,i / df), flen-1),
 * psi = TaylorF2 to 1PN + the leading eccentric term at reference frequency f0. m1, m2, e0: device double arrays
 * (e0 may be NULL); kmax: device int array (may be NULL). */
int gwgs_spa_templates(const double* m1, const double* m2, const double* e0, int n, long flen, double df,
                       double flow, double f0, void* H, long ldh, int* kmax, void* stream);
/* pipeline/imr_interp.py:18-40 (coarse_generate, the templates at coarse nodes for --templates interp): the same
 * chirps at the nx ascending nodes x (device float64): amp (n, nx) = f^(-7/6) below f_isco, 0 above; coef (n, nx - 1,
 * 4) = (c3, c2, c1, c0) of each interval's phase cubic in d = f - x_j, the Hermite cubic through -psi and its
 * analytic derivative at both ends (the reference fits a spline to PyCBC's phase at the nodes). float64. */
int gwgs_spa_nodes(const double* m1, const double* m2, const double* e0, int n, const double* x, int nx, double f0,
                   double* amp, double* coef, void* stream);
/* pipeline/imr_interp.py:43-85 (FineInterpolator): dense rows (n, flen) complex64 on [kmin, kend) from nodes (amp,
 * coef as gwgs_spa_nodes writes them) and the fine-bin grid (node int32, d and t float64 per bin of [kmin, kend):
 * its interval, f - x_j and the fraction (f - x_j) / (x_j+1 - x_j)): h = (A_j (1 - t) + A_j+1 t) exp(i p_j(d)) in
 * double; zero outside [kmin, kend). */
int gwgs_interp_templates(const double* amp, const double* coef, int n, int nx, const int* node, const double* d,
                          const double* t, long kmin, long kend, long flen, void* H, long ldh, void* stream);
/* sig[i, s] = 4 df sum_{kmin <= k < kend} |H[i, k]|^2 IP[s, k] (double accumulation, chunk partials summed in a
 * fixed order), s < nseg <= 16; work: gwgs_sigmasq_workspace(n, kmin, kend, nseg) doubles */
long gwgs_sigmasq_workspace(int n, long kmin, long kend, int nseg);
int gwgs_sigmasq(const void* H, long ldh, long kmin, long kend, int n, const float* IP, long ldip, int nseg,
                 double df, double* work, double* sig, long lds, void* stream);

/* ---- data conditioning (pipeline/gpu_prep.py) ------------------------------------------------------------------
 * welch (gpu_prep.py:87-125): segs[r, s, i] = x[r, start + s stride + i] * hann_L(i) (float64 out) */
int gwgs_welch_segments(const float* x, long ldx, int rows, long start, int nseg, int L, int stride, double* segs,
                        void* stream);
/* welch reduction: per (row, bin) of X (rows, nseg, nf) complex128: p = |X|^2 dt^2 (halved at DC and Nyquist),
 * median (numpy's) / bias or mean, times fac; nseg <= 512 */
int gwgs_welch_reduce(const void* X, int rows, int nseg, int nf, double dt, int mean, double bias, double fac,
                      double* out, void* stream);
/* interpolate (gpu_prep.py:128-135): np.interp of P (rows, nold) at k df_old onto i df_new, i < nnew */
int gwgs_interp(const double* P, int rows, int nold, double df_old, double df_new, long nnew, double* out,
                void* stream);
/* inverse_spectrum_truncation (gpu_prep.py:138-158) passes: P^-1/2 on [kmin, khi) as complex128 */
int gwgs_isqrt_band(const double* P, int rows, long flen, long kmin, long khi, void* out, void* stream);
/* q *= df; zero [L/2, N - L/2), Hann taper of the kept ends when hann != 0 */
int gwgs_trunc(double* q, int rows, long N, long L, int hann, double df, void* stream);
/* out = 1 / |X scale|^2 */
int gwgs_inv_abs2(const void* X, int rows, long flen, double scale, double* out, void* stream);
/* gpu_stage2 output (gpu_prep.py:448-451): D[s, k] = (X[s, k] inv_sr) / P[pick[s], k] (complex64), IP = 1/P as
 * float32, both 0 where P <= 0 or k < kmin */
int gwgs_whiten(const void* X, float inv_sr, const double* P, const int* pick, int nseg, long flen, long kmin,
                void* D, float* IP, void* stream);
/* detect_loud_glitches (gpu_prep.py:175-204): taper + centre in a zero buffer; whiten; |.| over the data */
int gwgs_gate_pad(const float* s, long n, long cl, long pad_start, long Lp, double* out, void* stream);
int gwgs_gate_whiten(void* st, const double* psd, long flen, long kmin, double dt, double norm, void* stream);
int gwgs_gate_mag(const double* wt, long pad_start, long n, long cl, double df, float* mag, void* stream);
/* gate_taper (gpu_prep.py:212-231): x[offset + i] *= inverted Tukey window(i), i in [i1, i2) */
int gwgs_gate_apply(float* x, long n, long offset, long i1, long i2, long pad, long win, void* stream);
/* resample_ldas (gpu_prep.py:67-84, pycbc's 'ldas' resampling): out[i] = sum_k c[k] x[factor i + ceil(ntaps / 2) - k]
 * in float64 without fused multiply-add, zero where that index is below 2 (ntaps / 2) or at or past L;
 * ceil(L / factor) outputs. */
int gwgs_fir_decimate(const double* x, long L, const double* c, int ntaps, int factor, double* out, void* stream);

/* ---- differentiable refinement (no reference function: the backward kernel of the peak SNR; its forward value
 * is pipeline/fp_filter.py:759-764's rebuilt SNR, evaluated off the sample grid and off the bank) -------------------
 * E[b, i] = D[k0 + i] e^{2 pi i (k0 + i)(t[b] + delta[b]) / n}, Ed = (2 pi i (k0 + i) / n) E (npk, lde) complex64:
 * the data phased to each peak's time; R = E Vc^T is the basis response (a plain GEMM). */
int gwgs_phasor_data(const void* D, long k0, long kb, const long long* t, const double* delta, long long n, int npk,
                     void* E, void* Ed, long lde, void* stream);
/* Per peak (one wave): rho^2 = |z|^2 (4 df)^2 / sigma^2 with z = a . R, sigma^2 = 4 df a^T G conj(a); z; the delta
 * derivative of rho^2; the phi derivative of Re(e^{-i phi} z)^2 (4 df)^2 / sigma^2 (phi NULL: phi = arg z); and
 * (grad != NULL) d rho^2 / d a as complex gradients d/dRe + i d/dIm. Double accumulation. */
int gwgs_snr_grad(const void* R, const void* Rd, long ldr, const void* a, long lda, const void* G, long ldg, int r,
                  int npk, double df, const float* phi, float* rho2, void* z, float* drho2_dt, float* drho2_dphi,
                  void* grad, long ldgr, void* stream);

/* ---- time slides (pipeline/coinc_rank.py:100-138 find_pairs) -------------------------------------------------
 * L triggers sorted by key lkey = tid * big + t (with copies at t - P and t + P), lt their unwrapped times, li their
 * indices; per H trigger (key hkey, time hx) the pairs with |tH - tL - k step| <= window, |k| <= K, within
 * [hkey - reach, hkey + reach]. count: counts (nh + 1 int64 workspace) become exclusive offsets, *total the number
 * of pairs; write: (H index, L index, k, dt = k step - (tH - tL)) at those offsets. */
int gwgs_slide_count(const double* hkey, const double* hx, int nh, const double* lkey, const double* lt,
                     const int* li, long nl, double reach, double step, int K, double window, long long* counts,
                     long long* total, void* stream);
int gwgs_slide_write(const double* hkey, const double* hx, int nh, const double* lkey, const double* lt,
                     const int* li, long nl, double reach, double step, int K, double window,
                     const long long* offsets, int* oh, int* ol, int* ok, double* odt, void* stream);

/* ---- pre-consolidation of the candidates (pipeline/fp_filter.py:665-695 pre_keep, the keep-loudest consolidation
 * of fp_filter.py:629-663 applied to one tile and segment's candidates on the device) ----------------------------
 * vw (n, nwsum) float32, iw (n, nwsum) int32: one batch's window maxima and window samples as the GEMM's epilogue
 * writes them. Every candidate vw[g] > cut that the consolidation drops is overwritten with -1 in place, so the
 * compaction that follows (gwgs_select_peaks) skips it. segi: nseg rows of 10 int64 (host-built, device copy):
 * sbase (segment start + first analysed sample, in samples of its dataset), wt0 and nwin (its keep-loudest windows
 * [wt0, wt0 + nwin), wt = trunc(rint((t - gps) sr) / wdiv)), col0 and ncol (its columns of vw), wbase (its first
 * window in the batch's nwin windows), gate0 and ngate (its dataset's gates in `gates`, pairs (time, w + taper +
 * 0.5) in float64), fl0 and nfl (its dataset's floor in `floors`, nfl = 0: none). segd: nseg float64 dataset start
 * times. Kept: apx = sqrt(vw) >= exact_cut; rebuilt candidates with apx > thr, (apx >= nthr or exempt), outside
 * every gate and, with flags & 2 (an interval), apx > floor[wt] and among the num loudest of their segment's window
 * (apx descending, then row-major order); exempt (flags & 1, apx >= exempt_snr) ones are not ranked. work:
 * gwgs_pre_keep_workspace(n * nwsum, nwin) bytes, 8-byte aligned (only with flags & 2). */
long gwgs_pre_keep_workspace(long total, long nwin);
int gwgs_pre_keep(float* vw, const int* iw, int n, int nwsum, int nseg, const long* segi, const double* segd,
                  const double* gates, const double* floors, double sr, double wdiv, int num, float cut,
                  float exact_cut, float thr, float nthr, float exempt_snr, int flags, int nwin, void* work,
                  long work_bytes, void* stream);

/* ---- sine-Gaussian chi-squared (pipeline/sg_kernel.py sg_tiles / sg_chisq_points; pipeline/fp_filter.py:463-486
 * sg_chisq_one) ---------------------------------------------------------------------------------------------------
 * For every (peak p, tile j), b = p * ntile + j, over bins k in [kk[b], k1[b]) with k >= kg[b] (k1 <= flen):
 * zsum[b] = sum g_k D_k e^{2 pi i (k t[p] mod N) / N} (double2) and gsum[b] = sum g_k^2 IP_k, with
 * g_k = (sqrt(pi) / 2 tau) exp(-(pi tau (k df - fcen))^2) (1 + exp(-q2 k df / fcen)) (per-b fcen, tau, q2 in
 * float64). D complex64 and IP float32 of one segment (flen bins, time length N). The caller finishes
 * z_j = 4 df zsum / sqrt(4 df gsum) and the value sum_j |z_j|^2 / (2 ntile). */
int gwgs_sg_tiles(const void* D, const float* IP, long flen, long N, double df, const long* kk, const long* k1,
                  const long* kg, const double* fcen, const double* tau, const double* q2, const long* t, int npk,
                  int ntile, void* zsum, double* gsum, void* stream);

#ifdef __cplusplus
}
#endif
#endif
