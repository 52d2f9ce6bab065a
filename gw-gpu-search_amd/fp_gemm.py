"""Fused MFMA GEMM + |z|^2 + max per BN output samples, for fp_filter.py (the reference's fp_gemm.py:1-64).

z_re = Are @ Z, z_im = Aim @ Z; each launch keeps only the max |z|^2 and its sample per BN = 128 samples (mode 0,
fused_peaks) or per cluster window of cw samples (mode 1, fused_window_peaks, the window max of fp_filter.py:745-746
done in the epilogue). The kernel is hand-written HIP for gfx950 (csrc/gemm_peaks.hip): fp32 operands on
v_mfma_f32_32x32x2_f32 (the plain form), bf16 operands on v_mfma_f32_32x32x16_bf16, fp16 operands (the
reference's type) on v_mfma_f32_32x32x16_f16, or fp8 e4m3 operands on
v_mfma_scale_f32_32x32x64_f8f6f4 (unit block scales; per-template factors applied to |z|^2 in the epilogue), fp32
accumulation.

Layout: the kernel reads the basis series time-major, Zt (T, K) with K contiguous. fused_peaks also accepts the
reference's (K, Tn) layout and transposes it first (one extra pass); fp_filter hands Zt over directly.
"""
from __future__ import annotations

import time

import numpy as np
import torch

import _dev
import _lib

BN = 128
DTYPES = {torch.float32: 0, torch.bfloat16: 1, torch.float8_e4m3fn: 2, torch.float16: 3}


def _kpad(K):
    return -(-K // 16) * 16


def _prep_operands(Are, Aim, Z, Tn, zt):
    """Pad K to a multiple of 16 and rows of Zt to a multiple of BN; return (Zt, Ar, Ai, K, dtype code). float8_e4m3fn
    operands are passed to the kernel as their bytes."""
    dt = Are.dtype
    assert dt in DTYPES, f'unsupported operand dtype {Are.dtype}'
    K = Are.shape[1]
    Kp = _kpad(K)
    Tpad = -(-Tn // BN) * BN
    if zt:
        assert Z.shape[0] >= Tn and Z.shape[1] == K
        Zt = Z
    else:
        assert Z.shape[0] == K and Z.shape[1] >= Tn
        Zt = Z[:, :Tn].t()
    if Kp != K or Zt.shape[0] < Tpad or Zt.dtype != dt or not Zt.is_contiguous():
        Zn = torch.zeros((Tpad, Kp), dtype=dt, device=Z.device)
        Zn[:Tn, :K] = Zt[:Tn]
        Zt = Zn
    if Kp != K or Are.dtype != dt or not Are.is_contiguous():
        A2 = torch.zeros((Are.shape[0], Kp), dtype=dt, device=Are.device); A2[:, :K] = Are; Are = A2
        A3 = torch.zeros((Aim.shape[0], Kp), dtype=dt, device=Aim.device); A3[:, :K] = Aim; Aim = A3
    code = DTYPES[dt]
    if code == 2:
        Zt, Are, Aim = Zt.view(torch.uint8), Are.view(torch.uint8), Aim.contiguous().view(torch.uint8)
    return Zt, Are, Aim.contiguous(), Kp, code


def _cfg(cfg):
    """cfg: None, an int (time tiles per block) or a tuple (tiles_per_block[, variant]). variant 0 picks by K
    (register-resident templates up to K = 256 bf16 / 128 fp32, else the K loop); 1 forces the K loop (bf16 / fp8:
    template fragments loaded into registers one chunk ahead); 2 the K loop with both operands staged through LDS;
    3 the register-fragment K loop without the one-chunk-ahead buffer; 4 (complex-stacked operands) the two-stage
    LDS ring with the imaginary-part template fragments rotated in registers; 5 (complex-stacked operands, also what
    0 selects for them in the K loop) the three-stage ring filled by LDS DMA."""
    if cfg is None:
        return 0, 0
    if isinstance(cfg, (tuple, list)):
        return int(cfg[0]), int(cfg[1]) if len(cfg) > 1 else 0
    return int(cfg), 0


def gemm_peaks_raw(dtype_code, Zt, ldz, Ar, Ai, lda, M, K, Tn, mode, per, cfg, V, I, zt_offset_bytes=0,
                   rowscale=None):
    """Direct C-ABI call on pre-laid-out operands (fp_filter's hot loop). dtype_code 0 fp32, 1 bf16, 2 fp8 e4m3
    (torch.uint8 bytes); rowscale: per-template factor on |z|^2 (fp8 row scales)."""
    tpb, variant = _cfg(cfg)
    _lib.call('gwgs_gemm_peaks', dtype_code, variant, Zt.data_ptr() + zt_offset_bytes, ldz, Ar.data_ptr(),
              Ai.data_ptr(), lda, M, K, Tn, mode, per, tpb, V.data_ptr(), I.data_ptr(), V.stride(0),
              _dev.ptr(rowscale), _dev.stream())


# mode flag: the operands are complex-stacked (Ai[m, 2j] = -Ar[m, 2j+1], Ai[m, 2j+1] = Ar[m, 2j], as stack_a
# writes them), so the bf16 / fp8 K loops form the imaginary part from rotated Zt fragments and never read Ai
STACKED = 4

ITEM = np.dtype([('zrow0', '<i8'), ('t_first', '<i4'), ('t_end', '<i4'), ('col0', '<i4'), ('seg', '<i4'),
                 ('Tn', '<i4'), ('pad', '<i4')])


def tiles_per_item(cfg, per):
    """128-sample tiles per GEMM work item: cfg's tiles_per_block rounded up to whole windows (default 2 windows)."""
    tpb, _ = _cfg(cfg)
    tpb = tpb or 2 * per
    return -(-tpb // per) * per


def items_table(segments, tpb):
    """Device item table for gemm_peaks_multi. segments: (zrow0, ntile, col0, seg, Tn) per segment; each is cut into
    runs of tpb tiles (csrc/gemm_peaks.hip GemmItem)."""
    rows = []
    for zrow0, ntile, col0, seg, Tn in segments:
        for t in range(0, ntile, tpb):
            rows.append((zrow0, t, min(t + tpb, ntile), col0, seg, Tn, 0))
    tab = np.array(rows, dtype=ITEM)
    return _dev.to_dev(tab.view(np.uint8)), len(rows)


def gemm_peaks_multi(dtype_code, Zt, ldz, Ar, Ai, lda, M, K, items, mode, per, cfg, V, I, rowscale=None):
    """One launch over several segments (item table from items_table); rowscale (nseg, M) or None."""
    _, variant = _cfg(cfg)
    tab, nitems = items
    _lib.call('gwgs_gemm_peaks_multi', dtype_code, variant, Zt.data_ptr(), ldz, Ar.data_ptr(), Ai.data_ptr(), lda,
              M, K, tab.data_ptr(), nitems, mode, per, V.data_ptr(), I.data_ptr(), V.stride(0), _dev.ptr(rowscale),
              _dev.stream())


LM = 4   # local maxima per (template, window) in gemm_peaks_multi_lm


def gemm_peaks_multi_lm(dtype_code, Zt, ldz, Ar, Ai, lda, M, K, items, stacked, per, sep, cfg, V, I, VL, IL,
                        rowscale=None, gate=0.0, floor=0.0):
    """gemm_peaks_multi in window mode, also writing each window's LM largest local maxima of |z|^2 separated by
    more than sep samples (VL, row scaled, -1 for none) and their samples (IL) as (M, nwin * LM); local maxima below
    gate x max(window maximum, floor) (row scaled) need not be listed."""
    _, variant = _cfg(cfg)
    tab, nitems = items
    _lib.call('gwgs_gemm_peaks_multi_lm', dtype_code, variant, Zt.data_ptr(), ldz, Ar.data_ptr(), Ai.data_ptr(),
              lda, M, K, tab.data_ptr(), nitems, int(bool(stacked)), per, int(sep), float(gate), float(floor),
              V.data_ptr(), I.data_ptr(), V.stride(0), VL.data_ptr(), IL.data_ptr(), VL.stride(0),
              _dev.ptr(rowscale), _dev.stream())


def gemm_peaks_multi_lm_dyn(dtype_code, Zt, ldz, Ar, Ai, lda, M, K, items, nitems_dev, max_items, stacked, per,
                            sep, cfg, V, I, VL, IL, rowscale=None, gate=0.0, floor=0.0):
    """gemm_peaks_multi_lm over an item table whose length is the int32 nitems_dev on the device (at most
    max_items), as hot_items writes it."""
    _, variant = _cfg(cfg)
    _lib.call('gwgs_gemm_peaks_multi_lm_dyn', dtype_code, variant, Zt.data_ptr(), ldz, Ar.data_ptr(), Ai.data_ptr(),
              lda, M, K, items.data_ptr(), nitems_dev.data_ptr(), int(max_items), int(bool(stacked)), per, int(sep),
              float(gate), float(floor), V.data_ptr(), I.data_ptr(), V.stride(0), VL.data_ptr(), IL.data_ptr(),
              VL.stride(0), _dev.ptr(rowscale), _dev.stream())


# one segment of a batch for hot_items (csrc/gemm_peaks.hip HotSeg)
HOTSEG = np.dtype([('zrow0', '<i8'), ('col0', '<i4'), ('Tn', '<i4')])


def hot_segments(row0, col0, Tns):
    """Device table of a batch's segments for hot_items: first Zt row, first window column, analysed samples."""
    tab = np.zeros(len(Tns), dtype=HOTSEG)
    tab['zrow0'], tab['col0'], tab['Tn'] = row0[:len(Tns)], col0[:len(Tns)], Tns
    return _dev.to_dev(tab.view(np.uint8))


def hot_items(idx, val, count, nwsum, cut, segs, per, flags, items, nitems):
    """One-window GEMM items (ITEM rows in items, their count in the int32 nitems) for the windows holding a
    candidate of a compacted list (idx, val, count as peak_kernel.select_launch writes them) with val >= cut. segs:
    hot_segments of the batch; flags: nwsum int32 of workspace; items: room for nwsum ITEM rows."""
    _lib.call('gwgs_hot_items', idx.data_ptr(), val.data_ptr(), count.data_ptr(), idx.numel(), nwsum, float(cut),
              segs.data_ptr(), segs.numel() // HOTSEG.itemsize, per, flags.data_ptr(), items.data_ptr(),
              nitems.data_ptr(), _dev.stream())


def fused_peaks(Are, Aim, Z, M, Tn, cfg=None, zt=False, rowscale=None):
    """Are, Aim (M x K), Z (K x Tn) [or Zt (>= Tn x K) with zt=True], float32, bfloat16, float16 (the reference's
    operand type, on the fp16 MFMA) or float8_e4m3fn. Returns V, I: max |z|^2 and its sample per BN = 128 samples, (M, ceil(Tn/BN)); columns >= Tn are
    masked to -1 as in fp_gemm.py:29. cfg: (tiles_per_block,) from pick_cfg, or None."""
    Zt, Ar, Ai, K, code = _prep_operands(Are, Aim, Z, Tn, zt)
    nblk = -(-Tn // BN)
    V = torch.empty((M, nblk), device=Z.device, dtype=torch.float32)
    I = torch.empty((M, nblk), device=Z.device, dtype=torch.int32)
    gemm_peaks_raw(code, Zt, Zt.stride(0), Ar, Ai, Ar.stride(0), M, K, Tn, 0, 1, cfg, V, I, rowscale=rowscale)
    return V, I


def fused_window_peaks(Are, Aim, Z, M, Tn, cw, cfg=None, zt=False, rowscale=None):
    """As fused_peaks, with the cluster-window max fused: (M, Tn // cw) maxima and their samples, first index on
    ties at block and window level. Tn must be a multiple of cw and cw of BN (fp_filter.py:359)."""
    assert cw % BN == 0, 'cluster window must be a multiple of %d samples (GEMM output block)' % BN
    assert Tn % cw == 0
    Zt, Ar, Ai, K, code = _prep_operands(Are, Aim, Z, Tn, zt)
    nw = Tn // cw
    V = torch.empty((M, nw), device=Z.device, dtype=torch.float32)
    I = torch.empty((M, nw), device=Z.device, dtype=torch.int32)
    gemm_peaks_raw(code, Zt, Zt.stride(0), Ar, Ai, Ar.stride(0), M, K, Tn, 1, cw // BN, cfg, V, I,
                   rowscale=rowscale)
    return V, I


TPB_CANDIDATES = (16, 32, 64, 128)
VARIANTS = (0, 1, 2, 3, 4, 5)


def pick_cfg(dev='cuda', M=1024, K=128, Tn=1 << 20, dtype=torch.bfloat16, cw=2048, reps=3, stacked=True,
             lists=False):
    """The project's autotuner (the reference's sweep, fp_gemm.py:47-64): time every candidate launch shape once warm
    and keep the fastest on a representative shape. Candidates: time tiles per work item (window mode needs a
    multiple of cw / BN) x kernel variant. stacked: complex-stacked operands (Ai derived from Ar, mode | STACKED),
    the filter's own case. lists: time the launch with the local-maximum lists (gemm_peaks_multi_lm, whose kernel
    does not depend on the variant: tiles per item only), as the filter runs it when its first pass keeps them.
    Returns (cfg, {str(cfg): seconds or error}); raises RuntimeError if none runs."""
    # seeded operands drawn on the device (a CPU draw of the Tn x K operand took seconds of the search's set-up)
    g = torch.Generator(device=dev).manual_seed(0)
    K = _kpad(K) if K % 2 == 0 else _kpad(K + 1)
    A = torch.randn((M, K), generator=g, device=dev)
    Ai = torch.empty_like(A)
    Ai[:, 0::2], Ai[:, 1::2] = -A[:, 1::2], A[:, 0::2]
    code = DTYPES[dtype]
    Tn = Tn // cw * cw
    Zs = torch.randn((Tn, K), generator=g, device=dev)
    if code == 2:
        Ar_d, Ai_d, Z_d = (x.clamp(-448, 448).to(dev).to(torch.float8_e4m3fn).view(torch.uint8) for x in (A, Ai, Zs))
    else:
        Ar_d, Ai_d, Z_d = (x.to(dev, dtype) for x in (A, Ai, Zs))
    per = cw // BN
    V = torch.empty((M, Tn // cw), device=dev); I = torch.empty((M, Tn // cw), dtype=torch.int32, device=dev)
    mode = 1 | (STACKED if stacked else 0)
    if lists:
        VL = torch.empty((M, Tn // cw * LM), device=dev)
        IL = torch.empty((M, Tn // cw * LM), dtype=torch.int32, device=dev)

    def run(items, c):
        if lists:
            gemm_peaks_multi_lm(code, Z_d, K, Ar_d, Ai_d, K, M, K, items, stacked, per, 10, c, V, I, VL, IL,
                                gate=0.99, floor=0.0)
        else:
            gemm_peaks_multi(code, Z_d, K, Ar_d, Ai_d, K, M, K, items, mode, per, c, V, I)
    # every candidate run once (loads its kernel, brings the clock up), then timed in two interleaved rounds, the
    # faster round kept, so that no candidate pays for coming first
    res, cand = {}, []
    for tpb in TPB_CANDIDATES:
        for variant in (VARIANTS[-1:] if lists else VARIANTS):
            c = (max(per, tpb // per * per), variant)
            if str(c) in res or c in [x for x, _ in cand]:
                continue
            try:
                items = items_table([(0, Tn // BN, 0, 0, Tn)], c[0])
                run(items, c)
                torch.cuda.synchronize()
                cand.append((c, items))
            except Exception as ex:   # a shape that fails is recorded, not fatal
                res[str(c)] = repr(ex)[:300]
    for _ in range(2):
        for c, items in cand:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                run(items, c)
            torch.cuda.synchronize()
            t = (time.perf_counter() - t0) / reps
            res[str(c)] = min(t, res.get(str(c), t))
    best = min((c for c, _ in cand), key=lambda c: res[str(c)], default=None)
    if best is None:
        raise RuntimeError('no GEMM config ran: %r' % res)
    return best, res
