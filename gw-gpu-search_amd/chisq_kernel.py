"""GPU power chi-squared at peaks (the reference's chisq_kernel.py:1-121): the peaks of one template row are taken
in groups of up to P samples; for each group and each equal-power bin j one wave sums q[row, k] e^{+2 pi i k t / N}
over the bin for each sample (csrc/chisq.hip chisq_bins_k: lanes on interleaved bins, the phasor reseeded exactly
from the 64-bit integer phase k t mod N at each run of 32 steps and advanced by a rotation in between, a wave
butterfly sum), and chisq_finish_k accumulates |Z_j|^2 and Z_j over the bins in float64, in bin order. N need not
be a power of two (the reference's 32-bit wrap needed one). P = 7 is one peak's 7 re-centring samples."""
from __future__ import annotations

import numpy as np
import torch

import _dev
import _lib

P = 7                 # samples per group; the same constant as csrc/chisq.hip's P
THREADS = 256


def groups_of_sorted(rs):
    """Groups of <= P consecutive peaks within each run of equal rows of a row-sorted peak list: (row, start,
    count) per group."""
    npk = len(rs)
    if npk == 0:
        z = np.zeros(0, np.int64)
        return z, z, z
    run0 = np.r_[0, np.flatnonzero(np.diff(rs)) + 1]
    pos = np.arange(npk) - np.repeat(run0, np.diff(np.r_[run0, npk]))       # position inside the row's run
    starts = np.flatnonzero(pos % P == 0)
    counts = np.diff(np.r_[starts, npk])
    return rs[starts], starts.astype(np.int64), counts.astype(np.int64)


def plan_groups(rows, times):
    """chisq_kernel.py:96-107: sort peaks by (row, time), cut each row's run into groups of <= P. Returns the sort
    order and per-group (row, start, count) over the sorted peaks."""
    rows = np.asarray(rows, np.int64); times = np.asarray(times, np.int64)
    npk = len(rows)
    o = np.lexsort((times, rows))
    grow, gstart, gcount = groups_of_sorted(rows[o])
    return o, grow, gstart, gcount


def _launch(q, times_sorted, grow, gstart, gcount, gedges_dev, want_z, d, N, span):
    """Enqueue gwgs_chisq_points for planned groups; gedges_dev (ngroups, nedge) int32 on the device. Returns the
    device outputs in sorted peak order. span: summed bin span of the groups (for the kernel meter's bytes)."""
    npk = len(times_sorted)
    out = torch.empty(npk, dtype=torch.float32, device=q.device)
    zo = torch.empty(npk, dtype=torch.complex64, device=q.device) if want_z else None
    dd = None
    if d is not None:
        dd = _dev.as_dev(d, torch.complex64)
    tdev, g_row, g_start, g_cnt = _dev.to_dev_many(np.asarray(times_sorted, np.int64), *(np.asarray(x, np.int32)
                                                                                         for x in (grow, gstart, gcount)))
    ge = gedges_dev if isinstance(gedges_dev, torch.Tensor) else _dev.to_dev(gedges_dev)
    ge = ge.contiguous()
    nwork = int(_lib.load().gwgs_chisq_points_workspace(len(gstart), ge.shape[1]))
    work = torch.empty(max(nwork, 8), dtype=torch.uint8, device=q.device)
    e0 = _dev.METER.start()
    _lib.call('gwgs_chisq_points', q.data_ptr(), q.shape[1], _dev.ptr(dd), int(N), g_row.data_ptr(),
              g_start.data_ptr(), g_cnt.data_ptr(), tdev.data_ptr(), ge.data_ptr(), ge.shape[1], len(gstart),
              out.data_ptr(), _dev.ptr(zo), work.data_ptr(), _dev.stream())
    # HBM bytes: per group the row (and the data) over its bins, 8 B each
    _dev.METER.stop('chisq', e0, span * (16 if d is not None else 8))
    return out, zo


def launch_sorted(q, rows, times, E_dev, e_index, want_z=False, d=None, n_time=None, row_span=None):
    """Enqueue the sums without waiting, for peaks already sorted by row (any time order within a row): edges come as
    device rows E_dev (nu, nedge) int32 with e_index[p] the row of peak p. Returns device (out, zo) in the given
    peak order. row_span: bins from the first to the last edge per row of q (for the kernel meter; optional)."""
    N = q.shape[1] if d is None else n_time
    assert N is not None and N > 0
    assert q.dtype == torch.complex64 and q.is_contiguous()
    rows = np.asarray(rows, np.int64)
    grow, gstart, gcount = groups_of_sorted(rows)
    gi = _dev.to_dev(np.asarray(e_index, np.int64)[gstart])
    span = 0.0 if row_span is None else float(np.asarray(row_span)[grow].sum())
    return _launch(q, np.asarray(times, np.int64), grow, gstart, gcount, E_dev[gi], want_z, d, N, span)


def launch_sorted_multi(h, keys, n_rows, times, E_dev, e_index, d_tab, n_time, row_span=None, regen=None):
    """launch_sorted in fused mode over peaks of several segments: keys = segment n_rows + row, sorted (any time
    order within a key); the template rows are h's, the data row of segment s is at d_tab[s] (int64 host array of
    device addresses of complex64 rows). regen (an exact_stage.Coarse): the rows are regenerated from their nodes
    instead of read from h (gwgs_regen_points_multi). Returns device (out, zo) in the given peak order."""
    assert h.dtype == torch.complex64 and h.is_contiguous()
    keys = np.asarray(keys, np.int64)
    gkey, gstart, gcount = groups_of_sorted(keys)
    npk = len(keys)
    out = torch.empty(npk, dtype=torch.float32, device=h.device)
    zo = torch.empty(npk, dtype=torch.complex64, device=h.device)
    if len(gkey) == 0:
        return out, zo
    gi = np.asarray(e_index, np.int64)[gstart]
    tdev, g_row, g_seg, g_start, g_cnt, g_i, tab = _dev.to_dev_many(
        np.asarray(times, np.int64), (gkey % n_rows).astype(np.int32), (gkey // n_rows).astype(np.int32),
        gstart.astype(np.int32), gcount.astype(np.int32), gi, np.asarray(d_tab, np.int64))
    ge = E_dev[g_i].contiguous()
    nwork = int(_lib.load().gwgs_chisq_points_workspace(len(gstart), ge.shape[1]))
    work = torch.empty(max(nwork, 8), dtype=torch.uint8, device=h.device)
    span = 0.0 if row_span is None else float(np.asarray(row_span)[gkey % n_rows].sum())
    e0 = _dev.METER.start()
    if regen is not None:
        g = regen.grid
        _lib.call('gwgs_regen_points_multi', regen.coef32.data_ptr(), regen.amp2.data_ptr(), g.nx,
                  g.node_d.data_ptr(), g.d32.data_ptr(), g.t32.data_ptr(), g.kmin, tab.data_ptr(), g_seg.data_ptr(),
                  int(n_time), g_row.data_ptr(), g_start.data_ptr(), g_cnt.data_ptr(), tdev.data_ptr(), ge.data_ptr(),
                  ge.shape[1], len(gstart), out.data_ptr(), zo.data_ptr(), work.data_ptr(), _dev.stream())
        # HBM bytes: the data per bin (the grid and the node tables are shared by the rows and stay in cache)
        _dev.METER.stop('chisq_regen', e0, span * 8)
    else:
        _lib.call('gwgs_chisq_points_multi', h.data_ptr(), h.shape[1], tab.data_ptr(), g_seg.data_ptr(), int(n_time),
                  g_row.data_ptr(), g_start.data_ptr(), g_cnt.data_ptr(), tdev.data_ptr(), ge.data_ptr(),
                  ge.shape[1], len(gstart), out.data_ptr(), zo.data_ptr(), work.data_ptr(), _dev.stream())
        _dev.METER.stop('chisq', e0, span * 16)
    return out, zo


def bin_power_sums(q, rows, times, edges_per_peak, want_z=False, d=None, n_time=None):
    """sum_j |Z_j|^2 for each peak; q is (nrows, >= N or flen) complex64 on the GPU. rows, times: int arrays per
    peak; edges_per_peak: (npeaks, nedge) int bin edges into q's row (bins with e_j >= e_j+1 are empty).
    want_z: also return z(t) = sum_j Z_j. d, n_time: fused mode, q holds the template rows H (nrows, flen), d the
    data (flen,), the kernel forms conj(H) d itself and n_time is the time length N of the phase factor.
    Returns host numpy arrays in request order (float32; complex64)."""
    N = q.shape[1] if d is None else n_time
    assert N is not None and N > 0
    assert q.dtype == torch.complex64 and q.is_contiguous()
    rows = np.asarray(rows, np.int64); times = np.asarray(times, np.int64)
    npk = len(rows)
    if npk == 0:
        return (np.zeros(0, np.float32), np.zeros(0, np.complex64)) if want_z else np.zeros(0, np.float32)
    E = np.asarray(edges_per_peak)
    o, grow, gstart, gcount = plan_groups(rows, times)
    gedges = np.ascontiguousarray(E[o][gstart].astype(np.int32))     # edges depend on the row only
    span = float((gedges[:, -1] - gedges[:, 0]).clip(min=0).sum())
    out, zo = _launch(q, times[o], grow, gstart, gcount, _dev.to_dev(gedges), want_z, d, N, span)
    ho, hz = _dev.Fetch(out, zo).get()
    res = np.empty(npk, np.float32)
    res[o] = ho
    if not want_z:
        return res
    zr = np.empty(npk, np.complex64)
    zr[o] = hz
    return res, zr
