"""GPU power chi-squared at peaks (the reference's chisq_kernel.py:1-121): one block per (template row, up to P
peaks); for each equal-power bin j the block sums q[row, k] e^{+2 pi i k t / N} over the bin for each of its peaks
and accumulates |Z_j|^2 and Z_j. csrc/chisq.hip chisq_pts_k: runs of 32 bins per thread with the phasor reseeded
exactly from the 64-bit integer phase k t mod N and advanced by a rotation in between, wave64 butterfly sums, and
float64 accumulation over bins. N need not be a power of two (the reference's 32-bit wrap needed one)."""
from __future__ import annotations

import numpy as np
import torch

import _dev
import _lib

P = 8
THREADS = 256


def plan_groups(rows, times):
    """chisq_kernel.py:96-107: sort peaks by (row, time), cut each row's run into groups of <= P. Returns the sort
    order and per-group (row, start, count) over the sorted peaks."""
    rows = np.asarray(rows, np.int64); times = np.asarray(times, np.int64)
    npk = len(rows)
    o = np.lexsort((times, rows))
    rs = rows[o]
    if npk == 0:
        z = np.zeros(0, np.int64)
        return o, z, z, z
    run0 = np.r_[0, np.flatnonzero(np.diff(rs)) + 1]
    pos = np.arange(npk) - np.repeat(run0, np.diff(np.r_[run0, npk]))       # position inside the row's run
    starts = np.flatnonzero(pos % P == 0)
    counts = np.diff(np.r_[starts, npk])
    return o, rs[starts], starts.astype(np.int64), counts.astype(np.int64)


def bin_power_sums(q, rows, times, edges_per_peak, want_z=False, d=None, n_time=None):
    """sum_j |Z_j(t)|^2 for each peak; q is (nrows, >= N or flen) complex64 on the GPU. rows, times: int arrays per
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
    tdev = _dev.i64(times[o])
    out = torch.zeros(npk, dtype=torch.float32, device=q.device)
    zo = torch.zeros(npk, dtype=torch.complex64, device=q.device) if want_z else None
    dd = None
    if d is not None:
        dd = _dev.as_dev(d, torch.complex64)
    g_row, g_start, g_cnt, g_e = _dev.i32(grow), _dev.i32(gstart), _dev.i32(gcount), _dev.i32(gedges)
    _lib.call('gwgs_chisq_points', q.data_ptr(), q.shape[1], _dev.ptr(dd), int(N), g_row.data_ptr(),
              g_start.data_ptr(), g_cnt.data_ptr(), tdev.data_ptr(), g_e.data_ptr(), gedges.shape[1], len(gstart),
              out.data_ptr(), _dev.ptr(zo), _dev.stream())
    res = np.empty(npk, np.float32)
    res[o] = out.cpu().numpy()
    if not want_z:
        return res
    zr = np.empty(npk, np.complex64)
    zr[o] = zo.cpu().numpy()
    return res, zr
