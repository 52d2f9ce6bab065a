"""Equal-power chi-squared bin edges per template (the reference's chisq_edges.py:1-143; pycbc's non-SPA rule as
multi_filter.edges_per_template): edge j = first k with cumsum(|h|^2 / psd)[k] > j / nb * total, total over
[0, k1); clipped to [kmin, k1]; padded with k1 beyond each row's own bin count; last column k1.

One GPU kernel (csrc/chisq.hip edges_fused_k): one block per requested row, float64 chunk sums, a deterministic
block scan, and a walk inside the chunk that holds each target. Same edges as the float64 cumulative sum up to
summation order (+-1 bin at exact near-ties). edges_blocked keeps the reference's name and contract and runs the
same kernel once per distinct row."""
from __future__ import annotations

import numpy as np
import torch

import _dev
import _lib

BL = 256


def edges_rows_dev(rows_dev_idx, nb_, kmin, k1, Hd, ips):
    """Enqueue the edges of the given rows of Hd; returns the (n, B + 1) int32 device tensor without waiting."""
    n = len(rows_dev_idx)
    B = int(np.max(nb_)) if n else 0
    out = torch.empty((n, B + 1), dtype=torch.int32, device=Hd.device)
    if n == 0:
        return out
    assert Hd.dtype == torch.complex64 and Hd.is_contiguous()
    ipd = _dev.as_dev(ips, torch.float32)
    r, kk, nb = _dev.to_dev_many(*(np.asarray(x, np.int32) for x in (rows_dev_idx, k1, nb_)))
    e0 = _dev.METER.start()
    _lib.call('gwgs_edges_fused', Hd.data_ptr(), Hd.shape[1], r.data_ptr(), ipd.data_ptr(), kk.data_ptr(),
              nb.data_ptr(), B + 1, int(kmin), n, out.data_ptr(), _dev.stream())
    _dev.METER.stop('edges', e0, 12.0 * float(np.asarray(k1, np.int64).sum()))      # h (8 B) and ips (4 B) per bin
    return out


def edges_rows_dev_multi(rows, rseg, nb_, kmin, k1, Hd, ip_tab):
    """edges_rows_dev over requests of several segments: request g takes row rows[g] of Hd and 1/psd from the
    segment rseg[g], whose float32 row pointer is ip_tab[rseg[g]] (int64 host array of device addresses). Enqueued
    without waiting; returns the (n, B + 1) int32 device tensor."""
    n = len(rows)
    B = int(np.max(nb_)) if n else 0
    out = torch.empty((n, B + 1), dtype=torch.int32, device=Hd.device)
    if n == 0:
        return out
    assert Hd.dtype == torch.complex64 and Hd.is_contiguous()
    r, sg, kk, nb, tab = _dev.to_dev_many(*(np.asarray(x, np.int32) for x in (rows, rseg, k1, nb_)),
                                          np.asarray(ip_tab, np.int64))
    e0 = _dev.METER.start()
    _lib.call('gwgs_edges_fused_multi', Hd.data_ptr(), Hd.shape[1], r.data_ptr(), tab.data_ptr(), sg.data_ptr(),
              kk.data_ptr(), nb.data_ptr(), B + 1, int(kmin), n, out.data_ptr(), _dev.stream())
    _dev.METER.stop('edges', e0, 12.0 * float(np.asarray(k1, np.int64).sum()))
    return out


def _edges_rows(rows_dev_idx, nb_, kmin, k1, Hd, ips):
    out = edges_rows_dev(rows_dev_idx, nb_, kmin, k1, Hd, ips)
    return out.cpu().numpy().astype(np.int64)


def edges_fused(xp, rows, nb_, kmin, k1, Hd, ips):
    """One peak per row: rows index Hd directly. Returns (npk, B + 1) int64 (chisq_edges.py:250-264). xp is
    accepted for the reference's signature and ignored (arrays are torch tensors on the GPU)."""
    rows = np.asarray(rows, np.int64)
    return _edges_rows(rows, np.asarray(nb_, np.int64), kmin, np.asarray(k1, np.int64), Hd, ips)


def edges_blocked(xp, rows, nb_, kmin, k1, Hd, ips):
    """rows: per-peak row index into Hd; nb_, k1: per-peak bin count and end index (functions of the row). Work is
    done once per distinct row, then copied to that row's peaks (chisq_edges.py:168-208)."""
    rows = np.asarray(rows, np.int64)
    nb_ = np.asarray(nb_, np.int64); k1 = np.asarray(k1, np.int64)
    ur, first, inv = np.unique(rows, return_index=True, return_inverse=True)
    E = _edges_rows(ur, nb_[first], kmin, k1[first], Hd, ips)
    return E[inv]
