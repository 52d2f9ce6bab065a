"""Template banks on the GPU: closed-form stationary-phase chirps (csrc/templates.hip), sorted by chirp time and cut
into tiles as fp_filter does (fp_filter.py:416-448), and sigma^2 / norm against each segment's inverse PSD
(fp_filter.py:710-714)."""
from __future__ import annotations

import numpy as np
import torch

import _dev
import _lib

F0_ECC = 15.0     # reference frequency of the eccentric phase term (Hz)


def generate(m1, m2, e0, flen, df, flow, out=None):
    """Templates (n, flen) complex64 on the GPU and their end bins kmax (host int64)."""
    n = len(m1)
    H = out if out is not None else torch.empty((n, flen), dtype=torch.complex64, device='cuda')
    a, b = _dev.as_dev(np.asarray(m1, np.float64)), _dev.as_dev(np.asarray(m2, np.float64))
    e = _dev.as_dev(np.asarray(e0, np.float64)) if e0 is not None and np.any(np.asarray(e0) != 0) else None
    km = torch.empty(n, dtype=torch.int32, device='cuda')
    _lib.call('gwgs_spa_templates', a.data_ptr(), b.data_ptr(), _dev.ptr(e), n, flen, float(df), float(flow),
              F0_ECC, H.data_ptr(), H.stride(0), km.data_ptr(), _dev.stream())
    return H, km.cpu().numpy().astype(np.int64)


def sigmasq(H, IPall, df, kmin=0, kend=None):
    """sig (n, nseg) float64 = 4 df |H|^2 @ IPall^T, accumulated in double over [kmin, kend) (default: all of H's
    columns; H must be zero outside it); norm = 4 df / sqrt(sig)."""
    n = H.shape[0]
    nseg = IPall.shape[0]
    kend = H.shape[1] if kend is None else int(kend)
    lib = _lib.load()
    sig = torch.empty((n, nseg), dtype=torch.float64, device=H.device)
    work = torch.empty(max(lib.gwgs_sigmasq_workspace(n, int(kmin), kend, min(nseg, 16)), 1), dtype=torch.float64,
                       device=H.device)
    for s0 in range(0, nseg, 16):
        s1 = min(nseg, s0 + 16)
        _lib.call('gwgs_sigmasq', H.data_ptr(), H.stride(0), int(kmin), kend, n, IPall[s0].data_ptr(),
                  IPall.stride(0), s1 - s0, float(df), work.data_ptr(), sig[:, s0:].data_ptr(), sig.stride(0),
                  _dev.stream())
    return sig, (4.0 * df) / torch.sqrt(sig)


def tiles(tau0, tile):
    """fp_filter.py:419-420, 448: stable argsort by chirp time, cut into runs of `tile` templates."""
    order = np.argsort(tau0, kind='stable')
    return [order[b0:b0 + tile] for b0 in range(0, len(order), tile)]
