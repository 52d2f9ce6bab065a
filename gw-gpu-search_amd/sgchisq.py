"""Sine-Gaussian chi-squared veto (pycbc.vetoes.sgchisq.SingleDetSGChisq.values) for many peaks of one segment in one
launch: the reference's sg_kernel.py (sg_chisq_points, tile_ranges) and the --sgchisq-locations parsing of
fp_filter.py:449-457, on csrc/sgchisq.hip. The tile ranges are computed on the host in float64 with the reference's
index arithmetic (each tile's lower edge bounded by the previous tile's first bin; bins below round(flo / df) get
weight 0); the kernel sums with an exact per-bin phasor and float64 accumulation; the values are finished on the
host in float64."""
from __future__ import annotations

import re

import numpy as np
import torch

import _dev
import _lib


def tile_ranges(fpeak, vals, k_start, df):
    """Per peak and tile (sg_kernel.py:66-81): bin range kk, k1, weight start kg (int64) and fcen, tau = q / (2 pi
    fcen), q^2 (float64), each (npk, ntile)."""
    fpeak = np.asarray(fpeak, np.float64)
    npk, nt = len(fpeak), len(vals)
    kk_o, k1_o, kg_o = (np.empty((npk, nt), np.int64) for _ in range(3))
    fc_o, tau_o, q2_o = (np.empty((npk, nt)) for _ in range(3))
    kk = np.asarray(k_start, np.int64).copy()
    for j, (q, off) in enumerate(vals):
        fcen = fpeak + off
        flo = np.maximum(kk * df, fcen - 50)
        fhi = fcen + 50
        kk = (flo / df).astype(np.int64)
        kk_o[:, j], k1_o[:, j], kg_o[:, j] = kk, (fhi / df).astype(np.int64), np.round(flo / df).astype(np.int64)
        fc_o[:, j], tau_o[:, j], q2_o[:, j] = fcen, q / 2 / np.pi / fcen, q * q
    return kk_o, k1_o, kg_o, fc_o, tau_o, q2_o


class PendingSG:
    """sg_chisq values being computed on the device; result() waits for them and finishes in float64."""

    def __init__(self, n, ok, idx, nt, df, fetch):
        self.n, self.ok, self.idx, self.nt, self.df, self.fetch = n, ok, idx, nt, df, fetch

    def result(self):
        out = np.ones(self.n, np.float64)
        if self.fetch is not None:
            zh, gh = self.fetch.get()
            zz = (zh[0::2] + 1j * zh[1::2]).reshape(-1, self.nt) * (4.0 * self.df)
            gsig = np.sqrt(4 * self.df * gh).reshape(-1, self.nt)
            out[self.idx] = (np.abs(zz / gsig) ** 2).sum(1) / (2 * self.nt)
        return out, self.ok


def sg_chisq_launch(D, IP, t, fpeak, vals, k_start, df, tlen):
    """Enqueue sg_chisq_points' kernel and the copy of its sums; returns a PendingSG (no host wait)."""
    flen = int(D.shape[-1])
    assert D.dtype == torch.complex64 and IP.dtype == torch.float32 and D.is_contiguous() and IP.is_contiguous()
    assert IP.shape[-1] == flen and flen == tlen // 2 + 1
    fpeak = np.asarray(fpeak, np.float64)
    t = np.asarray(t, np.int64)
    ok = np.all([fpeak + off + 50 <= flen * df * 0.9 for _, off in vals], axis=0)
    idx = np.flatnonzero(ok)
    nt = len(vals)
    if len(idx) == 0:
        return PendingSG(len(t), ok, idx, nt, df, None)
    kk, k1, kg, fc, tau, q2 = tile_ranges(fpeak[idx], vals, np.asarray(k_start)[idx], df)
    assert k1.max() <= flen and kk.min() >= 0
    tabs = _dev.to_dev_many(kk.ravel(), k1.ravel(), kg.ravel(), fc.ravel(), tau.ravel(), q2.ravel(), t[idx])
    z = torch.empty(2 * len(idx) * nt, dtype=torch.float64, device=D.device)
    gs = torch.empty(len(idx) * nt, dtype=torch.float64, device=D.device)
    e0 = _dev.METER.start()
    _lib.call('gwgs_sg_tiles', D.data_ptr(), IP.data_ptr(), flen, int(tlen), float(df),
              *[x.data_ptr() for x in tabs], len(idx), nt, z.data_ptr(), gs.data_ptr(), _dev.stream())
    # algorithmic bytes: D and IP (12 B) per bin of every (peak, tile)
    _dev.METER.stop('sg_tiles', e0, 12.0 * float((k1 - kk).sum()))
    return PendingSG(len(t), ok, idx, nt, df, _dev.Fetch(z, gs, after=_dev.mark()))


def sg_chisq_points(D, IP, t, fpeak, vals, k_start, df, tlen):
    """sg_chisq at samples t (int) of one segment for peaks sharing the tile list vals [(q, offset Hz)]; fpeak and
    k_start per peak. D complex64 (flen,), IP float32 (flen,) on the GPU. Returns (values float64, computed mask): a
    peak whose tiles reach above 0.9 x Nyquist is not computed (value 1), as the reference's sg_chisq_points."""
    return sg_chisq_launch(D, IP, t, fpeak, vals, k_start, df, tlen).result()


_COND = re.compile(r'^\s*(mass1|mass2|mtotal|mchirp|eta|q)\s*(<=|>=|<|>|==)\s*([-+0-9.eE]+)\s*$')


def _params(m1, m2):
    m1, m2 = np.asarray(m1, np.float64), np.asarray(m2, np.float64)
    mt = m1 + m2
    eta = m1 * m2 / mt ** 2
    return dict(mass1=m1, mass2=m2, mtotal=mt, eta=eta, mchirp=mt * eta ** 0.6, q=np.maximum(m1, m2) / np.minimum(m1, m2))


def parse_locations(descrs, m1, m2):
    """--sgchisq-locations (fp_filter.py:449-457): each 'region:q-offset,q-offset,...' assigns its tile list to the
    templates where region holds, later entries overriding earlier ones; templates no region covers get None. A
    region is 'else' or comparisons of mass1, mass2, mtotal, mchirp, eta or q with a number joined by '&' (the forms
    pycbc's parse_boolargs evaluates on bank parameters)."""
    par = _params(m1, m2)
    n = len(par['mass1'])
    locs = [None] * n
    for descr in descrs:
        region, values = descr.split(':')
        vals = [tuple(map(float, v.split('-'))) for v in values.split(',')]
        if region.strip() == 'else':
            mask = np.ones(n, bool)
        else:
            mask = np.ones(n, bool)
            for cond in region.split('&'):
                m = _COND.match(cond)
                assert m, f'unsupported sgchisq region {region!r}'
                x, op, v = par[m.group(1)], m.group(2), float(m.group(3))
                mask &= {'<': x < v, '>': x > v, '<=': x <= v, '>=': x >= v, '==': x == v}[op]
        for i in np.nonzero(mask)[0]:
            locs[i] = vals
    return locs
