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


# ---- coarse-node (interpolated) templates: the reference's --templates interp (imr_interp.py) -----------------------

def coarse_grid(flow, fhigh, n_chirp=6000, n_log=1500):
    """Node frequencies of the interpolated templates (imr_interp.py:11-15): n_chirp points uniform in f^(-5/3), so
    the inspiral phase advances by about the same amount between neighbours, joined with n_log points uniform in
    log f, and flow and fhigh, inside [flow, fhigh]."""
    u = np.linspace(flow ** (-5 / 3), fhigh ** (-5 / 3), n_chirp) ** (-3 / 5)
    x = np.unique(np.r_[u, np.geomspace(flow, fhigh, n_log), flow, fhigh])
    return x[(x >= flow) & (x <= fhigh)]


class Grid:
    """The fine bins [kmin, kend) against the nodes x, shared by every template of a bank (imr_interp.py:53-62):
    bin k at f = k df (clipped to [x_0, x_J]) lies in interval node[k - kmin] = j, d = f - x_j, t = d / (x_j+1 - x_j);
    kb[j] = the first bin of interval j (kb[J] = kend; an interval may hold no bin). Host arrays and their device
    copies (float64 for the dense evaluation, float32 for the regenerated exact stage)."""

    def __init__(self, x, df, kmin, kend):
        self.x = np.asarray(x, np.float64)
        self.nx, self.J = len(self.x), len(self.x) - 1
        self.df, self.kmin, self.kend = float(df), int(kmin), int(kend)
        f = np.clip(np.arange(self.kmin, self.kend) * self.df, self.x[0], self.x[-1])
        i = np.clip(np.searchsorted(self.x, f, side='right') - 1, 0, self.nx - 2)
        self.node = i.astype(np.int32)
        self.d = f - self.x[i]
        self.t = self.d / (self.x[i + 1] - self.x[i])
        self.kb = (self.kmin + np.searchsorted(self.node, np.arange(self.nx), side='left')).astype(np.int32)
        self.x_d, self.node_d, self.d_d, self.t_d, self.kb_d = _dev.to_dev_many(
            self.x, self.node, self.d, self.t, self.kb)
        self.d32, self.t32 = self.d_d.to(torch.float32), self.t_d.to(torch.float32)
        u = 1.0 - self.t_d
        self._w = (u * u, u * self.t_d, self.t_d * self.t_d)
        self._start = torch.as_tensor(self.kb[:-1] - self.kmin, dtype=torch.int64, device='cuda')
        last = self.kb[1:] - self.kmin - 1
        self._last = torch.as_tensor(np.maximum(last, 0), dtype=torch.int64, device='cuda')
        self._nonempty = torch.as_tensor(last >= self.kb[:-1] - self.kmin, device='cuda')
        self._node64 = self.node_d.to(torch.int64)
        self._mom = {}

    def moments(self, IP):
        """Per-segment moments of 1/psd in the (1-t)^2, t(1-t), t^2 basis (float64): Pk (kend - kmin, 3), each
        bin's in-interval inclusive prefix, and Mi (J, 3), each interval's total (equal to Pk at its last bin, 0 for
        an empty interval): the weight of |h|^2 IP over an interval is A_j^2 M0 + 2 A_j A_j+1 M1 + A_j+1^2 M2.
        Kept for the last few segments (keyed by the tensor, its address and its version counter, so new data in the
        same buffer is seen)."""
        key = (id(IP), IP.data_ptr(), IP._version)
        hit = self._mom.get(key)
        if hit is not None:
            return hit
        w = IP[self.kmin:self.kend].to(torch.float64)
        G = torch.cumsum(torch.stack([c * w for c in self._w]), 1)              # (3, K), scanned along rows
        Gb = torch.cat([torch.zeros((3, 1), dtype=torch.float64, device=G.device), G], 1)
        Pk = (G - Gb[:, self._start[self._node64]]).T.contiguous()
        Mi = torch.where(self._nonempty[:, None], Pk[self._last], torch.zeros_like(Pk[self._last])).contiguous()
        if len(self._mom) >= 32:
            self._mom.pop(next(iter(self._mom)))
        self._mom[key] = (Pk, Mi, IP)
        return Pk, Mi, IP


def nodes(m1, m2, e0, grid):
    """The chirps at the grid's nodes (csrc/templates.hip spa_nodes_k): amp (n, nx) float64, f^(-7/6) below f_isco
    and 0 above, and the phase cubics coef (n, J, 4) float64 as (c3, c2, c1, c0) in d = f - x_j (Hermite through the
    phase and its analytic derivative at both ends)."""
    n = len(m1)
    a, b = _dev.as_dev(np.asarray(m1, np.float64)), _dev.as_dev(np.asarray(m2, np.float64))
    e = _dev.as_dev(np.asarray(e0, np.float64)) if e0 is not None and np.any(np.asarray(e0) != 0) else None
    amp = torch.empty((n, grid.nx), dtype=torch.float64, device='cuda')
    coef = torch.empty((n, grid.J, 4), dtype=torch.float64, device='cuda')
    _lib.call('gwgs_spa_nodes', a.data_ptr(), b.data_ptr(), _dev.ptr(e), n, grid.x_d.data_ptr(), grid.nx, F0_ECC,
              amp.data_ptr(), coef.data_ptr(), _dev.stream())
    return amp, coef


def interp_end(m1, m2, grid):
    """Each interpolated chirp's end bin: its amplitude is 0 from the first node at or above f_isco, so bins from
    floor(that node / df) + 1 on are zero (at most kend). int64."""
    fisco = 1.0 / (6.0 ** 1.5 * np.pi * (np.asarray(m1) + np.asarray(m2)) * 4.925490947641267e-06)
    j = np.searchsorted(grid.x, fisco, side='left')
    end = np.where(j < grid.nx, np.floor(grid.x[np.minimum(j, grid.nx - 1)] / grid.df).astype(np.int64) + 1,
                   grid.kend)
    return np.minimum(end, grid.kend).astype(np.int64)


def interp_templates(amp, coef, grid, flen, out=None):
    """Dense rows (n, flen) complex64 of the interpolated templates (imr_interp.py FineInterpolator): amplitude
    linear and phase cubic per interval, evaluated in double on [kmin, kend), zero elsewhere."""
    n = amp.shape[0]
    H = out if out is not None else torch.empty((n, flen), dtype=torch.complex64, device='cuda')
    _lib.call('gwgs_interp_templates', amp.data_ptr(), coef.data_ptr(), n, grid.nx, grid.node_d.data_ptr(),
              grid.d_d.data_ptr(), grid.t_d.data_ptr(), grid.kmin, grid.kend, flen, H.data_ptr(), H.stride(0),
              _dev.stream())
    return H


def regen_tables(amp, coef):
    """The regenerated exact stage's per-interval tables (exact_stage.py:319-327): coef32 (n, J, 4) float32 with c0
    reduced to [-pi, pi) in float64 first, and amp2 (n, J, 2) float32 = (A_j, A_j+1)."""
    c = coef.clone()
    c[..., 3] = torch.remainder(c[..., 3] + np.pi, 2 * np.pi) - np.pi
    amp2 = torch.stack([amp[:, :-1], amp[:, 1:]], -1)
    return c.to(torch.float32).contiguous(), amp2.to(torch.float32).contiguous()
