"""Differentiable peak SNR of compressed templates and refinement of loud triggers off the discrete bank.

The rebuilt SNR of fp_filter.py:759-764 depends on a template only through its coefficient vector a in the tile's
SVD basis (z = sum_j a_j Zc_j). Here it is evaluated at a continuous time t + delta from the frequency domain (the
basis responses R from per-peak moments computed by the filter's fp32 basis transform, see ResponseContext; to 1e-9
of the exact sum), normalised by the template's own
sigma (sigma^2 = 4 df a^T G conj(a), G the tile's basis metric under the segment's 1/psd), and differentiated by a
hand-written kernel (snr_grad_k): rho^2 and its gradients with respect to a (2r real coordinates), delta and the
phase. PeakSNR2 wraps it as a torch.autograd.Function; refine() runs projected gradient ascent on the convex hull of
the nearest bank templates' coefficient vectors (and on delta), which refines the loudest triggers off the
discrete bank.
"""
from __future__ import annotations

import numpy as np
import torch

import _dev
import _lib


class ResponseContext:
    """One (tile, segment): the tile's basis Vc over [kmin, kend), the segment's data D and 1/psd IP, and the metric
    G[j, j'] = sum_k Vc[j, k] conj(Vc[j', k]) IP[k] (formed in complex128).

    The basis responses R_j(t + delta) = sum_k Vc[j, k] D[k] e^{2 pi i k (t + delta) / n} are evaluated, for the
    tile's power-of-two segment lengths, from moments: with kc the band's centre, H its half width and
    u_k = (k - kc) / H in [-1, 1],
        R_j(t + delta) = e^{2 pi i kc delta / n} sum_m (i a delta)^m / m! S_mj(t),   a = 2 pi H / n,
        S_mj(t) = sum_k Vc[j, k] D[k] u_k^m e^{2 pi i k t / n},
    where each S_m over every sample is the basis series of the data D u^m (the filter's fused fp32 four-step
    transform) read at the peaks' samples; M + 1 terms with (a dmax)^(M+1) / (M+1)! < 1e-9 for |delta| <= dmax.
    A peak's moments are computed once; every later evaluation at a new delta is a dot product of M + 1 terms.
    Other lengths form D e^{2 pi i k (t + delta) / n} and multiply by Vc (a library GEMM)."""

    DMAX = 2.0

    def __init__(self, tb, D, IP, df, tlen, kmin):
        self.k0, k1 = int(kmin), int(tb.kend)
        self.kb = k1 - self.k0
        self.r, self.rp, self.n, self.df = tb.r, tb.rp, int(tlen), float(df)
        self.tb = tb
        V64 = tb.Vc[:tb.r, self.k0:k1].to(torch.complex128)
        self.G = ((V64 * IP[self.k0:k1].to(torch.float64)) @ V64.conj().T).to(torch.complex64).contiguous()
        del V64
        self.D = D
        self._Vt = None
        self.ws = int(_lib.load().gwgs_basis_series_workspace(tb.r, self.n)) if self.kb > 0 else 0
        kc = 0.5 * (self.k0 + k1 - 1)
        H = max(0.5 * (k1 - 1 - self.k0), 0.5)
        self.kc, self.a = kc, 2 * np.pi * H / self.n
        x, M, term = self.a * self.DMAX, 0, 1.0
        while True:
            term *= x / (M + 1)
            if term < 1e-9:
                break
            M += 1
        self.M = M
        self.u = ((torch.arange(self.k0, k1, device='cuda', dtype=torch.float64) - kc) / H).to(torch.float32)
        self._tt = np.zeros(0, np.int64)
        self._S = None
        self._last_t = self._idx = None

    @property
    def Vt(self):
        """(kb, r) complex64: the band of Vc transposed, for the library-GEMM form."""
        if self._Vt is None:
            self._Vt = self.tb.Vc[:self.r, self.k0:self.k0 + self.kb].T.contiguous()
        return self._Vt

    def prepare(self, t):
        """Moments for the samples t (host or device int64): computed for those not held yet."""
        tt = np.unique(np.asarray(_dev.host(t), np.int64))
        new = np.setdiff1d(tt, self._tt)
        if len(new) == 0:
            return
        S_new = self._moments(new)
        self._last_t = None
        if self._S is None:
            self._tt, self._S = new, S_new
        else:
            allt = np.concatenate([self._tt, new])
            o = np.argsort(allt)
            self._tt = allt[o]
            self._S = torch.cat([self._S, S_new])[_dev.to_dev(o)]

    def _moments(self, tt):
        """(len(tt), M + 1, r) complex64 moments S_m(t) at sorted samples tt."""
        r, rp, K = self.r, self.rp, 2 * self.rp
        S = torch.empty((len(tt), self.M + 1, r), dtype=torch.complex64, device='cuda')
        work = torch.empty(max(self.ws, 8), dtype=torch.uint8, device='cuda')
        Dm = self.D.clone()
        band = slice(self.k0, self.k0 + self.kb)
        # runs of samples whose span keeps one output buffer under 2 GiB
        span = max(1, (1 << 31) // (K * 4))
        runs, i = [], 0
        while i < len(tt):
            j = int(np.searchsorted(tt, tt[i] + span, side='left'))
            runs.append((i, j)); i = j
        for m in range(self.M + 1):
            if m:
                Dm[band] *= self.u
            for i, j in runs:
                a0, Tn = int(tt[i]), int(tt[j - 1] - tt[i] + 1)
                Zt = torch.empty((Tn, K), dtype=torch.float32, device='cuda')
                _lib.call('gwgs_basis_series', 0, self.tb.Vc.data_ptr(), self.tb.Vc.shape[1], r, rp, Dm.data_ptr(),
                          self.k0, self.k0 + self.kb, self.n, a0, Tn, None, 0, None, work.data_ptr(), Zt.data_ptr(),
                          K, _dev.stream())
                rows = Zt[_dev.to_dev(tt[i:j] - a0)]
                S[i:j, m] = torch.complex(rows[:, 0:2 * r:2], rows[:, 1:2 * r:2])
        return S

    def response(self, t, delta):
        """Basis responses R, dR/ddelta (npk, r) complex64 at times t + delta (t int64, delta float64 tensors,
        |delta| <= DMAX)."""
        npk = t.shape[0]
        if self.ws > 0:
            if t is not self._last_t:
                self.prepare(t)
                self._idx = torch.searchsorted(_dev.to_dev(self._tt), t.to('cuda', torch.int64).contiguous())
                self._last_t = t
            S = self._S[self._idx]                                             # (npk, M + 1, r)
            d = delta.to('cuda', torch.float64)
            m = torch.arange(1, self.M + 1, device='cuda', dtype=torch.float64)
            # c_m = (i a delta)^m / m! by its recurrence, and dc_m / ddelta = i a c_(m-1)
            f = torch.ones((npk, self.M + 1), dtype=torch.complex128, device='cuda')
            f[:, 1:] = (1j * self.a) * d[:, None] / m[None, :]
            c = torch.cumprod(f, dim=1)
            cd = torch.zeros_like(c)
            cd[:, 1:] = (1j * self.a) * c[:, :-1]
            w0 = 2j * np.pi * self.kc / self.n
            ph = torch.exp(w0 * d)[:, None]
            cR = (ph * c).to(torch.complex64)
            cRd = (ph * (cd + w0 * c)).to(torch.complex64)
            return torch.bmm(cR[:, None, :], S)[:, 0], torch.bmm(cRd[:, None, :], S)[:, 0]
        E = torch.empty((npk, self.kb), dtype=torch.complex64, device='cuda')
        Ed = torch.empty_like(E)
        td, dd = t.to('cuda', torch.int64).contiguous(), delta.to('cuda', torch.float64).contiguous()
        _lib.call('gwgs_phasor_data', self.D.data_ptr(), self.k0, self.kb, td.data_ptr(), dd.data_ptr(), self.n, npk,
                  E.data_ptr(), Ed.data_ptr(), self.kb, _dev.stream())
        return E @ self.Vt, Ed @ self.Vt


def snr2(ctx: ResponseContext, a, t, delta, phi=None, want_grad=True):
    """rho^2, z, d rho^2 / d delta, the phase derivative and (want_grad) d rho^2 / d a for coefficient vectors
    a (npk, r) complex64 at times t + delta. Device tensors."""
    npk, r = a.shape
    assert r == ctx.r
    R, Rd = ctx.response(t, delta)
    a = a.contiguous()
    rho2 = torch.empty(npk, dtype=torch.float32, device='cuda')
    z = torch.empty(npk, dtype=torch.complex64, device='cuda')
    dt = torch.empty_like(rho2)
    dphi = torch.empty_like(rho2)
    grad = torch.empty((npk, r), dtype=torch.complex64, device='cuda') if want_grad else None
    ph = None if phi is None else phi.to('cuda', torch.float32).contiguous()
    _lib.call('gwgs_snr_grad', R.data_ptr(), Rd.data_ptr(), r, a.data_ptr(), r, ctx.G.data_ptr(), r, r, npk,
              ctx.df, _dev.ptr(ph), rho2.data_ptr(), z.data_ptr(), dt.data_ptr(), dphi.data_ptr(), _dev.ptr(grad), r,
              _dev.stream())
    return dict(rho2=rho2, z=z, drho2_ddelta=dt, drho2_dphi=dphi, grad=grad)


class PeakSNR2(torch.autograd.Function):
    """rho^2(a, delta): a as real (npk, r, 2) float32 (Re, Im), delta (npk,) float64; ctx and t are constants. The
    kernel returns the gradients with the value; backward scales them by the incoming gradient."""

    @staticmethod
    def forward(fctx, a_ri, delta, ctx, t):
        out = snr2(ctx, torch.view_as_complex(a_ri.detach().contiguous()), t, delta.detach())
        fctx.save_for_backward(torch.view_as_real(out['grad']), out['drho2_ddelta'])
        return out['rho2']

    @staticmethod
    def backward(fctx, g):
        ga, gt = fctx.saved_tensors
        return ga * g[:, None, None], (gt * g).to(torch.float64), None, None


def _project_simplex(w):
    """Euclidean projection of each row of w onto the probability simplex (sort-based)."""
    u, _ = torch.sort(w, dim=1, descending=True)
    css = u.cumsum(1) - 1
    ind = torch.arange(1, w.shape[1] + 1, device=w.device, dtype=w.dtype)
    cond = u - css / ind > 0
    rho = cond.to(torch.int64).cumsum(1).argmax(1, keepdim=True)               # last index where cond holds
    theta = css.gather(1, rho) / (rho + 1).to(w.dtype)
    return (w - theta).clamp(min=0)


def refine(ctx: ResponseContext, A, rows, t, neighbours=8, steps=20, step0=0.5, dmax=1.5):
    """Projected gradient ascent of rho^2 from bank templates `rows` (of the tile's coefficient matrix A (n, r)
    complex64) at integer samples t: a = sum_i w_i A[nb_i] over the `neighbours` templates nearest in normalised
    coefficient space, each phase-aligned to the starting template (w on the simplex: the bank's local coefficient
    hull), delta in [-dmax, dmax]. Each step is
    normalised by rho^2 and accepted per trigger only if rho^2 grows (the step then grows by 1.5, else halves).
    Returns dict(rho2_start, rho2, w, nb, delta, a)."""
    assert dmax <= ctx.DMAX
    rows = torch.as_tensor(rows, device='cuda', dtype=torch.int64)
    t = torch.as_tensor(t, device='cuda', dtype=torch.int64)
    npk = rows.shape[0]
    An = A / A.abs().pow(2).sum(1, keepdim=True).sqrt().clamp(min=1e-30)
    ov = An[rows] @ An.conj().T                                                 # (npk, n) overlaps in the basis
    L = min(neighbours, A.shape[0])
    nb = ov.abs().topk(L, dim=1).indices                                       # the row itself first
    # each neighbour's overall phase aligned to the starting template, so that convex combinations add coherently
    ph = torch.gather(ov, 1, nb).conj()
    ph = ph / ph.abs().clamp(min=1e-30)
    Ari = torch.view_as_real((A[nb] * ph.conj()[:, :, None]).contiguous())     # (npk, L, r, 2)
    w = torch.zeros((npk, L), dtype=torch.float32, device='cuda')
    w[torch.arange(npk, device='cuda'), (nb == rows[:, None]).to(torch.int64).argmax(1)] = 1.0
    delta = torch.zeros(npk, dtype=torch.float64, device='cuda')
    eta = torch.full((npk,), step0, dtype=torch.float64, device='cuda')

    def value_grad(w, delta):
        wv = w.clone().requires_grad_(True)
        dv = delta.clone().requires_grad_(True)
        a_ri = (wv[:, :, None, None] * Ari).sum(1)
        rho2 = PeakSNR2.apply(a_ri, dv, ctx, t)
        rho2.sum().backward()
        return rho2.detach(), wv.grad.detach(), dv.grad.detach()

    rho2, gw, gd = value_grad(w, delta)
    start = rho2.clone()
    for _ in range(steps):
        scale = (eta / rho2.to(torch.float64).clamp(min=1e-30))
        w_new = _project_simplex(w + (scale[:, None] * gw.to(torch.float64)).to(torch.float32))
        d_new = (delta + scale * gd.to(torch.float64)).clamp(-dmax, dmax)
        r_new, gw_new, gd_new = value_grad(w_new, d_new)
        up = r_new > rho2
        w = torch.where(up[:, None], w_new, w)
        delta = torch.where(up, d_new, delta)
        gw = torch.where(up[:, None], gw_new, gw)
        gd = torch.where(up, gd_new, gd)
        rho2 = torch.where(up, r_new, rho2)
        eta = torch.where(up, eta * 1.5, eta * 0.5)
    a = torch.view_as_complex((w[:, :, None, None] * Ari).sum(1).contiguous())
    return dict(rho2_start=start, rho2=rho2, w=w, nb=nb, delta=delta, a=a)


def refine_triggers(plans, datasets, outs, top=10000, neighbours=8, steps=20, chunk=256):
    """The loudest exact triggers of a filtering pass refined off the discrete bank: the `top` loudest exact triggers
    over banks and datasets (outs[b][d]: the trigger dict of plans[b] on datasets[d], fields as Triggers.to_dict),
    each refined by refine() from its own template on its tile's coefficient hull at its sample, grouped by
    (bank, dataset, segment, tile) and taken `chunk` triggers per call. Returns numpy arrays over the refined triggers:
    bank, dataset, index (into outs[b][d]), template_id, snr (the exact SNR), end_time, rho_start (the compressed
    template's SNR at the trigger's sample), rho (refined), delta (samples)."""
    cand = []
    for b in range(len(plans)):
        for d in range(len(datasets)):
            tr = outs[b][d]
            cand += [(float(tr['snr'][i]), b, d, int(i)) for i in np.flatnonzero(tr['exact'])]
    cand.sort(key=lambda c: -c[0])
    cand = cand[:top]
    seg0 = np.cumsum([0] + [len(ds.analyze) for ds in datasets])
    tile_of = [{int(t): (ti, r) for ti, tb in enumerate(p.tiles) for r, t in enumerate(tb.ids)} for p in plans]
    groups = {}
    for _, b, d, i in cand:
        tr = outs[b][d]
        ti, r = tile_of[b][int(tr['template_id'][i])]
        s = int(tr['segment'][i]) - int(seg0[d])
        groups.setdefault((b, d, s, ti), []).append((i, r, int(tr['sample'][i])))
    cols = {k: [] for k in ('bank', 'dataset', 'index', 'rho_start', 'rho', 'delta')}
    for (b, d, s, ti), lst in groups.items():
        tb, ds = plans[b].tiles[ti], datasets[d]
        ctx = ResponseContext(tb, ds.D[s], ds.IP[s], ds.df, ds.tlen, plans[b].kmin)
        if ctx.ws > 0:
            ctx.prepare(np.array([p[2] for p in lst], np.int64))       # every trigger's moments in one pass
        for c in range(0, len(lst), chunk):
            part = lst[c:c + chunk]
            res = refine(ctx, tb.A, [p[1] for p in part], [p[2] for p in part], neighbours=neighbours, steps=steps)
            cols['bank'] += [b] * len(part)
            cols['dataset'] += [d] * len(part)
            cols['index'] += [p[0] for p in part]
            cols['rho_start'].append(res['rho2_start'].sqrt().cpu().numpy())
            cols['rho'].append(res['rho2'].sqrt().cpu().numpy())
            cols['delta'].append(res['delta'].cpu().numpy())
    out = {k: (np.concatenate(v) if v and isinstance(v[0], np.ndarray) else np.asarray(v, np.int64))
           for k, v in cols.items()}
    for k in ('rho_start', 'rho', 'delta'):
        if len(out[k]) == 0:
            out[k] = np.zeros(0)
    out['template_id'] = np.asarray([outs[b][d]['template_id'][i] for b, d, i in
                                     zip(out['bank'], out['dataset'], out['index'])], np.int64)
    out['snr'] = np.asarray([outs[b][d]['snr'][i] for b, d, i in zip(out['bank'], out['dataset'], out['index'])],
                            np.float32)
    out['end_time'] = np.asarray([outs[b][d]['end_time'][i] for b, d, i in
                                  zip(out['bank'], out['dataset'], out['index'])], np.float64)
    return out
