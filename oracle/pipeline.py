"""The plain method end to end on the CPU (TEST INFRASTRUCTURE ONLY, see oracle/__init__.py): the checker for the
GPU path and the timed CPU baseline. Follows fp_filter.py:700-853 per tile and segment, with the GEMM and the basis
transforms in a chosen precision (float64 for the checker, float32 for the baseline) and the exact stage as the
float64 direct sum over chi-squared bins (tests/test_exact_stage.py:107-120).
"""
from __future__ import annotations

import numpy as np
import scipy.fft

from . import core

MSUN_S = 4.925490947641267e-06


def spa_phase(f, m1, m2, e0=0.0, f0=15.0):
    """TaylorF2 phase to 1PN (t_c = phi_c = 0) with the leading eccentric term of Yunes, Arun, Berti & Will 2009:
    3 / (128 eta v^5) [1 + (3715/756 + 55 eta/9) v^2 - (2355/1462) e0^2 (f/f0)^(-19/9)] - pi/4, v = (pi M f)^(1/3)."""
    M = (m1 + m2) * MSUN_S
    eta = m1 * m2 / (m1 + m2) ** 2
    v = (np.pi * M * f) ** (1.0 / 3.0)
    pn = 1.0 + (3715.0 / 756.0 + 55.0 * eta / 9.0) * v * v
    if e0:
        pn = pn - (2355.0 / 1462.0) * e0 * e0 * (f / f0) ** (-19.0 / 9.0)
    return 3.0 / (128.0 * eta * v ** 5) * pn - np.pi / 4.0


def coarse_grid(flow, fhigh, n_chirp=6000, n_log=1500):
    """imr_interp.py:11-15: nodes uniform in f^(-5/3) joined with a log grid and both ends, inside [flow, fhigh]."""
    u = np.linspace(flow ** (-5 / 3), fhigh ** (-5 / 3), n_chirp) ** (-3 / 5)
    g = np.geomspace(flow, fhigh, n_log)
    x = np.unique(np.r_[u, g, flow, fhigh])
    return x[(x >= flow) & (x <= fhigh)]


def fine_interp(amp, coef, x, df, kmin, kend, flen):
    """imr_interp.py:53-77 (FineInterpolator, numpy branch), float64: the bins k in [kmin, kend) at f = k df clipped
    to [x_0, x_J]; interval i = the last node at or below f (side='right' search), clipped to [0, J - 1];
    d = f - x_i, t = d / (x_i+1 - x_i); phase ((c3 d + c2) d + c1) d + c0 and amplitude amp_i (1 - t) + amp_i+1 t.
    coef (n, J, 4) as (c3, c2, c1, c0) per interval (the reference keeps (n, 4, J), highest order first). Returns
    (n, flen) complex128, zero outside [kmin, kend)."""
    f = np.clip(np.arange(kmin, kend) * df, x[0], x[-1])
    i = np.clip(np.searchsorted(x, f, side='right') - 1, 0, len(x) - 2)
    d = f - x[i]
    t = d / (x[i + 1] - x[i])
    c = coef[:, i, :]
    p = ((c[..., 0] * d + c[..., 1]) * d + c[..., 2]) * d + c[..., 3]
    a = amp[:, i] * (1 - t) + amp[:, i + 1] * t
    out = np.zeros((amp.shape[0], flen), np.complex128)
    out[:, kmin:kend] = a * np.exp(1j * p)
    return out


def spa_bank(m1, m2, e0, flen, df, flow):
    """complex128 (n, flen) templates f^(-7/6) exp(-i psi) on [int(flow/df), min(int(f_isco/df), flen - 1)) and
    their end bins."""
    n = len(m1)
    H = np.zeros((n, flen), np.complex128)
    kmax = np.zeros(n, np.int64)
    kmin = int(flow / df)
    for i in range(n):
        M = (m1[i] + m2[i]) * MSUN_S
        fisco = 1.0 / (6.0 ** 1.5 * np.pi * M)
        kmax[i] = min(int(fisco / df), flen - 1)
        f = np.arange(kmin, kmax[i]) * df
        H[i, kmin:kmax[i]] = f ** (-7.0 / 6.0) * np.exp(-1j * spa_phase(f, m1[i], m2[i], e0[i] if e0 is not None else 0.0))
    return H, kmax


def tiles_by_tau0(m1, m2, flow, tile):
    M = (np.asarray(m1) + np.asarray(m2)) * MSUN_S
    eta = np.asarray(m1) * np.asarray(m2) / (np.asarray(m1) + np.asarray(m2)) ** 2
    tau0 = 5.0 / (256.0 * np.pi * flow * eta) * (np.pi * M * flow) ** (-5.0 / 3.0)
    order = np.argsort(tau0, kind='stable')
    return [order[b0:b0 + tile] for b0 in range(0, len(order), tile)]


def filter_plain(data, m1, m2, e0, thr=5.5, margin=0.04, exact_snr=0.0, eps=1e-4, tile=512, cw=2048, nbins=16,
                 flow=15.0, dtype=np.float64, workers=1, sw=None, timing=None):
    """data: list of dicts with D (nseg, flen) complex64, IP (nseg, flen) float32, analyze, seg_start, tlen, df, sr,
    gps_start. Returns per dataset a dict of trigger arrays (tid, samp, seg, snr, ph, chi, dof, exact) sorted by
    (tid, samp), plus per dataset the candidate lists (tid, seg, samp, approx) for diagnostics."""
    import time
    T = timing if timing is not None else {}
    for k in ('templates_s', 'svd_s', 'basis_fft_s', 'gemm_s', 'exact_s'):
        T.setdefault(k, 0.0)
    d0 = data[0]
    flen, tlen, df = d0['D'].shape[1], d0['tlen'], d0['df']
    kmin = int(flow / df)
    IPall = np.concatenate([d['IP'] for d in data], 0)
    if sw is None:
        sw, _ = core.svd_metric(IPall)
    isw = np.where(sw > 0, 1.0 / np.maximum(sw, 1e-30), 0.0)
    seg0 = np.cumsum([0] + [len(d['analyze']) for d in data])
    outs = [dict(tid=[], samp=[], seg=[], snr=[], ph=[], chi=[], dof=[], exact=[]) for _ in data]
    cands = [[] for _ in data]
    ctype = np.complex128 if dtype == np.float64 else np.complex64
    for ids in tiles_by_tau0(m1, m2, flow, tile):
        t0 = time.perf_counter()
        H, kmax = spa_bank(np.asarray(m1)[ids], np.asarray(m2)[ids], None if e0 is None else np.asarray(e0)[ids],
                           flen, df, flow)
        H = H.astype(np.complex64)
        t1 = time.perf_counter(); T['templates_s'] += t1 - t0
        sig, norm = core.sigmasq(H, IPall, df)
        r, A, Vh, worst = core.svd_tile(H, sw, kmin, eps)
        t2 = time.perf_counter(); T['svd_s'] += t2 - t1
        for di, d in enumerate(data):
            for s, (a0, a1) in enumerate(d['analyze']):
                gs = seg0[di] + s
                nw = (a1 - a0) // cw
                if nw <= 0:
                    continue
                Tn = nw * cw
                t3 = time.perf_counter()
                q = np.zeros((r, tlen), ctype)
                q[:, :flen] = np.conj(Vh) * (d['D'][s] * isw)[None, :]
                zb = (scipy.fft.ifft(q, axis=1, norm='forward', workers=workers))[:, a0:a0 + Tn]
                t4 = time.perf_counter(); T['basis_fft_s'] += t4 - t3
                As = (A * norm[:, gs][:, None]).astype(ctype)
                vw = np.empty((len(ids), nw), dtype); iw = np.empty((len(ids), nw), np.int64)
                step = max(cw, (1 << 22) // max(len(ids), 1) // cw * cw)
                for c0 in range(0, Tn, step):
                    z = As @ zb[:, c0:c0 + step]
                    pw = (z.real * z.real + z.imag * z.imag).astype(dtype)
                    V, I = core.block_absmax(pw, pw.shape[1])
                    w0 = c0 // cw
                    v_, i_ = core.window_max(V, I + c0, pw.shape[1] // cw, cw // core.BN)
                    vw[:, w0:w0 + v_.shape[1]] = v_; iw[:, w0:w0 + v_.shape[1]] = i_
                tb, wb, samp, approx = core.candidates(vw, iw, thr, margin, a0)
                zreb = (As[tb] * zb[:, samp - a0].T).sum(1) if len(tb) else np.zeros(0, complex)
                t5 = time.perf_counter(); T['gemm_s'] += t5 - t4
                cands[di].append((ids[tb], np.full(len(tb), gs), samp, approx))
                o = outs[di]
                ex = approx >= exact_snr * (1 - margin)
                rb = ~ex & (approx > thr)
                if rb.any():
                    o['tid'].append(ids[tb[rb]]); o['samp'].append(samp[rb]); o['seg'].append(np.full(rb.sum(), gs))
                    o['snr'].append(approx[rb]); o['ph'].append(np.angle(zreb[rb])); o['chi'].append(np.zeros(rb.sum()))
                    o['dof'].append(np.zeros(rb.sum(), int)); o['exact'].append(np.zeros(rb.sum(), bool))
                tbx, sx = tb[ex], samp[ex]
                if len(tbx):
                    nbv = np.full(len(tbx), nbins)
                    E = core.edges_per_template(tbx, nbv, kmin, kmax[tbx], H, d['IP'][s])
                    zx, pw = core.dense_float64(H, d['D'][s], tbx, sx, E, np.full(len(ids), nbins), tlen)
                    nr = norm[tbx, gs]
                    keep, snr, chi, dof = core.finalize_exact(zx, pw, nr, nbv, thr)
                    o['tid'].append(ids[tbx[keep]]); o['samp'].append(sx[keep]); o['seg'].append(np.full(keep.sum(), gs))
                    o['snr'].append(snr); o['ph'].append(np.angle(zx[keep])); o['chi'].append(chi); o['dof'].append(dof)
                    o['exact'].append(np.ones(keep.sum(), bool))
                T['exact_s'] += time.perf_counter() - t5
    res = []
    for o in outs:
        c = {k: (np.concatenate(v) if v else np.zeros(0)) for k, v in o.items()}
        srt = np.lexsort((c['samp'], c['tid']))
        res.append({k: v[srt] for k, v in c.items()})
    return res, cands
