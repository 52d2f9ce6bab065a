"""CPU restatement of the reference's sine-Gaussian chi-squared veto (pycbc.vetoes.sgchisq.SingleDetSGChisq.values as
the reference computes it in float64, fp_filter.py:463-486 sg_chisq_one, with the batch form's tile arithmetic of
sg_kernel.py:66-81 tile_ranges). TEST INFRASTRUCTURE ONLY: see oracle/__init__.py.

Per peak at sample t of a segment (over-whitened data D, 1/psd IP, length N, bin width df) with reference frequency
fpeak and a tile list [(q, offset)]: tile j is centred at fcen = fpeak + offset, spans flo = max(kk df, fcen - 50)
to fcen + 50 Hz with kk the previous tile's first bin (the template's kmin for the first), bins [int(flo / df),
int(fhi / df)), weight g_k = (sqrt(pi) / 2 tau) exp(-(pi tau (f - fcen))^2) (1 + exp(-q^2 f / fcen)) with
tau = q / (2 pi fcen) and g_k = 0 below round(flo / df); sigma_j = sqrt(4 df sum g^2 IP),
z_j = 4 df sum g D e^{2 pi i (k t mod N) / N} / sigma_j, and the value is sum_j |z_j|^2 / (2 ntile). A peak whose
tiles reach above 0.9 x Nyquist (fcen + 50 > 0.9 flen df) is not computed (value 1).

Pinned in tests/test_sgchisq.py by a per-bin Python loop on small inputs and by the closed form: with
D_k = c IP_k g_k e^{-2 pi i k t / N} on one tile's bins, that tile's |z_j|^2 is |c|^2 sigma_j^2.
"""
from __future__ import annotations

import numpy as np


def tile_ranges(fpeak, vals, k_start, df):
    """sg_kernel.py:66-81 / fp_filter.py:472-477: per peak and tile, float64, (kk, k1, kg) int64 and (fcen, tau, q^2),
    each (npk, ntile); each tile's lower edge is bounded by the previous tile's first bin."""
    fpeak = np.asarray(fpeak, np.float64)
    npk, nt = len(fpeak), len(vals)
    out = [np.empty((npk, nt), np.int64) for _ in range(3)] + [np.empty((npk, nt)) for _ in range(3)]
    kk = np.asarray(k_start, np.int64).copy()
    for j, (q, off) in enumerate(vals):
        fcen = fpeak + off
        flo = np.maximum(kk * df, fcen - 50)
        fhi = fcen + 50
        kk = (flo / df).astype(np.int64)
        out[0][:, j], out[1][:, j], out[2][:, j] = kk, (fhi / df).astype(np.int64), np.round(flo / df).astype(np.int64)
        out[3][:, j], out[4][:, j], out[5][:, j] = fcen, q / 2 / np.pi / fcen, q * q
    return out


def computable(fpeak, vals, flen, df):
    """fp_filter.py:468-470: every tile below 0.9 x Nyquist."""
    fpeak = np.asarray(fpeak, np.float64)
    return np.all([fpeak + off + 50 <= flen * df * 0.9 for _, off in vals], axis=0)


def sg_chisq_one(D, IP, t, fpeak, vals, k_start, df, N):
    """fp_filter.py:463-486 in float64 for one peak: the value, or None when not computable. D complex (flen,),
    IP (flen,)."""
    flen = len(D)
    if not computable([fpeak], vals, flen, df)[0]:
        return None
    D = np.asarray(D, np.complex128)
    IP = np.asarray(IP, np.float64)
    kk = int(k_start)
    acc = 0.0
    for q, off in vals:
        fcen = fpeak + off
        flo = max(kk * df, fcen - 50)
        fhi = fcen + 50
        kk = int(flo / df)
        k1 = int(fhi / df)
        kg = int(round(flo / df))
        k = np.arange(kk, k1, dtype=np.int64)
        f = k * df
        tau = q / 2 / np.pi / fcen
        g = (np.pi ** 0.5 / 2 * tau) * np.exp(-(np.pi * tau * (f - fcen)) ** 2) * (1 + np.exp(-q ** 2 * f / fcen))
        g = np.where(k >= kg, g, 0.0)
        gsig = np.sqrt(4 * df * (g * g * IP[kk:k1]).sum())
        ph = ((k * int(t)) % N).astype(np.float64) * (2 * np.pi / N)
        z = (g * D[kk:k1] * np.exp(1j * ph)).sum() * (4.0 * df) / gsig
        acc += abs(z) ** 2
    return acc / (2 * len(vals))


def sg_chisq_loop(D, IP, t, fpeak, vals, k_start, df, N):
    """The same definition as a per-bin Python loop (math module), for small inputs."""
    import cmath
    import math
    flen = len(D)
    if any(fpeak + off + 50 > flen * df * 0.9 for _, off in vals):
        return None
    kk, acc = int(k_start), 0.0
    for q, off in vals:
        fcen = fpeak + off
        flo = max(kk * df, fcen - 50)
        kk, k1, kg = int(flo / df), int((fcen + 50) / df), int(round(flo / df))
        tau = q / (2 * math.pi * fcen)
        s2, z = 0.0, 0j
        for k in range(kk, k1):
            if k < kg:
                continue
            f = k * df
            g = math.sqrt(math.pi) / 2 * tau * math.exp(-(math.pi * tau * (f - fcen)) ** 2) * (1 + math.exp(-q * q * f / fcen))
            s2 += g * g * float(IP[k])
            z += g * complex(D[k]) * cmath.exp(2j * math.pi * ((k * t) % N) / N)
        gsig = math.sqrt(4 * df * s2)
        acc += abs(z * 4 * df / gsig) ** 2
    return acc / (2 * len(vals))
