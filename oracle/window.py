"""CPU restatement of the reference's overlap-save chunking (window_filter.py:39-47 chunk_segment and the index part of
fp_filter.py:324-357 to_chunks), numpy float64. TEST INFRASTRUCTURE ONLY: see oracle/__init__.py.

Pinned in tests/test_window.py by the property that defines it: for a template supported on [-hist, fut], the direct
matched filter of a chunk, on the chunk's valid samples, equals the direct matched filter of the whole segment at the
same absolute samples (float64, whole-segment vs chunk), with and without band limiting.
"""
from __future__ import annotations

import numpy as np


def chunk_starts(tlen, Lc, hist, fut, lo, hi):
    """window_filter.py:16-26."""
    lo, hi = max(lo, hist), min(hi, tlen - fut)
    s = np.arange(lo - hist, hi - hist, Lc - hist - fut)
    s[-1] = min(s[-1], tlen - Lc)
    return s


def valid(starts, Lc, hist, fut, hi):
    """window_filter.py:29-36."""
    a = starts + hist
    b = np.minimum(np.append(a[1:], starts[-1] + Lc - fut), hi)
    return a, np.maximum(b, a)


def chunk_segment(D, IP, tlen, dt, starts, Lc):
    """window_filter.py:39-47 in float64: d = irfft(D) / dt; per chunk Dc = dt rfft(d[s : s + Lc]); IP every
    tlen / Lc bins. Returns (nch, Lc // 2 + 1) complex128 and float64."""
    step = tlen // Lc
    assert step * Lc == tlen
    d = np.fft.irfft(np.asarray(D, np.complex128), n=tlen) / dt
    idx = np.asarray(starts)[:, None] + np.arange(Lc)[None, :]
    Dc = dt * np.fft.rfft(d[idx], axis=1)
    IPc = np.broadcast_to(np.asarray(IP, np.float64)[::step][:Lc // 2 + 1], Dc.shape)
    return Dc, IPc


def to_chunks(D, IP, analyze, seg_start, tlen, sr, spec):
    """fp_filter.py:324-357 for one dataset (D, IP: (nseg, flen)): spec 'LEN,HIST,FUT[,FMAX]' seconds; with FMAX the
    chunks keep the bins below FMAX (sampled at 2 FMAX, decimation dec = round(sr / (2 FMAX))). Returns a dict with
    D, IP (nch, nb), analyze and seg_start per chunk (in decimated samples), tlen, flen, df, sr."""
    v = [float(x) for x in spec.split(',')]
    Lc, hist, fut = int(v[0] * sr), int(v[1] * sr), int(v[2] * sr)
    dec = int(round(sr / (2 * v[3]))) if len(v) > 3 else 1
    assert dec >= 1 and Lc % (2 * dec) == 0 and hist % dec == 0 and fut % dec == 0
    nb = Lc // (2 * dec) + 1
    Ds, IPs, an, ss = [], [], [], []
    for s, (a0, a1) in enumerate(analyze):
        st = chunk_starts(tlen, Lc, hist, fut, a0, a1)
        va, vb = valid(st, Lc, hist, fut, a1)
        assert np.all(st % dec == 0) and np.all(va % dec == 0) and np.all(vb % dec == 0) and seg_start[s] % dec == 0
        Dc, IPc = chunk_segment(D[s], IP[s], tlen, 1 / sr, st, Lc)
        Ds.append(Dc[:, :nb]); IPs.append(IPc[:, :nb])
        an += [(int(x - y) // dec, int(z - y) // dec) for x, z, y in zip(va, vb, st)]
        ss += [int(seg_start[s] + y) // dec for y in st]
    return dict(D=np.concatenate(Ds), IP=np.concatenate(IPs), analyze=an, seg_start=ss, tlen=Lc // dec, flen=nb,
                df=sr / Lc, sr=sr / dec, dec=dec)
