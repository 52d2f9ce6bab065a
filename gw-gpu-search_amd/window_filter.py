"""Overlap-save chunks of a conditioned segment (the reference's window_filter.py:1-47, imported by fp_filter for
--chunk / --chunk2): chunk starts whose valid ranges [s + hist, s + Lc - fut) tile the analysed span, the range each
chunk alone reports, and the over-whitened chunks Dc = dt rfft(d[s : s + Lc]) with 1/psd sampled at the chunk
resolution. A template supported on [-hist, fut] gives, on a chunk's valid samples, the same SNR as on the whole
segment. to_chunks turns datasets into chunk datasets (fp_filter.py:324-357), optionally band-limited. Index
arithmetic is host integer code; the transforms are hipFFT (torch.fft) on the device."""
from __future__ import annotations

import numpy as np
import torch


def chunk_starts(tlen, Lc, hist, fut, lo=None, hi=None):
    """Chunk starts whose valid ranges tile [lo, hi) (default: every sample with hist of history and fut of future);
    the last chunk is moved back to end at the segment end (window_filter.py:16-26)."""
    lo = hist if lo is None else max(lo, hist)
    hi = tlen - fut if hi is None else min(hi, tlen - fut)
    V = Lc - hist - fut
    assert V > 0 and hi > lo
    s = np.arange(lo - hist, hi - hist, V)
    s[-1] = min(s[-1], tlen - Lc)
    return s


def valid(starts, Lc, hist, fut, hi=None):
    """Per chunk the half-open range [a, b) it alone reports: no overlaps, nothing at or after hi
    (window_filter.py:29-36)."""
    a = starts + hist
    b = np.append(a[1:], starts[-1] + Lc - fut)
    if hi is not None:
        b = np.minimum(b, hi)
    return a, np.maximum(b, a)


def chunk_segment(D, IP, tlen, dt, starts, Lc, xp=None):
    """(nch, Lc//2 + 1) over-whitened chunks Dc = dt rfft(d[s : s + Lc]) (d = irfft(D) / dt) and IP sampled every
    tlen / Lc bins (window_filter.py:39-47). D, IP: torch tensors on the GPU; xp is accepted for the reference's
    signature."""
    step = tlen // Lc
    assert step * Lc == tlen, 'chunk length must divide the segment length'
    d = torch.fft.irfft(D.to(torch.complex64), n=tlen) / dt
    st = torch.as_tensor(np.asarray(starts), device=d.device)
    idx = st[:, None] + torch.arange(Lc, device=d.device)[None, :]
    Dc = (dt * torch.fft.rfft(d[idx], dim=1)).to(torch.complex64)
    IPc = IP[::step][:Lc // 2 + 1].expand(Dc.shape[0], Lc // 2 + 1)
    return Dc, IPc


def to_chunks(datasets, spec):
    """fp_filter.py:324-357 (--chunk / --chunk2 LEN,HIST,FUT[,FMAX] seconds): every segment of every dataset becomes
    overlap-save chunks of its over-whitened series, and from then on a chunk is a segment. With FMAX the chunks keep
    the bins below FMAX only, i.e. are sampled at 2 FMAX (decimation dec = round(sr / (2 FMAX)); the same SNR for
    templates without power above FMAX). Returns (chunk datasets, HIST in seconds); the templates filtered on them
    must be shorter than HIST."""
    import fp_filter
    v = [float(x) for x in spec.split(',')]
    c_len, c_hist, c_fut = v[:3]
    out = []
    for m in datasets:
        sr = m.sr
        Lc, hist, fut = int(c_len * sr), int(c_hist * sr), int(c_fut * sr)
        dec = int(round(sr / (2 * v[3]))) if len(v) > 3 else 1
        assert dec >= 1 and Lc % (2 * dec) == 0 and hist % dec == 0 and fut % dec == 0
        nb = Lc // (2 * dec) + 1
        Ds, IPs, an, ss, first = [], [], [], [], []
        for s, (a0, a1) in enumerate(m.analyze):
            first.append(len(an))
            st = chunk_starts(m.tlen, Lc, hist, fut, a0, a1)
            va, vb = valid(st, Lc, hist, fut, a1)
            assert np.all(st % dec == 0) and np.all(va % dec == 0) and np.all(vb % dec == 0)
            assert m.seg_start[s] % dec == 0
            Dc, IPc = chunk_segment(m.D[s], m.IP[s], m.tlen, 1 / sr, st, Lc)
            Ds.append(Dc[:, :nb].contiguous())
            IPs.append(IPc[:, :nb].contiguous())
            an += [(int(x - y) // dec, int(z - y) // dec) for x, z, y in zip(va, vb, st)]
            ss += [int(m.seg_start[s] + y) // dec for y in st]
        out.append(fp_filter.Dataset(m.tag, torch.cat(Ds), torch.cat(IPs), nb, Lc // dec, sr / Lc, sr / dec,
                                     m.gps_start, an, ss, list(m.gates), first))
    return out, c_hist
