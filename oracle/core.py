"""CPU restatement of the reference's filtering path, function by function (numpy; float64 unless a step says
otherwise). TEST INFRASTRUCTURE ONLY: see oracle/__init__.py.

Shared conventions (reference gpu_prep.py:21, 380, 421, 448-451; fp_filter.py:606-608, 711-714):
  D = stilde * IP with stilde = (1/sr) rfft(strain * DYN_RANGE_FAC) and IP = 1/psd (0 below kmin);
  sigma^2 = 4 df sum |h_k|^2 IP_k, norm = 4 df / sigma;
  z(t) = sum_{k < flen} conj(h_k) D_k exp(+2 pi i k t / N) (unnormalised one-sided inverse DFT), snr = |z| norm.
"""
from __future__ import annotations

import numpy as np

BN = 128                 # GEMM output block (fp_gemm.py:11)


# ---------------------------------------------------------------------------------------------------------------------
# Conditioning (gpu_prep.py)
# ---------------------------------------------------------------------------------------------------------------------

def median_bias(n):
    """gpu_prep.py:47-53: the bias of the median of n chi^2_2 variables, 1 + sum_i 1/(2i+1) - 1/(2i); ln 2 for
    n >= 1000."""
    if n >= 1000:
        return np.log(2)
    ans = 1.0
    for i in range(1, (n - 1) // 2 + 1):
        ans += 1.0 / (2 * i + 1) - 1.0 / (2 * i)
    return ans


def welch(x, seg_len, seg_stride, dt, num_segments=None, avg='median'):
    """gpu_prep.py:87-125 (pycbc.psd.welch, symmetric Hann window, centred data fit), float64 throughout. x: (n,) or
    (rows, n). Returns (psd, delta_f)."""
    X = np.atleast_2d(np.asarray(x, np.float64))
    n = X.shape[1]
    if num_segments is None:
        num_segments = int(n // seg_stride)
        if (num_segments - 1) * seg_stride + seg_len > n:
            num_segments -= 1
    data_len = (num_segments - 1) * seg_stride + seg_len
    if data_len < n:
        diff = n - data_len
        start = diff // 2 + (diff % 2)
        X = X[:, start:n - diff // 2]
    assert X.shape[1] == data_len
    w = np.hanning(seg_len)
    idx = (np.arange(num_segments) * seg_stride)[:, None] + np.arange(seg_len)[None, :]
    out = []
    for row in X:
        st = np.fft.rfft(row[idx] * w, axis=1) * dt
        p = st.real ** 2 + st.imag ** 2
        p[:, 0] /= 2
        p[:, -1] /= 2
        q = np.median(p, axis=0) / median_bias(num_segments) if avg == 'median' else p.mean(0)
        out.append(q)
    delta_f = 1.0 / dt / seg_len
    fac = 2 * delta_f * seg_len / (w * w).sum()
    P = np.stack(out) * fac
    return (P[0] if np.ndim(x) == 1 else P), delta_f


def interpolate(psd, df_old, df_new, length=None):
    """gpu_prep.py:128-135 (pycbc.psd.interpolate): np.interp in float64 onto k df_new."""
    P = np.atleast_2d(psd)
    new_n = (P.shape[-1] - 1) * df_old / df_new + 1 if length is None else length
    samples = np.arange(0, np.rint(new_n)) * df_new
    fo = np.arange(P.shape[-1]) * df_old
    out = np.stack([np.interp(samples, fo, r.astype(np.float64)) for r in P])
    return out[0] if np.ndim(psd) == 1 else out


def inverse_spectrum_truncation(psd, df, max_filter_len, flow, trunc_method=None):
    """gpu_prep.py:138-158 (pycbc.psd.inverse_spectrum_truncation) in float64: inv_asd = psd^-1/2 on [kmin, N/2)
    (zero elsewhere, DC and Nyquist included), its unnormalised irfft times df, zeroed outside +-max_filter_len/2
    (optional Hann taper on the kept ends), rfft / (df N), then 1 / |.|^2."""
    P = np.atleast_2d(np.asarray(psd, np.float64))
    flen = P.shape[1]
    N = (flen - 1) * 2
    kmin = int(flow / df) if flow else 1
    inv_asd = np.zeros(P.shape, np.complex128)
    inv_asd[:, kmin:N // 2] = (1.0 / P[:, kmin:N // 2]) ** 0.5
    q = np.fft.irfft(inv_asd, n=N, axis=1, norm='forward') * df
    ts, te = max_filter_len // 2, N - max_filter_len // 2
    if trunc_method == 'hann':
        tw = np.hanning(max_filter_len)
        q[:, 0:ts] *= tw[-ts:]
        q[:, te:N] *= tw[0:max_filter_len // 2]
    if ts < te:
        q[:, ts:te] = 0
    pt = np.fft.rfft(q, axis=1) / (df * N)
    out = 1.0 / np.abs(pt * pt.conj())
    return out[0] if np.ndim(psd) == 1 else out


def segment_slices(n, sr, start_gps, seg_len, start_pad, end_pad, trig_start=0, trig_end=0):
    """gpu_prep.py:234-271 (pycbc.strain.StrainSegments index arithmetic). Returns [((s0, s1), (a0, a1))]."""
    end_gps = start_gps + n / sr
    if not trig_start:
        trig_start = int(start_gps) + start_pad
    if not trig_end:
        trig_end = int(end_gps) - end_pad
    assert trig_start >= int(start_gps) + start_pad and trig_end <= int(end_gps) - end_pad
    seg_width = seg_len - start_pad - end_pad
    analyzable = trig_end - trig_start
    data_start = (trig_start - start_pad) - int(start_gps)
    data_end = trig_end + end_pad - int(start_gps)
    data_dur = data_end - data_start
    data_start, data_end = data_start * sr, data_end * sr
    num_segs = int(np.ceil(float(analyzable) / float(seg_width)))
    seg_offset = int(np.ceil(analyzable / float(num_segs)))
    segs, anas = [], []
    for k in range(num_segs - 1):
        s0 = int(data_start + (k * seg_offset) * sr)
        segs.append((s0, int(s0 + seg_len * sr)))
        a0 = int(start_pad * sr)
        anas.append((a0, int(a0 + seg_offset * sr)))
    s1 = int(data_end)
    segs.append((int(s1 - seg_len * sr), s1))
    remaining = data_dur - ((num_segs - 1) * seg_offset + start_pad)
    anas.append((int((seg_len - remaining) * sr), int((seg_len - end_pad) * sr)))
    ts_i = (trig_start - int(start_gps)) * sr
    te_i = (trig_end - int(start_gps)) * sr
    out = []
    for (s0, s1), (a0, a1) in zip(segs, anas):
        if ts_i > a0 + s0:
            a0 += ts_i - (a0 + s0)
        if te_i < a1 + s0:
            a1 -= (a1 + s0) - te_i
        if a0 < a1:
            assert s0 >= 0 and s1 <= n
            out.append(((s0, s1), (a0, a1)))
    return out


def overlapping_psd_spans(n, seg_len, seg_stride, num_segments):
    """gpu_prep.py:274-282 (pycbc.psd.generate_overlapping_psds): (start, end) sample spans of the PSDs."""
    psd_len = (num_segments - 1) * seg_stride + seg_len
    assert n >= psd_len
    if n == psd_len:
        return [(0, n)]
    m = int(2 * (n - 1) / psd_len)
    stride = int((n - psd_len) / m)
    return [(n - psd_len, n) if i == m - 1 else (stride * i, psd_len + stride * i) for i in range(m)]


def pick_psd(segs, spans):
    """gpu_prep.py:438-447: per segment the PSD span of largest overlap, first wins ties."""
    pick = []
    for (a, b), _ in segs:
        best, ov = None, 0
        for j, (c, e) in enumerate(spans):
            o = min(b, e) - max(a, c)
            if o > 0 and o > ov:
                best, ov = j, o
        pick.append(best)
    return pick


def condition(s, sr, gps_start, seg_len, start_pad, end_pad, psd_seg_len, psd_stride, psd_nseg, psd_inv_len, flow,
              trunc_method=None, avg='median'):
    """gpu_prep.py:405-454 (gpu_stage2 after padding removal), with the PSD in float64. s: float32 strain already
    x DYN_RANGE_FAC. Returns dict(D complex64 (nseg, flen), IP float32, flen, tlen, df, analyze, seg_start, psd)."""
    n = len(s)
    seg = segment_slices(n, sr, gps_start, seg_len, start_pad, end_pad)
    tlen = int(seg_len * sr)
    flen = tlen // 2 + 1
    df = 1.0 / seg_len
    idx = np.array([a for (a, _), _ in seg])[:, None] + np.arange(tlen)[None, :]
    stilde = np.fft.rfft(s[idx].astype(np.float64), axis=1) / sr
    pl, ps = int(psd_seg_len * sr), int(psd_stride * sr)
    spans = overlapping_psd_spans(n, pl, ps, psd_nseg)
    parts = np.stack([s[a:b] for a, b in spans]).astype(np.float64)
    psd, pdf = welch(parts, pl, ps, 1.0 / sr, num_segments=psd_nseg, avg=avg)
    if pdf != df:
        psd = interpolate(psd, pdf, df, flen)
    if psd_inv_len:
        psd = inverse_spectrum_truncation(psd, df, int(psd_inv_len * sr), flow, trunc_method)
    P = psd[np.array(pick_psd(seg, spans))]
    kmin = int(flow / df)
    ip = np.where(P > 0, 1.0 / np.where(P > 0, P, 1.0), 0.0)
    ip[:, :kmin] = 0
    D = (stilde * ip).astype(np.complex64)
    return dict(D=D, IP=ip.astype(np.float32), flen=flen, tlen=tlen, df=df, sr=float(sr), gps_start=float(gps_start),
                analyze=[a for _, a in seg], seg_start=[s0 for (s0, _), _ in seg], psd=psd, stilde=stilde)


def resample_ldas(x, factor):
    """gpu_prep.py:67-84 (pycbc resample_to_delta_t, method 'ldas'): causal FIR with scipy's firwin(20 factor + 1,
    1 / factor, Kaiser beta 5) taps, the first 2 (ntaps // 2) outputs zeroed, rolled left by ceil(ntaps / 2),
    decimated: out[i] = sum_k c[k] x[factor i + ceil(ntaps / 2) - k], float64, zero where that index is below
    2 (ntaps // 2) or past the end."""
    import scipy.signal
    x = np.asarray(x, np.float64)
    ntaps = factor * 20 + 1
    c = scipy.signal.firwin(ntaps, 1.0 / factor, window=('kaiser', 5))
    L = len(x)
    n = np.arange(0, L, factor) + (ntaps + 1) // 2
    ok = (n >= 2 * (ntaps // 2)) & (n < L)
    out = np.zeros(len(n))
    acc = np.zeros(int(ok.sum()))
    for k in range(ntaps):
        acc += c[k] * x[n[ok] - k]
    out[ok] = acc
    return out


def cluster_over_window(idx, vals, window):
    """gpu_prep.py:161-172 (findchirp_cluster_over_window): keep the loudest of each run closer than window to the
    run's current keeper."""
    if len(idx) == 0:
        return np.zeros(0, dtype=int)
    keep = [0]
    for i in range(1, len(idx)):
        if idx[i] < idx[keep[-1]] + window:
            if vals[i] > vals[keep[-1]]:
                keep[-1] = i
        else:
            keep.append(i)
    return np.array(keep)


def whiten_for_gating(strain, sr, flow, corrupt_time, psd_duration=4.0, psd_stride=2.0):
    """gpu_prep.py:175-204 up to |whitened| (pycbc.strain.detect_loud_glitches), float64: taper the corrupt ends,
    zero-pad centred to the next power of two, whiten by a 4 s median-Welch PSD truncated with a Hann taper and
    normalised to unit variance. Returns |w| over the original n samples with the corrupt ends zeroed."""
    dt = 1.0 / sr
    s = np.asarray(strain, np.float64).copy()
    n = len(s)
    cl = int(corrupt_time * sr)
    w = np.arange(cl) / float(cl)
    s[0:cl] *= w
    s[n - cl:] *= w[::-1]
    L = 1 << n.bit_length()
    pad_start = int(L / 2 - n / 2)
    sp = np.zeros(L)
    sp[pad_start:pad_start + n] = s
    psd, pdf = welch(s[cl:n - cl], int(psd_duration * sr), int(psd_stride * sr), dt)
    df = 1.0 / (L * dt)
    psd = interpolate(psd, pdf, df)
    psd = inverse_spectrum_truncation(psd, df, int(psd_duration * sr), flow, 'hann')
    psd[0:int(flow / df)] = np.inf
    norm = sr / 2.0 - flow
    st = np.fft.rfft(sp) * dt
    st *= (psd * norm) ** (-0.5)
    wt = np.fft.irfft(st, n=L, norm='forward') * df
    mag = np.abs(wt[pad_start:pad_start + n])
    mag[0:cl] = 0
    mag[n - cl:] = 0
    return mag


def detect_loud_glitches(strain, sr, start_gps, threshold, cluster_window, flow, corrupt_time):
    """gpu_prep.py:175-209 with float64 GPS times in place of LIGOTimeGPS: times of the clustered |whitened| >
    threshold samples."""
    mag = whiten_for_gating(strain, sr, flow, corrupt_time)
    ind = np.nonzero(mag > threshold)[0]
    keep = cluster_over_window(ind.astype(np.int32), mag[ind], int(cluster_window * sr))
    return [start_gps + int(i) / sr for i in ind[keep]]


def gate_window(sr, gw, pw):
    """gpu_prep.py:222-225: inverted Tukey, pad_samples cosine ramps around win_samples - 2 pad_samples zeros."""
    win_samples = int(2 * sr * (gw + pw))
    pad_samples = int(sr * pw)
    pad = 0.5 * (1.0 + np.cos(np.pi * np.arange(pad_samples) / pad_samples))
    return np.concatenate((pad, np.zeros(win_samples - 2 * pad_samples), pad[::-1]))


def gate_taper(x, sr, start_gps, gates):
    """gpu_prep.py:212-231 (pycbc.strain.gate_data), in place on float32 x; product in float64, then rounded."""
    duration = len(x) / sr
    for gt, gw, pw in gates:
        t_start = gt - gw - pw - start_gps
        t_end = gt + gw + pw - start_gps
        if t_start > duration or t_end < 0.0:
            continue
        window = gate_window(sr, gw, pw)
        offset = int(t_start * sr)
        i1, i2 = max(0, -offset), min(len(window), len(x) - offset)
        sl = slice(i1 + offset, i2 + offset)
        x[sl] = (x[sl].astype(np.float64) * window[i1:i2]).astype(x.dtype)
    return x


# ---------------------------------------------------------------------------------------------------------------------
# Templates and the compressed filter (fp_filter.py, fp_gemm.py)
# ---------------------------------------------------------------------------------------------------------------------

def sigmasq(H, IPall, df):
    """fp_filter.py:710-714: sig = 4 df (|H|^2 @ IPall^T) (float64 here), norm = 4 df / sqrt(sig)."""
    Hd = np.asarray(H, np.complex128)
    sig = 4.0 * df * ((Hd.real ** 2 + Hd.imag ** 2) @ np.asarray(IPall, np.float64).T)
    return sig, 4.0 * df / np.sqrt(sig)


def svd_metric(IPall):
    """fp_filter.py:363-364: sw = sqrt(mean IP over all segments), isw = 1/sw (0 where sw = 0)."""
    sw = np.sqrt(np.asarray(IPall, np.float64).mean(0))
    isw = np.where(sw > 0, 1.0 / np.maximum(sw, 1e-30), 0.0)
    return sw, isw


def svd_tile(H, sw, kmin, eps):
    """fp_filter.py:576-593 in complex128: Gram of the sw-weighted rows, normalised; eigh; the worst lost power after
    k + 1 vectors over the tile's templates; r = #(worst > eps) + 1. Returns r, A (n, r), Vh (r, flen), worst[r-1],
    with z = A @ IFFT(conj(Vh) D / sw)."""
    Hw = np.asarray(H, np.complex128) * sw[None, :]
    Hw[:, :kmin] = 0
    G = Hw @ Hw.conj().T
    nrm = np.sqrt(np.maximum(G.diagonal().real, 1e-30))
    lam, W = np.linalg.eigh(G / (nrm[:, None] * nrm[None, :]))
    lam, W = np.maximum(lam[::-1], 0), W[:, ::-1]
    worst = (1 - np.cumsum(np.abs(W) ** 2 * lam[None, :], axis=1)).max(0)
    n = H.shape[0]
    r = min(int((worst > eps).sum()) + 1, n)
    Wr = W[:, :r]
    s_r = np.sqrt(np.maximum(lam[:r], 1e-12 * lam[0]))
    Bm = (Wr.conj().T / s_r[:, None]) / nrm[None, :]
    Vh = Bm @ Hw
    A = (Wr * s_r[None, :]).conj() * nrm[:, None]
    return r, A, Vh, float(worst[r - 1])


def basis_series(Vh, Dws, a0, Tn, N):
    """fp_filter.py:597-608 in complex128: zb[j, t] = sum_{k<flen} conj(Vh[j,k]) Dws[k] e^{2 pi i k t / N} for
    t in [a0, a0 + Tn). Dws = D / sw."""
    q = np.zeros((Vh.shape[0], N), np.complex128)
    q[:, :Vh.shape[1]] = np.conj(Vh) * Dws[None, :]
    return (np.fft.ifft(q, axis=1) * N)[:, a0:a0 + Tn]


def stacked_A(A, nrm_s, scale, Kp):
    """fp_filter.py:615-621 (float64 here): As = A (norm_i scale); Are = [Re As, -Im As, 0...], Aim = [Im As, Re As,
    0...], so z = Are @ Z + i Aim @ Z for Z = [Re zb; Im zb] / scale (rows padded with zeros to Kp)."""
    r = A.shape[1]
    As = A * (nrm_s * scale)[:, None]
    Are = np.zeros((A.shape[0], Kp)); Aim = np.zeros_like(Are)
    Are[:, :r] = As.real; Are[:, r:2 * r] = -As.imag
    Aim[:, :r] = As.imag; Aim[:, r:2 * r] = As.real
    return Are, Aim


def stacked_Z(zb, scale, Kp):
    """fp_filter.py:609-612 (without the fp16 cast): rows 0..r-1 = Re zb / scale, r..2r-1 = Im zb / scale."""
    r, Tn = zb.shape
    Z = np.zeros((Kp, Tn))
    Z[:r] = zb.real / scale
    Z[r:2 * r] = zb.imag / scale
    return Z


def block_absmax(pw, Tn, bn=BN):
    """fp_gemm.py:29-33 epilogue given the |z|^2 array pw (M, >= Tn): columns >= Tn masked to -1, per block of bn
    columns the max and its first index (pid_n bn + argmax)."""
    M = pw.shape[0]
    nblk = -(-Tn // bn)
    P = np.full((M, nblk * bn), -1.0, dtype=pw.dtype)
    P[:, :Tn] = pw[:, :Tn]
    P = P.reshape(M, nblk, bn)
    j = P.argmax(2)                                   # numpy argmax: first index of the max
    V = np.take_along_axis(P, j[..., None], 2)[..., 0]
    I = (np.arange(nblk)[None, :] * bn + j).astype(np.int32)
    return V, I


def gemm_absmax(Are, Aim, Z, Tn, bn=BN, dtype=np.float64):
    """fp_gemm.py:14-44 (gemm_absmax / fused_peaks): z_re = Are @ Z, z_im = Aim @ Z in dtype, pw = z_re^2 + z_im^2,
    then block_absmax. Returns V (M, ceil(Tn/bn)), I int32."""
    zr = np.asarray(Are, dtype) @ np.asarray(Z[:, :Tn], dtype)
    zi = np.asarray(Aim, dtype) @ np.asarray(Z[:, :Tn], dtype)
    return block_absmax(zr * zr + zi * zi, Tn, bn)


def window_max(V, I, nw, per):
    """fp_filter.py:745-746: per cluster window (per = cw / BN blocks) the max of the block maxima and the index of
    that block's argmax; the first block wins ties (the same rule as peak_kernel.py:21-33). Trailing blocks beyond
    nw windows are dropped."""
    n = V.shape[0]
    Vw = V[:, :nw * per].reshape(n, nw, per)
    jj = Vw.argmax(2)
    vw = np.take_along_axis(Vw, jj[..., None], 2)[..., 0]
    iw = np.take_along_axis(I[:, :nw * per].reshape(n, nw, per), jj[..., None], 2)[..., 0]
    return vw, iw


def candidates(vw, iw, thr, margin, a0):
    """fp_filter.py:747-753: windows with vw > (thr (1 - margin))^2 (strict), row-major order; for float32 vw the
    float64 square is rounded to float32 (torch's compare of a float32 tensor with a Python float). Returns tb (row),
    wb (window), samp = a0 + iw, approx = sqrt(vw) (float32)."""
    cut = np.float32((thr * (1 - margin)) ** 2) if vw.dtype == np.float32 else (thr * (1 - margin)) ** 2
    tb, wb = np.nonzero(vw > cut)
    return tb, wb, (a0 + iw[tb, wb]).astype(np.int64), np.sqrt(vw[tb, wb])


def window_peaks(z, a0, cw, nwin):
    """peak_kernel.py:8-47: per (row, window) max |z|^2 over z[row, a0 + w cw : a0 + (w+1) cw] and its index within
    the window, first index on ties. z complex; |z|^2 in z's real precision."""
    seg = z[:, a0:a0 + nwin * cw].reshape(z.shape[0], nwin, cw)
    p = seg.real * seg.real + seg.imag * seg.imag
    j = p.argmax(2)
    return np.take_along_axis(p, j[..., None], 2)[..., 0], j.astype(np.int32)


# ---------------------------------------------------------------------------------------------------------------------
# Exact values at peaks: chi-squared edges and bin sums (multi_filter.py, chisq_kernel.py, exact_stage.py)
# ---------------------------------------------------------------------------------------------------------------------

def edges_per_template(rows, nb_, kmin, k1, Hd, ips):
    """multi_filter.py:264-282: edge j = first k with cumsum_k(|h|^2 IP) / total > j / nb (searchsorted right on
    the float64 cumsum, total = cumsum[k1 - 1]); clipped to [kmin, k1]; columns j >= nb and the last column are k1.
    Returns (npk, B + 1) int64."""
    rows = np.asarray(rows); nb_ = np.asarray(nb_, np.int64); k1 = np.asarray(k1, np.int64)
    B = int(nb_.max())
    out = np.empty((len(rows), B + 1), np.int64)
    ipd = np.asarray(ips, np.float64)
    cache = {}
    for i, (r, nb, kk) in enumerate(zip(rows, nb_, k1)):
        if r not in cache:
            h = np.asarray(Hd[r], np.complex128)
            cache[r] = np.cumsum((h.real ** 2 + h.imag ** 2) * ipd)
        cs = cache[r]
        tot = cs[kk - 1]
        j = np.arange(B)
        e = np.searchsorted(cs / tot, j / nb, side='right')
        e = np.clip(e, kmin, kk)
        out[i, :B] = np.where(j < nb, e, kk)
        out[i, B] = kk
    return out


def bin_sums(q, k0, t, edges, N):
    """Per bin j of edges (e_0 < ... < e_nb): Z_j = sum_{e_j <= k < e_j+1} q[k - k0] exp(2 pi i ((k t) mod N) / N)
    in complex128 (tests/test_chisq_kernel_gpu.py:15-22; tests/test_exact_stage.py:107-120)."""
    e = np.asarray(edges, np.int64)
    k = np.arange(e[0], e[-1], dtype=np.int64)
    x = np.asarray(q[e[0] - k0:e[-1] - k0], np.complex128) * np.exp(2j * np.pi * ((k * int(t)) % N) / N)
    b = np.searchsorted(e, k, side='right') - 1
    nb = len(e) - 1
    return np.bincount(b, x.real, nb) + 1j * np.bincount(b, x.imag, nb)


def bin_power_sums(q, rows, times, edges_per_peak, want_z=False, d=None, n_time=None):
    """chisq_kernel.py:85-121 semantics in float64: per peak sum_j |Z_j(t)|^2 (and z = sum_j Z_j) with
    Z_j = sum over bin j of q[row, k] e^{2 pi i k t / N}; in fused mode q rows are templates h and the summand is
    conj(h_k) d_k. Empty bins (e_j >= e_j+1) contribute nothing. Request order."""
    N = q.shape[1] if d is None else n_time
    npk = len(rows)
    pw = np.zeros(npk); z = np.zeros(npk, np.complex128)
    for i in range(npk):
        row = np.asarray(q[rows[i]], np.complex128)
        qq = row if d is None else np.conj(row) * np.asarray(d, np.complex128)
        e = np.asarray(edges_per_peak[i], np.int64)
        e = np.maximum.accumulate(e)
        Z = bin_sums(qq, 0, times[i], e, N) if e[-1] > e[0] else np.zeros(1, complex)
        pw[i] = (np.abs(Z) ** 2).sum(); z[i] = Z.sum()
    return (pw, z) if want_z else pw


def dense_float64(H, D, rows, samples, E, nb, N):
    """tests/test_exact_stage.py:107-120, the definition: Z_j over the bins of E[i, :max(nb,1)+1] of
    conj(H_k) D_k e^{2 pi i k t / N} in float64; z = sum Z_j, pw = sum |Z_j|^2."""
    z = np.empty(len(rows), np.complex128); pw = np.empty(len(rows))
    for i, (r, t) in enumerate(zip(rows, samples)):
        e = np.asarray(E[i, :max(int(nb[r]), 1) + 1], np.int64)
        q = np.conj(np.asarray(H[r], np.complex128)) * np.asarray(D, np.complex128)
        Z = bin_sums(q, 0, t, e, N)
        z[i] = Z.sum(); pw[i] = (np.abs(Z) ** 2).sum()
    return z, pw


def fpeak(E, nb, df):
    """exact_stage.py:236-238; fp_filter.py:826-827: (2 e[nb-1] - e[nb-2]) df."""
    r = np.arange(len(E))
    nb = np.asarray(nb, np.int64)
    return (2 * E[r, np.maximum(nb - 1, 0)] - E[r, np.maximum(nb - 2, 0)]) * df


def newsnr(snr, rchisq):
    """coinc_rank.py:39-44 (pycbc newsnr): snr / ((1 + rchisq^3) / 2)^(1/6) where rchisq > 1."""
    ns = np.array(snr, dtype=np.float64)
    rchisq = np.asarray(rchisq, np.float64)
    m = rchisq > 1
    ns[m] = ns[m] / ((1 + rchisq[m] ** 3) / 2.0) ** (1.0 / 6)
    return ns


def finalize_exact(zx, pw, nr, nbins, thr):
    """fp_filter.py:837-853: snr = |zx| norm; keep snr > thr (strict); chisq = (nb pw - |zx|^2) norm^2 (float32);
    dof = 2 nb - 2. Returns keep mask and (snr, chisq, dof) of the kept peaks."""
    snr = np.abs(zx) * nr
    keep = snr > thr
    chi = ((nbins * pw - np.abs(zx) ** 2) * nr ** 2).astype(np.float32)
    dof = (2 * nbins - 2).astype(np.int32)
    return keep, snr[keep], chi[keep], dof[keep]


def direct_z(H, D, N):
    """FP64 direct frequency-domain matched filter numerator (multi_filter.py:203-204; tests/test_exact_stage.py:86-89):
    z[i, t] = sum_{k < flen} conj(H[i,k]) D[k] e^{2 pi i k t / N}, all t in [0, N). complex128 (rows, N)."""
    H = np.atleast_2d(H)
    q = np.zeros((H.shape[0], N), np.complex128)
    q[:, :H.shape[1]] = np.conj(np.asarray(H, np.complex128)) * np.asarray(D, np.complex128)[None, :]
    return np.fft.ifft(q, axis=1) * N


def find_pairs(H, L, P, step, K, window):
    """coinc_rank.py:100-138 (find_pairs) by brute force: every (h, l, k) of the same template with
    |tH - (tL + k step) mod P| <= window, |k| <= K, taking the unwrapped L time among tL - P, tL, tL + P within
    K step + window of tH. Returns (h, l, k, dt = k step - (tH - tL)) sorted by (h, unwrapped tL)."""
    out = []
    reach = K * step + window
    for h, (th, tidh) in enumerate(zip(H['x'], H['tid'])):
        cand = []
        for l, (tl, tidl) in enumerate(zip(L['x'], L['tid'])):
            if tidl != tidh:
                continue
            for tu in (tl - P, tl, tl + P):
                if abs(th - tu) <= reach:
                    dt = th - tu
                    k = np.rint(dt / step)
                    if abs(dt - k * step) <= window and abs(k) <= K:
                        cand.append((tu, l, int(k), k * step - dt))
        cand.sort(key=lambda c: c[0])
        out += [(h, l, k, d) for _, l, k, d in cand]
    a = np.array(out, dtype=float).reshape(-1, 4)
    return a[:, 0].astype(int), a[:, 1].astype(int), a[:, 2].astype(int), a[:, 3]
