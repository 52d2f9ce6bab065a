"""Data conditioning on the GPU (the reference's gpu_prep.py for the steps on its GPU path, :87-231, 234-282,
405-454): segment FFTs, overlapping median-Welch PSDs in float64, interpolation, inverse-spectrum truncation, then
D = stilde / psd and IP = 1 / psd, zero below the low-frequency cutoff; and the autogating (|whitened strain| above
threshold, clustered, inverted-Tukey gates).

FFTs are hipFFT (through torch.fft); every other pass is a kernel in csrc/prep.hip. Segment and PSD-span index
arithmetic is integer host code, as in the reference. The strain is float32 already multiplied by DYN_RANGE_FAC;
the PyCBC-side stages of the reference (store read, lal IIR high-pass, resampling, injections from files) are not
on this path.
"""
from __future__ import annotations

import numpy as np
import torch

import _dev
import _lib
import peak_kernel

DYN_RANGE_FAC = 5.9029581035870565e+20


def rows_fft(fn, x, **kw):
    """fn (torch.fft.rfft or irfft) along the last dimension of 2-D x, over its rows in power-of-two blocks: hipFFT
    plans are then made once per block size and reused, where a new row count (the segment or PSD-span count of a
    batch) would make and upload a new plan each time."""
    R = x.shape[0]
    if R & (R - 1) == 0:
        return fn(x, dim=1, **kw)
    first = fn(x[:1], dim=1, **kw)
    out = torch.empty((R,) + tuple(first.shape[1:]), dtype=first.dtype, device=x.device)
    r0 = 0
    for b in reversed(range(R.bit_length())):
        if R >> b & 1:
            fn(x[r0:r0 + (1 << b)], dim=1, out=out[r0:r0 + (1 << b)], **kw)
            r0 += 1 << b
    return out


def median_bias(n):
    """gpu_prep.py:47-53."""
    if n >= 1000:
        return np.log(2)
    ans = 1.0
    for i in range(1, (n - 1) // 2 + 1):
        ans += 1.0 / (2 * i + 1) - 1.0 / (2 * i)
    return ans


def welch(x, seg_len, seg_stride, dt, num_segments=None, avg='median'):
    """gpu_prep.py:87-125 in float64 on the GPU. x: (n,) or (rows, n) float32 tensor. Returns (psd float64 tensor,
    delta_f)."""
    X = x[None] if x.ndim == 1 else x
    X = X.to(torch.float32).contiguous()
    rows, n = X.shape
    if num_segments is None:
        num_segments = int(n // seg_stride)
        if (num_segments - 1) * seg_stride + seg_len > n:
            num_segments -= 1
    data_len = (num_segments - 1) * seg_stride + seg_len
    start = 0
    if data_len < n:
        diff = n - data_len
        start = diff // 2 + (diff % 2)
    assert data_len <= n, 'Incorrect choice of segmentation parameters'
    segs = torch.empty((rows, num_segments, seg_len), dtype=torch.float64, device=X.device)
    _lib.call('gwgs_welch_segments', X.data_ptr(), X.stride(0), rows, start, num_segments, seg_len, seg_stride,
              segs.data_ptr(), _dev.stream())
    st = rows_fft(torch.fft.rfft, segs.view(rows * num_segments, seg_len)).view(rows, num_segments, -1)
    nf = seg_len // 2 + 1
    out = torch.empty((rows, nf), dtype=torch.float64, device=X.device)
    w = np.hanning(seg_len)
    delta_f = 1.0 / dt / seg_len
    fac = 2 * delta_f * seg_len / (w * w).sum()
    _lib.call('gwgs_welch_reduce', st.data_ptr(), rows, num_segments, nf, float(dt), int(avg != 'median'),
              float(median_bias(num_segments)), float(fac), out.data_ptr(), _dev.stream())
    return (out[0] if x.ndim == 1 else out), delta_f


def interpolate(psd, df_old, df_new, length=None):
    """gpu_prep.py:128-135: np.interp (float64) onto k df_new."""
    P = (psd[None] if psd.ndim == 1 else psd).to(torch.float64).contiguous()
    new_n = (P.shape[-1] - 1) * df_old / df_new + 1 if length is None else length
    nn = int(np.rint(new_n))
    out = torch.empty((P.shape[0], nn), dtype=torch.float64, device=P.device)
    _lib.call('gwgs_interp', P.data_ptr(), P.shape[0], P.shape[1], float(df_old), float(df_new), nn, out.data_ptr(),
              _dev.stream())
    return out[0] if psd.ndim == 1 else out


def inverse_spectrum_truncation(psd, df, max_filter_len, low_frequency_cutoff, trunc_method=None):
    """gpu_prep.py:138-158 in float64: psd^-1/2 on [kmin, N/2), unnormalised irfft * df, zero outside +-L/2 (Hann
    taper of the kept ends with trunc_method 'hann'), rfft / (df N), 1 / |.|^2."""
    P = (psd[None] if psd.ndim == 1 else psd).to(torch.float64).contiguous()
    rows, flen = P.shape
    N = (flen - 1) * 2
    kmin = int(low_frequency_cutoff / df) if low_frequency_cutoff else 1
    inv = torch.empty((rows, flen), dtype=torch.complex128, device=P.device)
    _lib.call('gwgs_isqrt_band', P.data_ptr(), rows, flen, kmin, N // 2, inv.data_ptr(), _dev.stream())
    q = rows_fft(torch.fft.irfft, inv, n=N, norm='forward').contiguous()
    _lib.call('gwgs_trunc', q.data_ptr(), rows, N, int(max_filter_len), int(trunc_method == 'hann'), float(df),
              _dev.stream())
    pt = rows_fft(torch.fft.rfft, q)
    out = torch.empty((rows, flen), dtype=torch.float64, device=P.device)
    _lib.call('gwgs_inv_abs2', pt.data_ptr(), rows, flen, 1.0 / (df * N), out.data_ptr(), _dev.stream())
    return out[0] if psd.ndim == 1 else out


def segment_slices(n, sr, start_gps, seg_len, start_pad, end_pad, trig_start=0, trig_end=0):
    """gpu_prep.py:234-271 (pycbc.strain.StrainSegments index arithmetic): [((s0, s1), (a0, a1))]."""
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
            assert s0 >= 0 and s1 <= n, 'segment outside the data needs zero padding (not supported)'
            out.append(((s0, s1), (a0, a1)))
    return out


def overlapping_psd_spans(n, seg_len, seg_stride, num_segments):
    """gpu_prep.py:274-282 (pycbc.psd.generate_overlapping_psds)."""
    psd_len = (num_segments - 1) * seg_stride + seg_len
    assert n >= psd_len
    if n == psd_len:
        return [(0, n)]
    m = int(2 * (n - 1) / psd_len)
    stride = int((n - psd_len) / m)
    return [(n - psd_len, n) if i == m - 1 else (stride * i, psd_len + stride * i) for i in range(m)]


def pick_psds(segs, spans):
    """gpu_prep.py:438-447 (associate_psds_to_segments): largest overlap, first wins ties."""
    pick = []
    for (a, b), _ in segs:
        best, ov = None, 0
        for j, (c, e) in enumerate(spans):
            o = min(b, e) - max(a, c)
            if o > 0 and o > ov:
                best, ov = j, o
        pick.append(best)
    return pick


def cluster_over_window(idx, vals, window):
    """gpu_prep.py:161-172 (findchirp_cluster_over_window; a sequential host loop over the few threshold hits)."""
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


def whitened_magnitude(s, sr, flow, corrupt_time, psd_duration=4.0, psd_stride=2.0):
    """gpu_prep.py:175-202: |strain whitened by a 4 s median-Welch PSD (Hann-truncated), unit variance| over the n
    samples, zero at the corrupt ends. float32 tensor (n,)."""
    n = s.shape[0]
    dt = 1.0 / sr
    cl = int(corrupt_time * sr)
    L = 1 << n.bit_length()
    pad_start = int(L / 2 - n / 2)
    sp = torch.empty(L, dtype=torch.float64, device=s.device)
    _lib.call('gwgs_gate_pad', s.data_ptr(), n, cl, pad_start, L, sp.data_ptr(), _dev.stream())
    core = sp[pad_start + cl:pad_start + n - cl].to(torch.float32)
    psd, pdf = welch(core, int(psd_duration * sr), int(psd_stride * sr), dt)
    df = 1.0 / (L * dt)
    psd = interpolate(psd, pdf, df)
    psd = inverse_spectrum_truncation(psd, df, int(psd_duration * sr), flow, 'hann').contiguous()
    norm = sr / 2.0 - flow
    st = torch.fft.rfft(sp).contiguous()
    _lib.call('gwgs_gate_whiten', st.data_ptr(), psd.data_ptr(), st.shape[0], int(flow / df), dt, norm, _dev.stream())
    wt = torch.fft.irfft(st, n=L, norm='forward').contiguous()
    mag = torch.empty(n, dtype=torch.float32, device=s.device)
    _lib.call('gwgs_gate_mag', wt.data_ptr(), pad_start, n, cl, float(df), mag.data_ptr(), _dev.stream())
    return mag


def detect_loud_glitches(s, sr, start_gps, threshold, cluster_window, flow, corrupt_time):
    """gpu_prep.py:175-209 with float64 GPS times: times of the clustered |whitened| > threshold samples."""
    mag = whitened_magnitude(s, sr, flow, corrupt_time)
    ind = peak_kernel.select_above(mag, float(threshold))
    vals = mag[torch.from_numpy(ind).to(mag.device)].cpu().numpy() if len(ind) else np.zeros(0, np.float32)
    keep = cluster_over_window(ind.astype(np.int32), vals, int(cluster_window * sr))
    return [start_gps + int(i) / sr for i in ind[keep]]


def gate_taper(x, sr, start_gps, gates):
    """gpu_prep.py:212-231 (pycbc gate_data): inverted Tukey windows, in place on the float32 strain tensor."""
    n = x.shape[0]
    duration = n / sr
    for gt, gw, pw in gates:
        t_start = gt - gw - pw - start_gps
        t_end = gt + gw + pw - start_gps
        if t_start > duration or t_end < 0.0:
            continue
        win = int(2 * sr * (gw + pw))
        pad = int(sr * pw)
        offset = int(t_start * sr)
        i1, i2 = max(0, -offset), min(win, n - offset)
        _lib.call('gwgs_gate_apply', x.data_ptr(), n, offset, i1, i2, pad, win, _dev.stream())
    return x


def autogate(s, sr, start_gps, threshold, cluster=5.0, width=0.25, taper=0.25, pad=16.0, flow=15.0,
             max_iterations=1):
    """gpu_prep.py:383-392: detect, gate, repeat. Returns the gate list [(t, width, taper)] (s gated in place)."""
    gates = []
    for _ in range(max_iterations):
        gt = detect_loud_glitches(s, sr, start_gps, threshold, cluster, flow, pad)
        gp = [(g, width, taper) for g in gt]
        gates += gp
        gate_taper(s, sr, start_gps, gp)
        if not gt:
            break
    return gates


def condition(s, sr, gps_start, seg_len, start_pad, end_pad, psd_seg_len, psd_stride, psd_nseg, psd_inv_len, flow,
              trunc_method=None, avg='median', tag='', gates=(), trig_start=0, trig_end=0):
    """gpu_stage2 (gpu_prep.py:405-454) from the float32 strain tensor s (x DYN_RANGE_FAC, on the GPU): segment
    FFTs, PSD spans, float64 median Welch, interpolation, truncation, and D, IP. Returns an fp_filter.Dataset."""
    from fp_filter import Dataset
    s = s.to(torch.float32).contiguous()
    n = s.shape[0]
    seg = segment_slices(n, sr, gps_start, seg_len, start_pad, end_pad, trig_start, trig_end)
    tlen = int(seg_len * sr)
    flen = tlen // 2 + 1
    df = 1.0 / seg_len
    starts = [a for (a, _), _ in seg]
    x = torch.stack([s[a:a + tlen] for a in starts])
    X = rows_fft(torch.fft.rfft, x).contiguous()               # complex64, scaled by 1/sr in the whitening pass
    pl, ps = int(psd_seg_len * sr), int(psd_stride * sr)
    spans = overlapping_psd_spans(n, pl, ps, psd_nseg)
    parts = torch.stack([s[a:b] for a, b in spans])
    psd, pdf = welch(parts, int(psd_seg_len * sr + 0.5), int(psd_stride * sr + 0.5), 1.0 / sr, num_segments=psd_nseg,
                     avg=avg)
    if pdf != df:
        psd = interpolate(psd, pdf, df, flen)
    if psd_inv_len:
        psd = inverse_spectrum_truncation(psd, df, int(psd_inv_len * sr), flow, trunc_method)
    psd = psd.contiguous()
    pick = _dev.i32(pick_psds(seg, spans))
    kmin = int(flow / df)
    D = torch.empty((len(seg), flen), dtype=torch.complex64, device=s.device)
    IP = torch.empty((len(seg), flen), dtype=torch.float32, device=s.device)
    _lib.call('gwgs_whiten', X.data_ptr(), float(np.float32(1.0 / sr)), psd.data_ptr(), pick.data_ptr(), len(seg),
              flen, kmin, D.data_ptr(), IP.data_ptr(), _dev.stream())
    return Dataset(tag, D, IP, flen, tlen, df, float(sr), float(gps_start), [a for _, a in seg], starts, list(gates))


def condition_many(xs, sr, gps_starts, seg_len, start_pad, end_pad, psd_seg_len, psd_stride, psd_nsegs, psd_inv_len,
                   flow, trunc_method=None, avg='median', tags=None, gates=None):
    """condition() for several strains on one segment grid in a handful of batched launches: every segment of every
    strain in one segment transform, the PSD spans of all strains with the same Welch segment count in one Welch
    estimate, interpolation and truncation, and one whitening pass over all segments with each segment's PSD row
    picked by largest overlap within its own strain (gpu_prep.py:405-454 per dataset). xs: float32 strain tensors
    (x DYN_RANGE_FAC, on the GPU); psd_nsegs: the Welch segment count of each strain. No host synchronisation: the
    index arithmetic is host code and the pick table goes over in a pinned copy. Returns one fp_filter.Dataset per
    strain; their D and IP are row blocks of one tensor each."""
    from fp_filter import Dataset
    tlen = int(seg_len * sr)
    flen = tlen // 2 + 1
    df = 1.0 / seg_len
    kmin = int(flow / df)
    pl, ps = int(psd_seg_len * sr), int(psd_stride * sr)
    per = []
    for x, gps, nseg in zip(xs, gps_starts, psd_nsegs):
        n = x.shape[0]
        seg = segment_slices(n, sr, gps, seg_len, start_pad, end_pad)
        per.append((seg, overlapping_psd_spans(n, pl, ps, nseg)))
    xs = [x.to(torch.float32).contiguous() for x in xs]
    X = rows_fft(torch.fft.rfft, torch.stack([x[a:a + tlen] for x, (seg, _) in zip(xs, per)
                                              for (a, _), _ in seg])).contiguous()
    groups = {}
    for i, nseg in enumerate(psd_nsegs):
        groups.setdefault(int(nseg), []).append(i)
    row0 = [0] * len(xs)
    psds, row = [], 0
    for nseg, idx in groups.items():
        parts = torch.stack([xs[i][a:b] for i in idx for a, b in per[i][1]])
        psd, pdf = welch(parts, int(psd_seg_len * sr + 0.5), int(psd_stride * sr + 0.5), 1.0 / sr,
                         num_segments=nseg, avg=avg)
        if pdf != df:
            psd = interpolate(psd, pdf, df, flen)
        if psd_inv_len:
            psd = inverse_spectrum_truncation(psd, df, int(psd_inv_len * sr), flow, trunc_method)
        for i in idx:
            row0[i] = row
            row += len(per[i][1])
        psds.append(psd)
    P = torch.cat(psds).contiguous()
    pick = np.concatenate([np.asarray(pick_psds(seg, spans), np.int64) + row0[i]
                           for i, (seg, spans) in enumerate(per)]).astype(np.int32)
    pick_d = _dev.to_dev(pick)
    S = X.shape[0]
    D = torch.empty((S, flen), dtype=torch.complex64, device=X.device)
    IP = torch.empty((S, flen), dtype=torch.float32, device=X.device)
    _lib.call('gwgs_whiten', X.data_ptr(), float(np.float32(1.0 / sr)), P.data_ptr(), pick_d.data_ptr(), S, flen, kmin,
              D.data_ptr(), IP.data_ptr(), _dev.stream())
    out, s0 = [], 0
    for i, (seg, _) in enumerate(per):
        s1 = s0 + len(seg)
        out.append(Dataset(tags[i] if tags else '', D[s0:s1], IP[s0:s1], flen, tlen, df, float(sr),
                           float(gps_starts[i]), [a for _, a in seg], [a for (a, _), _ in seg],
                           list(gates[i]) if gates else []))
        s0 = s1
    return out


# ---------------------------------------------------------------------------------------------------------------------
# The reference's stage API (gpu_prep.py:214-470): argv-driven conditioning of one dataset. The CPU stages read the
# store and apply the IIR high-pass; the GPU stages resample, add injections, gate, and condition. xp (cupy in the
# reference) is accepted for the signature and ignored: arrays are torch tensors on the GPU.
# ---------------------------------------------------------------------------------------------------------------------

UNSUPPORTED = ('--gating-file', '--zpk-z', '--normalize-strain', '--strain-low-pass', '--taper-data',
               '--witness-frame-type', '--sgburst-injection-file', '--psd-model', '--psd-file', '--asd-file',
               '--fake-strain', '--filter-inj-only', '--allow-zero-padding', '--frame-files', '--frame-cache',
               '--frame-type')


def parser():
    """The pycbc_inspiral options the conditioning reads (gpu_prep.py:24-44); the options it does not implement are
    accepted by the parser and refused by options()."""
    import argparse
    p = argparse.ArgumentParser(add_help=False)
    a = p.add_argument
    a('--hdf-store'); a('--channel-name'); a('--gps-start-time', type=int); a('--gps-end-time', type=int)
    a('--pad-data', type=int, default=0); a('--strain-high-pass', type=float); a('--sample-rate', type=float)
    a('--low-frequency-cutoff', type=float); a('--segment-length', type=int)
    a('--segment-start-pad', type=int); a('--segment-end-pad', type=int)
    a('--trig-start-time', type=int, default=0); a('--trig-end-time', type=int, default=0)
    a('--psd-estimation'); a('--psd-segment-length', type=float); a('--psd-segment-stride', type=float)
    a('--psd-num-segments', type=int); a('--psd-inverse-length', type=float); a('--invpsd-trunc-method')
    a('--autogating-threshold', type=float); a('--autogating-cluster', type=float, default=5.0)
    a('--autogating-width', type=float, default=0.25); a('--autogating-taper', type=float, default=0.25)
    a('--autogating-pad', type=float, default=16); a('--autogating-max-iterations', type=int, default=1)
    a('--gating-method', default='taper')
    a('--injection-file'); a('--injection-scale-factor', type=float, default=1); a('--injection-sample-rate', type=float)
    for k in UNSUPPORTED:
        a(k, dest='unsupported_' + k[2:].replace('-', '_'))
    return p


def options(argv):
    """gpu_prep.py:306-311: parsed options; AssertionError on an option this path does not implement."""
    opt, _ = parser().parse_known_args(argv)
    bad = [k for k, v in vars(opt).items() if k.startswith('unsupported_') and v]
    assert not bad, f'gpu_prep does not implement {bad}'
    assert opt.psd_estimation in ('median', 'mean') and opt.gating_method == 'taper' and opt.hdf_store
    return opt


class _Laps:
    """Stage timings as the reference's timing dicts (gpu_prep.py:296-303): seconds since the previous lap, with the
    device synchronised first."""

    def __init__(self, gpu):
        import time
        self.gpu, self.T, self.t, self._time = gpu, {}, time.perf_counter(), time.perf_counter

    def __call__(self, k):
        if self.gpu:
            torch.cuda.synchronize()
        self.T[k] = round(self._time() - self.t, 4)
        self.t = self._time()


def write_store(path, channel, segments):
    """A strain store in the layout the stage API reads (the reference's HDF5 store, gpu_prep.py:321-330, as a numpy
    .npz archive: h5py is not available here): per channel the arrays '{channel}/segments/start' and '.../end' (GPS
    seconds) and '{channel}/{i}' (float64 strain of segment i). segments: [(gps_start, gps_end, strain)]. Written to
    a temporary name and renamed."""
    arrs = {f'{channel}/segments/start': np.array([a for a, _, _ in segments], np.float64),
            f'{channel}/segments/end': np.array([b for _, b, _ in segments], np.float64)}
    for i, (_, _, x) in enumerate(segments):
        arrs[f'{channel}/{i}'] = np.asarray(x, np.float64)
    tmp = path + '.part'
    with open(tmp, 'wb') as f:
        np.savez(f, **arrs)
    import os
    os.replace(tmp, path)


def highpass_cpu(x, sr, epoch, freq, filter_order=8, attenuation=0.1):
    """The IIR high-pass of pycbc.filter.highpass (gpu_prep.py:56-64 calls lal's HighPassREAL8/REAL4TimeSeries with
    1 - attenuation at freq and order filter_order). lal is not available here: this is a Butterworth high-pass of
    that order applied forward and backward (zero phase), whose two-pass amplitude at freq is 1 - attenuation, in the
    input's precision. Parity with lal's design is unpinned."""
    import scipy.signal
    x = np.asarray(x)
    # two passes: |H(f)|^2 = 1 / (1 + (fc / f)^(2 order)) at f = freq equals 1 - attenuation
    a2 = 1.0 - attenuation
    fc = freq * (1.0 / a2 - 1.0) ** (1.0 / (2 * filter_order))
    sos = scipy.signal.butter(filter_order, fc, btype='highpass', fs=sr, output='sos')
    return scipy.signal.sosfiltfilt(sos, x.astype(np.float64)).astype(x.dtype)


def cpu_stage1(argv):
    """gpu_prep.py:314-337: read the dataset's span (with padding) from the store and apply the first (float64)
    high-pass. Returns (x float64 host array at the store's rate, timing)."""
    L = _Laps(False)
    opt = options(argv)
    pad = opt.pad_data
    t0 = opt.gps_start_time - pad
    with np.load(opt.hdf_store) as f:
        ch = opt.channel_name
        starts, ends = f[f'{ch}/segments/start'], f[f'{ch}/segments/end']
        diff = t0 - starts
        loc = np.where(diff >= 0)[0]
        assert len(loc), 'no store segment starts at or before the dataset'
        sidx = loc[diff[loc].argmin()]
        assert ends[sidx] >= opt.gps_end_time + pad, 'the store segment ends before the dataset'
        d = f[f'{ch}/{sidx}']
        sr0 = len(d) / (ends[sidx] - starts[sidx])
        x = d[int((t0 - starts[sidx]) * sr0):int((opt.gps_end_time + pad - starts[sidx]) * sr0)]
    L('read')
    if opt.strain_high_pass:
        x = highpass_cpu(x, sr0, t0, opt.strain_high_pass)
    L('highpass1_cpu')
    return x, L.T


def n_resampled(argv):
    o = options(argv)
    return int(round((o.gps_end_time - o.gps_start_time + 2 * o.pad_data) * o.sample_rate))


def injection_runs(argv):
    """gpu_prep.py:345-356. Injections from a file need PyCBC's InjectionSet, which is not available here: raises
    NotImplementedError when --injection-file is given (pass a dense injection signal to gpu_stage1 instead)."""
    if options(argv).injection_file:
        raise NotImplementedError('--injection-file needs PyCBC (absent); pass inj= to gpu_stage1 / prep')
    return [], 0.0


def injection_dense(runs, n):
    x = np.zeros(n)
    for a, v in runs:
        x[a:a + len(v)] = v
    return x


def resample_ldas(x, factor):
    """gpu_prep.py:67-84 on the GPU (gwgs_fir_decimate): x float64 tensor; the FIR taps are scipy's firwin, as the
    reference's."""
    import scipy.signal
    ntaps = factor * 20 + 1
    c = torch.from_numpy(scipy.signal.firwin(ntaps, 1.0 / factor, window=('kaiser', 5))).to(x.device)
    x = x.to(torch.float64).contiguous()
    out = torch.empty((x.shape[0] + factor - 1) // factor, dtype=torch.float64, device=x.device)
    _lib.call('gwgs_fir_decimate', x.data_ptr(), x.shape[0], c.data_ptr(), ntaps, factor, out.data_ptr(),
              _dev.stream())
    return out


def gpu_stage1(xp, x, argv, inj=None):
    """gpu_prep.py:359-388: resample to --sample-rate, add the injection signal (dense float64 at the new rate; from
    --injection-file only through PyCBC, see injection_runs), x DYN_RANGE_FAC to float32, autogating. Returns (gated
    float32 strain on the GPU, gates [(time, width, taper)], timing)."""
    L = _Laps(True)
    opt = options(argv)
    t0 = opt.gps_start_time - opt.pad_data
    sr = opt.sample_rate
    sr0 = len(x) / (opt.gps_end_time - opt.gps_start_time + 2 * opt.pad_data)
    factor = int(round(sr0 / sr))
    xg = torch.as_tensor(np.asarray(x, np.float64)).cuda() if not isinstance(x, torch.Tensor) else x.to(torch.float64)
    if factor != 1:
        xg = resample_ldas(xg, factor)
    L('resample')
    if inj is None and opt.injection_file:
        injection_runs(argv)                      # raises: PyCBC's injection set is not available
    if inj is not None:
        xg = xg + torch.as_tensor(np.asarray(inj, np.float64)).to(xg.device)
        L('injections')
    s = (xg * DYN_RANGE_FAC).to(torch.float32)
    del xg
    gates = []
    if opt.autogating_threshold is not None:
        for _ in range(opt.autogating_max_iterations):
            gt = detect_loud_glitches(s, sr, t0, opt.autogating_threshold, opt.autogating_cluster,
                                      opt.strain_high_pass or 0.0, opt.pad_data + opt.autogating_pad)
            gp = [(g_, opt.autogating_width, opt.autogating_taper) for g_ in gt]
            gates += gp
            gate_taper(s, sr, t0, gp)
            if not gt:
                break
    L('autogating')
    return s, [(float(g), w, tp) for g, w, tp in gates], L.T


def cpu_stage2(h, argv):
    """gpu_prep.py:391-397: the second (float32) high-pass, after gating, on a host array."""
    L = _Laps(False)
    opt = options(argv)
    if opt.strain_high_pass:
        h = highpass_cpu(h, opt.sample_rate, opt.gps_start_time - opt.pad_data, opt.strain_high_pass)
    L('highpass2_cpu')
    return h, L.T


def gpu_stage2(xp, h, argv):
    """gpu_prep.py:405-454: remove the padding, segment FFTs, overlapping PSDs (float64 median Welch here),
    interpolation, truncation, D and IP. h: float32 strain (host array or GPU tensor) with its padding. Returns the
    reference's dict: D, IP (GPU tensors), flen, tlen, df, sr, gps_start, analyze, seg_start, timing."""
    L = _Laps(True)
    opt = options(argv)
    sr, pad = opt.sample_rate, opt.pad_data
    s = torch.as_tensor(np.asarray(h, np.float32)).cuda() if not isinstance(h, torch.Tensor) else h
    if pad:
        s = s[int(pad * sr):int(s.shape[0] - sr * pad)]
    d = condition(s, sr, float(opt.gps_start_time), opt.segment_length, opt.segment_start_pad, opt.segment_end_pad,
                  opt.psd_segment_length, opt.psd_segment_stride, opt.psd_num_segments, opt.psd_inverse_length,
                  opt.low_frequency_cutoff, opt.invpsd_trunc_method, opt.psd_estimation, trig_start=opt.trig_start_time,
                  trig_end=opt.trig_end_time)
    L('output')
    return dict(D=d.D, IP=d.IP, flen=d.flen, tlen=d.tlen, df=d.df, sr=float(sr), gps_start=float(opt.gps_start_time),
                analyze=d.analyze, seg_start=d.seg_start, timing=L.T)


def prep(argv, xp=None, inj=None):
    """gpu_prep.py:457-470: condition one dataset, all stages in this process. Returns gpu_stage2's dict plus
    'gates' and the merged timing."""
    x, T = cpu_stage1(argv)
    s, gates, T1 = gpu_stage1(xp, x, argv, inj)
    import time
    t = time.perf_counter()
    h = s.cpu().numpy()
    del s
    h, T2 = cpu_stage2(h, argv)
    T2['highpass2_cpu'] = round(time.perf_counter() - t, 4)        # includes the device-to-host copy
    out = gpu_stage2(xp, h, argv)
    out['gates'] = gates
    out['timing'] = {**T, **T1, **T2, **out['timing']}
    return out
