"""Writes golden vectors from the reference's own numpy-only functions (run only where the reference's sources are
mounted; the tests read the written vectors, never the reference). Only functions whose integer/index conventions no closed form
pins are sampled: gpu_prep.segment_slices, overlapping_psd_spans, inverse_spectrum_truncation (edge bins, Hann
halves), welch (float32 path), median_bias, interpolate, cluster_over_window, window_filter.chunk_starts/valid,
resample_ldas (the 'ldas' FIR decimation), and
gpu_prep.gpu_stage2 (numpy backend) end to end on a seeded strain: one PSD span, and several spans with per-segment
PSD picks (stage2_strain() below rebuilds the inputs from their seeds; D and IP are stored on a bin subsample).

Run: python tests/golden/make_golden.py  ->  tests/golden/reference_vectors.npz
"""
import json
import os
import sys

import numpy as np

REF = '/root/reference/pipeline'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'reference_vectors.npz')


# gpu_stage2 cases: (name, seed, sample rate, seconds before padding, pad, pycbc_inspiral options)
STAGE2 = [
    ('one_span', 5, 512, 64, 4, ['--segment-length', '64', '--segment-start-pad', '8', '--segment-end-pad', '8',
                                 '--psd-segment-length', '4', '--psd-segment-stride', '2', '--psd-num-segments', '31',
                                 '--psd-inverse-length', '4', '--psd-estimation', 'median']),
    ('multi_span', 6, 256, 1100, 8, ['--segment-length', '256', '--segment-start-pad', '16', '--segment-end-pad', '16',
                                     '--psd-segment-length', '16', '--psd-segment-stride', '8', '--psd-num-segments',
                                     '63', '--psd-inverse-length', '16', '--invpsd-trunc-method', 'hann',
                                     '--psd-estimation', 'median']),
]
STAGE2_GPS = 1000000000
STAGE2_STRIDE = 13        # D and IP are stored on every 13th bin (and in full on the 64 bins from kmin)


def stage2_strain(seed, sr, seconds, pad):
    """float32 strain with its padding: white noise times a slow envelope (so PSD spans differ and the per-segment
    pick matters), plus two lines. The tests import this function to rebuild the inputs."""
    rng = np.random.default_rng(seed)
    n = int((seconds + 2 * pad) * sr)
    t = np.arange(n) / sr
    env = 1.0 + 0.6 * np.sin(2 * np.pi * t / 700.0 + 0.3)
    x = rng.standard_normal(n) * env + 3.0 * np.sin(2 * np.pi * 60.0 * t) + 0.5 * np.sin(2 * np.pi * 33.3 * t)
    return (x * 1e3).astype(np.float32)


def stage2_argv(sr, seconds, pad, opts):
    return ['--hdf-store', 'synthetic.hdf', '--channel-name', 'H1:STRAIN', '--gps-start-time', str(STAGE2_GPS),
            '--gps-end-time', str(STAGE2_GPS + seconds), '--pad-data', str(pad), '--sample-rate', str(sr),
            '--low-frequency-cutoff', '15'] + opts


def main():
    sys.path.insert(0, REF)
    import gpu_prep as g          # numpy + argparse at import time
    import window_filter as w     # numpy at import time
    rng = np.random.default_rng(20261203)
    out = {}
    # segment arithmetic: production-like pads and odd lengths
    cases = [(2048 * 4096, 2048, 0, 1024, 144, 16, 0, 0), (2048 * 12272, 2048, 1380077576, 1024, 144, 16, 0, 0),
             (2048 * 512, 2048, 1e9, 512, 48, 16, 0, 0), (2048 * 700, 2048, 1e9, 256, 48, 16, 0, 0),
             (2048 * 3000, 2048, 1e9, 1024, 48, 16, int(1e9) + 100, int(1e9) + 2900), (4096 * 300, 4096, 5e8, 64, 8, 8, 0, 0)]
    segs = [g.segment_slices(*c) for c in cases]
    out['seg_cases'] = np.array(cases, dtype=np.float64)
    out['seg_json'] = np.array(json.dumps(segs))
    span_cases = [(2048 * 512, 32768, 16384, 63), (2048 * 4096, 32768, 16384, 63), (2048 * 12256, 32768, 16384, 63),
                  (2048 * 100, 8192, 4096, 47), (1 << 20, 1 << 15, 1 << 14, 20)]
    out['span_cases'] = np.array(span_cases, dtype=np.int64)
    out['span_json'] = np.array(json.dumps([g.overlapping_psd_spans(*c) for c in span_cases]))
    out['median_bias'] = np.array([g.median_bias(n) for n in range(1, 80)] + [g.median_bias(1001)])
    # inverse spectrum truncation, float64, with and without the Hann taper
    f = np.arange(2049) / 8.0
    P = 1e-3 * ((np.maximum(f, 10) / 20) ** -8 + 3 * (np.maximum(f, 10) / 100) ** -2 + 1 + (f / 300) ** 2)
    P *= 1 + 0.3 * rng.random(len(f))
    out['ist_psd'] = P
    out['ist_plain'] = g.inverse_spectrum_truncation(np, P.copy(), 1 / 8.0, 512, 15.0, None)
    out['ist_hann'] = g.inverse_spectrum_truncation(np, P.copy(), 1 / 8.0, 512, 15.0, 'hann')
    out['ist_noflow'] = g.inverse_spectrum_truncation(np, P.copy(), 1 / 8.0, 256, None, None)
    # welch on float32 input (the reference's precision path), median of an odd and an even count, and mean
    x = (rng.standard_normal(3 * 8192 + 100) * 2).astype(np.float32)
    out['welch_x'] = x
    out['welch_med'], _ = g.welch(np, x, 4096, 2048, 1 / 2048.0)
    out['welch_even'], _ = g.welch(np, x, 4096, 4096, 1 / 2048.0, num_segments=6)
    out['welch_mean'], _ = g.welch(np, x, 4096, 2048, 1 / 2048.0, avg='mean')
    out['interp_in'] = P[:257]
    out['interp_out'] = g.interpolate(np, P[:257], 0.25, 1 / 64.0, 4097)
    idx = np.sort(rng.choice(100000, 300, replace=False)).astype(np.int32)
    vals = rng.random(300)
    out['cl_idx'], out['cl_vals'] = idx, vals
    out['cl_keep'] = g.cluster_over_window(idx, vals, 2048)
    chunks = []
    for (tlen, Lc, hist, fut, lo, hi) in [(1 << 21, 1 << 17, 1 << 15, 1 << 12, None, None),
                                          (1 << 21, 1 << 16, 20000, 3000, 300000, 1900000),
                                          (1 << 20, 1 << 18, 1 << 16, 1 << 13, 150000, None)]:
        st = w.chunk_starts(tlen, Lc, hist, fut, lo, hi)
        a, b = w.valid(st, Lc, hist, fut, hi)
        chunks.append(dict(case=[tlen, Lc, hist, fut, lo, hi], starts=st.tolist(), a=a.tolist(), b=b.tolist()))
    out['chunk_json'] = np.array(json.dumps(chunks))
    xr = rng.standard_normal(6001)
    out['resample_x'] = xr
    out['resample_4'] = g.resample_ldas(np, xr.copy(), 4)
    out['resample_2'] = g.resample_ldas(np, xr.copy(), 2)
    meta = []
    for name, seed, sr, seconds, pad, opts in STAGE2:
        h = stage2_strain(seed, sr, seconds, pad)
        r = g.gpu_stage2(np, h.copy(), stage2_argv(sr, seconds, pad, opts))
        kmin = int(15.0 / r['df'])
        out[f's2_{name}_D'] = r['D'][:, ::STAGE2_STRIDE]
        out[f's2_{name}_IP'] = r['IP'][:, ::STAGE2_STRIDE]
        out[f's2_{name}_Dk'] = r['D'][:, kmin - 32:kmin + 32]
        out[f's2_{name}_IPk'] = r['IP'][:, kmin - 32:kmin + 32]
        meta.append(dict(name=name, flen=r['flen'], tlen=r['tlen'], df=r['df'], sr=r['sr'], gps_start=r['gps_start'],
                         analyze=[list(map(int, a)) for a in r['analyze']], seg_start=list(map(int, r['seg_start']))))
    out['s2_json'] = np.array(json.dumps(meta))
    np.savez_compressed(OUT, **out)
    print('wrote', OUT, sorted(out))


if __name__ == '__main__':
    main()
