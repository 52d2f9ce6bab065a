"""Writes golden vectors from the reference's own numpy-only functions (run only where the reference's sources are
mounted; the tests read the written vectors, never the reference). Only functions whose integer/index conventions no closed form
pins are sampled: gpu_prep.segment_slices, overlapping_psd_spans, inverse_spectrum_truncation (edge bins, Hann
halves), welch (float32 path), median_bias, interpolate, cluster_over_window, and window_filter.chunk_starts/valid.

Run: python tests/golden/make_golden.py  ->  tests/golden/reference_vectors.npz
"""
import json
import os
import sys

import numpy as np

REF = '/root/reference/pipeline'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'reference_vectors.npz')


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
    np.savez_compressed(OUT, **out)
    print('wrote', OUT, sorted(out))


if __name__ == '__main__':
    main()
