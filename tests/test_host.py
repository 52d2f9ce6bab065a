"""Host logic (CPU): argv helpers, chunk arithmetic against the reference's golden vectors, peak grouping, trigger
output rules, GEMM config parsing."""
import json
import os

import numpy as np

GOLD = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'reference_vectors.npz'))


def test_chunk_arithmetic_golden():
    import window_filter as w
    for c in json.loads(str(GOLD['chunk_json'])):
        tlen, Lc, hist, fut, lo, hi = c['case']
        st = w.chunk_starts(tlen, Lc, hist, fut, lo, hi)
        a, b = w.valid(st, Lc, hist, fut, hi)
        assert st.tolist() == c['starts'] and a.tolist() == c['a'] and b.tolist() == c['b']
        # the valid ranges tile [max(lo, hist), min(hi, tlen - fut)) exactly once
        lo_ = hist if lo is None else max(lo, hist); hi_ = tlen - fut if hi is None else min(hi, tlen - fut)
        cov = np.zeros(tlen, int)
        for x, y in zip(a, b):
            cov[x:y] += 1
        assert np.all(cov[lo_:hi_] == 1) and cov[:lo_].sum() == 0


def test_segment_arithmetic_matches_oracle_and_golden():
    import gpu_prep
    for c, exp in zip(GOLD['seg_cases'], json.loads(str(GOLD['seg_json']))):
        n, sr, st, L, sp, ep, ts, te = c
        got = gpu_prep.segment_slices(int(n), int(sr), float(st), int(L), int(sp), int(ep), int(ts), int(te))
        assert [[list(a), list(b)] for a, b in got] == exp
    for c, exp in zip(GOLD['span_cases'], json.loads(str(GOLD['span_json']))):
        assert [list(s) for s in gpu_prep.overlapping_psd_spans(*[int(v) for v in c])] == exp
    np.testing.assert_array_equal(gpu_prep.cluster_over_window(GOLD['cl_idx'], GOLD['cl_vals'], 2048),
                                  GOLD['cl_keep'])
    np.testing.assert_array_equal([gpu_prep.median_bias(n) for n in range(1, 80)] + [gpu_prep.median_bias(1001)],
                                  GOLD['median_bias'])


def test_argv_helpers():
    import multi_filter as mf
    av = mf.setopt(['--a', '1'], '--a', '2')
    assert av == ['--a', '2'] and mf.setopt(av, '--b', '3') == ['--a', '2', '--b', '3']
    d = mf.dataset_argv(['--pad-data', '8', '--segment-start-pad', '144'], 'H1', 1000, 5000, 'x.hdf')
    assert d[d.index('--gps-start-time') + 1] == '1008' and d[d.index('--trig-start-time') + 1] == '1152'
    assert d[d.index('--trig-end-time') + 1] == str(5000 - 8 - 16)
    opt, _ = mf.pycbc_parser().parse_known_args(['--snr-threshold', '5.5', '--chisq-bins', '16', '--zz', '1'])
    assert opt.snr_threshold == 5.5 and opt.chisq_bins == '16'


def test_peak_groups_like_reference():
    """chisq_kernel.py:96-107: sort by (row, time), groups of <= P peaks of one row (the reference's blocks take 8;
    each peak's sums do not depend on the grouping, and P = 7 fits one peak's re-centring samples)."""
    import chisq_kernel
    P = chisq_kernel.P
    rows = np.array([3, 1, 1, 3, 1, 2] + [5] * 19)
    times = np.arange(len(rows))[::-1]
    o, grow, gstart, gcount = chisq_kernel.plan_groups(rows, times)
    assert np.all(np.diff(rows[o]) >= 0)
    n5 = [P] * (19 // P) + ([19 % P] if 19 % P else [])
    assert grow.tolist() == [1, 2, 3] + [5] * len(n5) and gcount.tolist() == [3, 1, 2] + n5
    for s, c in zip(gstart, gcount):
        assert len(set(rows[o][s:s + c])) == 1


def test_trigger_output_rules():
    """fp_filter.py:910-937: gate veto |t - g_t| > w + taper + 0.5, lexsort by (template_id, end_time), newsnr."""
    import fp_filter
    tr = fp_filter.Triggers()
    tr.add(tid=np.array([5, 2, 5]), t=np.array([10.0, 3.0, 1.0]), snr=np.array([6.0, 7.0, 8.0]),
           ph=np.zeros(3), chi=np.array([30.0, 0.0, 90.0]), dof=np.array([30, 30, 30]), sg=np.ones(3),
           ex=np.ones(3, bool), samp=np.array([1, 2, 3]), seg=np.zeros(3))
    out = tr.to_dict(gates=[(3.2, 0.25, 0.25)])
    assert out['template_id'].tolist() == [5, 5] and out['end_time'].tolist() == [1.0, 10.0]
    assert abs(out['newsnr'][0] - 8 / ((1 + 27) / 2) ** (1 / 6)) < 1e-12 and out['newsnr'][1] == 6.0


def test_trigger_order_equals_lexsort():
    """The single-key sort of the trigger list is the stable lexsort by (template_id, end_time), including ties,
    GPS-sized times and two end times one ulp apart (where the key cannot separate them and the lexsort runs)."""
    import fp_filter
    rng = np.random.default_rng(7)
    for trial in range(60):
        n = int(rng.integers(0, 2000))
        tid = rng.integers(0, int(rng.integers(1, 30000)), n)
        t = (1.2e9 if trial % 2 else 0.0) + np.round(rng.uniform(0, 2048, n) * 4096) / 4096
        if trial % 5 == 0 and n:
            t[: n // 2] = t[0]
        if trial % 7 == 0 and n > 3:
            t[1], tid[1] = np.nextafter(t[0], np.inf), tid[0]
        assert np.array_equal(fp_filter._order_by_template_time(tid, t), np.lexsort((t, tid))), trial


def test_gemm_cfg_parsing():
    import fp_gemm
    assert fp_gemm._cfg(None) == (0, 0) and fp_gemm._cfg(32) == (32, 0) and fp_gemm._cfg((16, 1)) == (16, 1)


def test_batched_segment_grids():
    """The batched form's segmenting: each chunk on the longest grid that leaves it a full segment beyond the 64 s of
    pads (the shortest grid from its own length up); every chunk's segments analyse L - 64 s, whatever the grid."""
    import gpu_prep, workloads
    exp = {80: None, 127: None, 128: 128, 192: 128, 319: 128, 320: 256, 448: 256, 575: 256, 576: 512, 1024: 512}
    for L, g in exp.items():
        c = workloads.chunk_conditioning(L)
        assert (c.seg_len if c else None) == g, L
    # the two-grid segmenting: 128 s segments up to 575 s
    assert [workloads.chunk_conditioning(L, (512, 128)).seg_len for L in (128, 320, 575, 576)] == [128, 128, 128, 512]
    sr = 16
    for L in range(128, 1025, 16):
        c = workloads.chunk_conditioning(L)
        seg = gpu_prep.segment_slices(L * sr, sr, 0.0, c.seg_len, c.start_pad, c.end_pad, 0, 0)
        cov = np.zeros(L * sr, int)
        for (a, _), (b0, b1) in seg:
            cov[a + b0:a + b1] += 1
        assert np.all(cov[c.start_pad * sr:(L - c.end_pad) * sr] == 1) and cov.sum() == (L - 64) * sr, L
