"""The device pre-consolidation (gwgs_pre_keep, fp_filter.py:665-695) through the C ABI against the oracle's
pre_keep, bit-exact on the kept set, and the filter with consolidation (fp_filter.py:629-663 after every tile) with
and without it against the oracle's consolidation of the unconsolidated triggers."""
import numpy as np
import pytest
import torch

from oracle import consolidate as oc

pytestmark = pytest.mark.gpu

SR = 2048.0
CW = 2048


def _batch(rng, n, segs, frac, lo=3.84, hi=7.0):
    """vw (n, nwsum): a fraction `frac` of entries hold rounded SNRs in [lo, hi) squared (ties), the rest noise
    below any cut; iw: each column's sample inside its segment window."""
    nwsum = sum(s['nw'] for s in segs)
    apx = np.round(rng.uniform(lo, hi, (n, nwsum)), 2).astype(np.float32)
    vw = np.where(rng.random((n, nwsum)) < frac, apx * apx, rng.uniform(0, 9, (n, nwsum))).astype(np.float32)
    iw = np.empty((n, nwsum), np.int32)
    c0 = 0
    for s in segs:
        s['col0'] = c0
        iw[:, c0:c0 + s['nw']] = np.arange(s['nw']) * CW + rng.integers(0, CW, (n, s['nw']))
        c0 += s['nw']
    return vw, iw


def _expected(vw, iw, segs, cut, thr, exact_snr, margin, kl):
    keep = vw > cut
    for s in segs:
        sub = vw[:, s['col0']:s['col0'] + s['nw']]
        m, j = np.nonzero(sub > cut)                       # row-major: the part's candidate order
        col = s['col0'] + j
        mask = oc.pre_keep(s['a0'] + iw[m, col].astype(np.int64), np.sqrt(vw[m, col]), s['gps'], s['seg_start'], SR,
                           s['gates'], thr, exact_snr, margin, kl.newsnr_threshold, kl.interval, kl.num,
                           kl.exempt_snr, s['floor'])
        keep[m[~mask], col[~mask]] = False
    return keep


def _run(vw, iw, segs, cut, thr, exact_snr, margin, kl):
    import consolidate
    pk = consolidate.PreKeep(kl, thr, exact_snr, margin, SR, CW)
    vw_d, iw_d = torch.from_numpy(vw).cuda(), torch.from_numpy(iw).cuda()
    fl = {}
    tabs = []
    for s in segs:
        f = s['floor']
        if f is not None and id(f) not in fl:
            fl[id(f)] = torch.from_numpy(np.asarray(f, np.float64)).cuda()
        tabs.append((s['seg_start'] + s['a0'], s['gps'], s['col0'], s['nw'], s['gates'],
                     None if f is None else fl[id(f)]))
    pk.launch(vw_d, iw_d, cut, tabs)
    torch.cuda.synchronize()
    return vw_d.cpu().numpy()


CASES = [  # (seed, rows, interval, num, exempt, floors, gates)
    (0, 64, 1.072, 3, 5.0, False, False),
    (1, 64, 1.072, 1, None, True, True),
    (2, 200, 1.072, 5, 5.0, True, True),
    (3, 64, None, 3, 5.0, False, True),
    (4, 300, 0.5, 2, None, True, False),
    (5, 128, 3.3, 30, 5.0, True, True),
]


@pytest.mark.parametrize('case', CASES)
def test_pre_keep_matches_oracle(gpu, case):
    import consolidate, fp_filter
    seed, n, interval, num, exempt, floors, gates = case
    print('pre_keep case', case, flush=True)
    rng = np.random.default_rng(seed)
    gps = [1380077576.0, 1380090000.0]
    gts = [[(g + rng.uniform(0, 40), 0.25, 0.25) for _ in range(3)] if gates else [] for g in gps]
    fls = [None, None]
    if floors:
        for d in range(2):
            f = np.round(rng.uniform(4.0, 6.5, int(rng.integers(3, 30))), 2)
            f[rng.random(len(f)) < 0.3] = -np.inf
            fls[d] = f
    segs = [dict(gps=gps[0], seg_start=0, a0=int(rng.integers(0, 5000)), nw=int(rng.integers(3, 20)), gates=gts[0],
                 floor=fls[0]),
            dict(gps=gps[0], seg_start=1636352, a0=int(rng.integers(0, 5000)), nw=int(rng.integers(3, 20)),
                 gates=gts[0], floor=fls[0]),
            dict(gps=gps[1], seg_start=524288, a0=0, nw=int(rng.integers(3, 20)), gates=gts[1], floor=fls[1])]
    vw, iw = _batch(rng, n, segs, 0.4)
    thr, exact_snr, margin = 4.0, 5.5, 0.04
    kl = consolidate.KeepLoudest(4.0, interval, num, exempt)
    cut = fp_filter.candidate_cut(thr, margin)
    out = _run(vw, iw, segs, cut, thr, exact_snr, margin, kl)
    exp = _expected(vw, iw, segs, cut, thr, exact_snr, margin, kl)
    got = out > cut
    assert got.sum() > 0 and (vw > cut).sum() > got.sum()
    np.testing.assert_array_equal(got, exp)
    np.testing.assert_array_equal(out[got], vw[got])        # kept entries untouched


def test_pre_keep_crowded_window(gpu):
    """Keep-loudest windows with more ranked candidates than the rank kernel stages in LDS (4096): 16000 rows over
    two output windows."""
    import consolidate, fp_filter
    rng = np.random.default_rng(11)
    gps = 1380077576.0
    segs = [dict(gps=gps, seg_start=0, a0=0, nw=2, gates=[], floor=None)]
    vw, iw = _batch(rng, 16000, segs, 0.95)
    kl = consolidate.KeepLoudest(4.0, 1.072, 30, None)
    cut = fp_filter.candidate_cut(4.0, 0.04)
    out = _run(vw, iw, segs, cut, 4.0, 5.5, 0.04, kl)
    exp = _expected(vw, iw, segs, cut, 4.0, 5.5, 0.04, kl)
    np.testing.assert_array_equal(out > cut, exp)
    rb = (vw > cut) & (np.sqrt(vw) > 4.0) & (np.sqrt(vw) < np.float32(5.28))
    wt = oc.window_of(gps + iw[rb] / SR, gps, SR, 1.072)
    assert np.bincount(wt).min() > 4096 and ((out > cut) & rb).sum() == 30 * len(np.bincount(wt))


def test_pre_keep_staged_window(gpu):
    """A keep-loudest window of 1,025 to 4,096 ranked candidates (above the rank kernel's LDS counting path of 1,024
    keys, within its staged max-reduction path): 6,000 rows over one output window."""
    import consolidate, fp_filter
    rng = np.random.default_rng(12)
    gps = 1380077576.0
    segs = [dict(gps=gps, seg_start=0, a0=0, nw=1, gates=[], floor=None)]
    vw, iw = _batch(rng, 6000, segs, 0.95)
    kl = consolidate.KeepLoudest(4.0, 1.072, 30, None)
    cut = fp_filter.candidate_cut(4.0, 0.04)
    out = _run(vw, iw, segs, cut, 4.0, 5.5, 0.04, kl)
    exp = _expected(vw, iw, segs, cut, 4.0, 5.5, 0.04, kl)
    np.testing.assert_array_equal(out > cut, exp)
    rb = (vw > cut) & (np.sqrt(vw) > 4.0) & (np.sqrt(vw) < np.float32(5.28))
    cnt = np.bincount(oc.window_of(gps + iw[rb] / SR, gps, SR, 1.072))
    assert 1024 < cnt.max() <= 4096 and ((out > cut) & rb).sum() == 30 * (cnt > 0).sum()


@pytest.fixture(scope='module')
def loud_data(gpu):
    """One detector, 64 s of seeded coloured noise with two injections, 64 templates: at SNR threshold 4 most
    windows of most templates are candidates."""
    import gpu_prep, synth
    bank = synth.circular_bank(64, seed=11, mmin=5, mmax=100)
    inj = [synth.Injection(30.0, float(bank.m1[5]), float(bank.m2[5]), 12.0, 0.3),
           synth.Injection(47.0, float(bank.m1[40]), float(bank.m2[40]), 25.0, 1.1)]
    st = synth.make_stream('H1', 64, SR, seed=1, injections=inj)
    ds = gpu_prep.condition(torch.from_numpy(st.x).cuda(), SR, st.gps_start, 64, 8, 8, 4, 2, 31, 4, 15.0, tag='H1')
    return bank, ds


@pytest.mark.parametrize('precision', ['fp32', 'bf16'])
def test_filter_consolidation(loud_data, precision):
    """Consolidation after each of 4 tiles, with and without the device pre-consolidation, leaves the same trigger
    file fields, and the same (template, end time) list as the oracle's consolidation of the unconsolidated
    triggers in one go."""
    import consolidate, fp_filter
    bank, ds = loud_data
    kl = consolidate.KeepLoudest(4.0, 1.072, 3, 5.0)
    IPall, sw, isw = fp_filter.svd_metric([ds])
    outs = {}
    locs = ['mtotal>12:20-15,20-30,20-45']
    for name, kw in (('none', {}), ('cons', dict(consolidate=kl)), ('pre', dict(consolidate=kl, pre_consolidate=True)),
                     ('pre_sg', dict(consolidate=kl, pre_consolidate=True, sgchisq_snr_threshold=4.5))):
        opts = fp_filter.Options(tile=16, precision=precision, snr_threshold=4.0, exact_snr=5.5, **kw)
        plan = fp_filter.BankPlan(bank.m1, bank.m2, None, ds.flen, ds.df, opts, sw, isw,
                                  sg_locations=locs if name == 'pre_sg' else None)
        f = fp_filter.Filter([ds], opts)
        outs[name] = (f.run(plan)[0], f.T)
    a, b = outs['cons'][0].to_dict(gates=ds.gates), outs['pre'][0].to_dict(gates=ds.gates)
    assert len(a['template_id']) > 0
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    # with the sine-Gaussian values arriving one batch late, tiles consolidate only the parts already complete:
    # the same peaks, now with sg_chisq for some
    c_ = outs['pre_sg'][0].to_dict(gates=ds.gates)
    for k in a:
        np.testing.assert_array_equal(a[k], c_[k], err_msg=k)
    assert c_['sg_chisq_computed'].sum() > 0 and outs['pre_sg'][1]['n_sg_chisq_points'] >= c_['sg_chisq_computed'].sum()
    assert outs['pre'][1]['n_candidates'] < outs['cons'][1]['n_candidates']
    # the unconsolidated peaks in the order they were found (ties in a window go to the earlier peak)
    tr = outs['none'][0]
    c = dict(snr=tr.cat('snr', np.float32), chi=tr.cat('chi', np.float32), dof=tr.cat('dof', np.int32),
             t=tr.cat('t', np.float64))
    keep, _ = oc.consolidate(c, ds.gates, ds.gps_start, ds.sr, kl.newsnr_threshold, kl.interval, kl.num,
                             kl.exempt_snr)
    assert keep.sum() < len(keep)
    exp = sorted(zip(tr.cat('tid', np.int64)[keep].tolist(), c['t'][keep].tolist()))
    assert sorted(zip(a['template_id'].tolist(), a['end_time'].tolist())) == exp


@pytest.mark.parametrize('kind', ['all_exact', 'none_above_cut', 'all_tied', 'all_exempt', 'all_gated'])
def test_pre_keep_edge_cases(gpu, kind):
    """Edge cases against the oracle: every candidate in the exact band (all kept), no candidate (vw untouched),
    one SNR for all (the num earliest in row-major order per window survive), every rebuilt candidate exempt, every
    candidate inside a gate."""
    import consolidate, fp_filter
    rng = np.random.default_rng(21)
    gps = 1380077576.0
    segs = [dict(gps=gps, seg_start=0, a0=0, nw=6, gates=[], floor=None),
            dict(gps=gps, seg_start=2097152, a0=4096, nw=5, gates=[], floor=None)]
    n = 40
    vw, iw = _batch(rng, n, segs, 0.7)
    cut = fp_filter.candidate_cut(4.0, 0.04)
    if kind == 'all_exact':
        vw[vw > cut] = np.float32(6.0) ** 2
    elif kind == 'none_above_cut':
        vw = np.minimum(vw, np.float32(3.0))
    elif kind == 'all_tied':
        vw[vw > cut] = np.float32(4.5) ** 2
    elif kind == 'all_exempt':
        vw[vw > cut] = np.where(vw[vw > cut] < np.float32(5.0) ** 2, np.float32(5.1) ** 2, vw[vw > cut])
    elif kind == 'all_gated':
        for s in segs:
            s['gates'] = [(gps + (s['seg_start'] + s['a0'] + 6 * 2048) / 2048.0, 8.0, 0.25)]
    kl = consolidate.KeepLoudest(4.0, 1.072, 2, 5.0)
    out = _run(vw, iw, segs, cut, 4.0, 5.5, 0.04, kl)
    exp = _expected(vw, iw, segs, cut, 4.0, 5.5, 0.04, kl)
    np.testing.assert_array_equal(out > cut, exp)
    if kind in ('all_exact', 'none_above_cut', 'all_exempt'):
        np.testing.assert_array_equal(out, vw)
    if kind == 'all_gated':
        assert not (out > cut)[np.sqrt(vw) < np.float32(5.28)].any()
