"""Chunked filtering on the GPU (fp_filter.py:324-357 --chunk / --chunk2, window_filter.py): the device chunks against
the oracle's, and the filter run on chunk datasets (full rate and band-limited) against the float64 direct filter of
the same chunks and against the whole-segment search."""
import numpy as np
import pytest
import torch

from oracle import core, pipeline, window as ow

pytestmark = pytest.mark.gpu

SR = 2048.0


@pytest.fixture(scope='module')
def heavy(gpu):
    """64 s, one detector, one SNR 20 injection; 32 templates of total mass 40-100 Msun (at most 4 s long)."""
    import gpu_prep, synth
    bank = synth.circular_bank(32, seed=3, mmin=20, mmax=50)
    st = synth.make_stream('H1', 64, SR, seed=1,
                           injections=[synth.Injection(30.0, float(bank.m1[4]), float(bank.m2[4]), 20.0, 0.3)])
    ds = gpu_prep.condition(torch.from_numpy(st.x).cuda(), SR, st.gps_start, 64, 8, 8, 4, 2, 31, 4, 15.0, tag='H1')
    return bank, ds


@pytest.mark.parametrize('spec', ['16,8,2', '16,8,2,256'])
def test_chunks_match_oracle(heavy, spec):
    import window_filter
    _, ds = heavy
    (c,), _ = window_filter.to_chunks([ds], spec)
    o = ow.to_chunks(ds.D.cpu().numpy(), ds.IP.cpu().numpy(), ds.analyze, ds.seg_start, ds.tlen, SR, spec)
    assert (c.tlen, c.flen, c.df, c.sr) == (o['tlen'], o['flen'], o['df'], o['sr'])
    assert c.analyze == o['analyze'] and c.seg_start == o['seg_start']
    Dc = c.D.cpu().numpy()
    assert np.sqrt(np.mean(np.abs(Dc - o['D']) ** 2) / np.mean(np.abs(o['D']) ** 2)) < 1e-5
    assert np.array_equal(c.IP.cpu().numpy(), o['IP'].astype(np.float32))


@pytest.mark.parametrize('spec', ['16,8,2', '16,8,2,256'])
def test_filter_on_chunks(heavy, spec):
    """Every exact trigger of the chunked search has the float64 direct-filter SNR of its chunk at its sample within
    1e-4, its end time is its chunk's absolute time, and the injection is found at the whole-segment search's SNR
    within 2% (the chunks' 1/psd is the segment's sampled every tlen / Lc bins)."""
    import fp_filter, window_filter
    bank, ds = heavy
    print('chunked filter', spec, flush=True)
    (c,), _ = window_filter.to_chunks([ds], spec)
    res = {}
    for name, d in (('whole', ds), ('chunks', c)):
        opts = fp_filter.Options(tile=32, precision='bf16', snr_threshold=5.0, cluster_window=1.0)
        IPall, sw, isw = fp_filter.svd_metric([d])
        plan = fp_filter.BankPlan(bank.m1, bank.m2, None, d.flen, d.df, opts, sw, isw)
        res[name] = fp_filter.Filter([d], opts).run(plan)[0].to_dict()
    tr = res['chunks']
    assert len(tr['template_id']) > 0
    H, _ = pipeline.spa_bank(bank.m1, bank.m2, None, c.flen, c.df, 15.0)
    D, IP = c.D.cpu().numpy(), c.IP.cpu().numpy()
    cache = {}
    worst = 0.0
    for tid, seg, samp, snr, t in zip(tr['template_id'], tr['segment'], tr['sample'], tr['snr'], tr['end_time']):
        if (tid, seg) not in cache:
            sig = 4 * c.df * float((np.abs(H[tid]) ** 2 * IP[seg]).sum())
            cache[(tid, seg)] = np.abs(core.direct_z(H[tid:tid + 1], D[seg], c.tlen)[0]) * 4 * c.df / np.sqrt(sig)
        ref = cache[(tid, seg)][samp]
        worst = max(worst, abs(snr - ref) / ref)
        assert t == ds.gps_start + (c.seg_start[seg] + samp) / c.sr
    assert worst < 1e-4, worst
    i, j = np.argmax(tr['snr']), np.argmax(res['whole']['snr'])
    assert tr['snr'][i] > 15 and abs(tr['snr'][i] / res['whole']['snr'][j] - 1) < 0.02
    assert abs(tr['end_time'][i] - res['whole']['end_time'][j]) < 0.01


def test_gated_chunks(heavy):
    """--gate (fp_filter.py:424-445, 724-756): the chunked search gated by the whole-segment search's loud triggers
    (SNR > 8) keeps exactly the ungated chunked triggers whose tile has a gate trigger of matching chirp mass near the
    chunk and which lie within the gate window of one, and skips the other (tile, chunk) pairs."""
    import fp_filter, window_filter
    bank, ds = heavy
    (c,), _ = window_filter.to_chunks([ds], '16,8,2')
    opts = fp_filter.Options(tile=8, precision='bf16', snr_threshold=5.0)
    IPall, sw, isw = fp_filter.svd_metric([ds])
    base = fp_filter.Filter([ds], opts).run(fp_filter.BankPlan(bank.m1, bank.m2, None, ds.flen, ds.df, opts, sw,
                                                               isw))[0].to_dict()
    loud = base['snr'] > 8
    mc_all = (bank.m1 * bank.m2) ** 0.6 / (bank.m1 + bank.m2) ** 0.2
    gate = fp_filter.Gate(np.array(['H1'] * int(loud.sum())), base['end_time'][loud], mc_all[base['template_id'][loud]],
                          mc_ratio=1.25, window=1.0)
    assert 0 < loud.sum()
    IPc, swc, iswc = fp_filter.svd_metric([c])
    plan = fp_filter.BankPlan(bank.m1, bank.m2, None, c.flen, c.df, opts, swc, iswc)
    full = fp_filter.Filter([c], opts).run(plan)[0].to_dict()
    fg = fp_filter.Filter([c], fp_filter.Options(tile=8, precision='bf16', snr_threshold=5.0, gate=gate))
    got = fg.run(plan)[0].to_dict()
    assert fg.T['gate_skips'] > 0 and fg.T['gate_runs'] > 0
    # the rule, restated: per trigger its tile's chirp-mass range, the gate triggers of H1 in it, in samples of its chunk
    tile_of = {int(i): tb for tb in plan.tiles for i in tb.ids}
    gs = np.round((gate.time - c.gps_start) * c.sr).astype(np.int64)
    w = int(1.0 * c.sr)
    keep = np.zeros(len(full['snr']), bool)
    for i, (tid, seg, samp) in enumerate(zip(full['template_id'], full['segment'], full['sample'])):
        mc = mc_all[tile_of[int(tid)].ids]
        ok = (gate.mchirp >= mc.min() / 1.25) & (gate.mchirp <= mc.max() * 1.25)
        rel = gs[ok] - c.seg_start[seg]
        a0, a1 = c.analyze[seg]
        rel = rel[(rel >= a0 - w) & (rel <= a1 + w)]
        keep[i] = len(rel) > 0 and np.abs(rel - samp).min() <= w
    assert 0 < keep.sum() < len(keep)
    exp = sorted(zip(full['template_id'][keep].tolist(), full['end_time'][keep].tolist(), full['snr'][keep].tolist()))
    assert sorted(zip(got['template_id'].tolist(), got['end_time'].tolist(), got['snr'].tolist())) == exp
