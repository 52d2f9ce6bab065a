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
