"""Overlap-save chunking (window_filter.py:39-47, fp_filter.py:324-357) on the CPU: the oracle's chunks pinned by the
overlap-save property, and the product's to_chunks (torch on the CPU) against the oracle."""
import numpy as np
import pytest
import torch

from oracle import core, pipeline, window as ow

SR = 2048.0


@pytest.fixture(scope='module')
def seg64():
    """64 s of seeded coloured noise with one SNR 20 injection, conditioned by the oracle; 6 templates of total mass
    40-50 Msun (at most 4 s long from 15 Hz)."""
    import synth
    bank = synth.circular_bank(6, seed=3, mmin=20, mmax=50)
    st = synth.make_stream('H1', 64, SR, seed=1,
                           injections=[synth.Injection(30.0, float(bank.m1[0]), float(bank.m2[0]), 20.0, 0.3)])
    ref = core.condition(st.x, SR, st.gps_start, 64, 8, 8, 4, 2, 31, 4, 15.0)
    return bank, ref


@pytest.mark.parametrize('spec', ['16,8,2', '16,8,2,256', '32,12,2'])
def test_overlap_save_property(seg64, spec):
    """On every chunk's valid samples the direct matched-filter numerator of the chunk, scaled by df_chunk / df, is
    the whole segment's at the same absolute sample, to within the templates' frequency-cut tails beyond HIST
    (measured 1.4e-3 to 3.1e-3 of max |z|; a start off by one sample or a missing dt is off by 1e-1 or more)."""
    bank, ref = seg64
    D, IP = ref['D'].astype(np.complex128), ref['IP'].astype(np.float64)
    H, _ = pipeline.spa_bank(bank.m1, bank.m2, None, ref['flen'], ref['df'], 15.0)
    zw = core.direct_z(H, D[0], ref['tlen'])
    c = ow.to_chunks(D, IP, ref['analyze'], ref['seg_start'], ref['tlen'], SR, spec)
    Hc, _ = pipeline.spa_bank(bank.m1, bank.m2, None, c['flen'], c['df'], 15.0)
    covered = np.zeros(ref['tlen'], int)
    worst = 0.0
    for j, (a, b) in enumerate(c['analyze']):
        zc = core.direct_z(Hc, c['D'][j], c['tlen']) * (c['df'] / ref['df'])
        ws = (c['seg_start'][j] + np.arange(a, b)) * c['dec'] - ref['seg_start'][0]
        worst = max(worst, float(np.abs(zc[:, a:b] - zw[:, ws]).max() / np.abs(zw).max()))
        covered[ws] += 1
    assert worst < 4e-3
    # the valid ranges tile the analysed span that has HIST of history and FUT of future once (decimated rate)
    v = [float(x) for x in spec.split(',')]
    a0, a1 = max(ref['analyze'][0][0], int(v[1] * SR)), min(ref['analyze'][0][1], ref['tlen'] - int(v[2] * SR))
    assert np.all(covered[a0:a1:c['dec']] == 1) and covered.sum() == len(range(a0, a1, c['dec']))


@pytest.mark.parametrize('spec', ['16,8,2', '16,8,2,512'])
def test_to_chunks_matches_oracle(seg64, spec):
    import fp_filter, window_filter
    _, ref = seg64
    ds = fp_filter.Dataset('H1', torch.from_numpy(ref['D']), torch.from_numpy(ref['IP']), ref['flen'], ref['tlen'],
                           ref['df'], SR, ref['gps_start'], list(ref['analyze']), list(ref['seg_start']))
    (c,), hist = window_filter.to_chunks([ds], spec)
    o = ow.to_chunks(ref['D'], ref['IP'], ref['analyze'], ref['seg_start'], ref['tlen'], SR, spec)
    assert hist == float(spec.split(',')[1])
    assert (c.tlen, c.flen, c.df, c.sr) == (o['tlen'], o['flen'], o['df'], o['sr'])
    assert c.analyze == o['analyze'] and c.seg_start == o['seg_start']
    Dc = c.D.numpy()
    assert Dc.shape == o['D'].shape and np.array_equal(c.IP.numpy(), o['IP'].astype(np.float32))
    err = np.sqrt(np.mean(np.abs(Dc - o['D']) ** 2) / np.mean(np.abs(o['D']) ** 2))
    assert err < 1e-5
