"""Trigger consolidation and pre-consolidation (fp_filter.py:629-695) on the CPU: the oracle against brute force and
the reference's randomized union property (tests/test_prekeep.py), the host consolidation against the oracle, the
candidate cut's float32 rounding, and the device tables' window ranges."""
import numpy as np
import pytest
import torch

from oracle import consolidate as oc

SR = 2048.0
GPS = 1380077576.0


def test_newsnr_closed_form():
    """PyCBC newsnr at q = 6, n = 2: snr (0.5 (1 + r^3))^(-1/6) for r > 1, snr otherwise; r = 3 gives
    snr 14^(-1/6), r = 1 (not > 1) leaves snr unchanged."""
    import consolidate
    snr = np.array([8.0, 8.0, 6.0, 7.5], np.float32)
    r = np.array([3.0, 1.0, 0.2, 2.0])
    exp = np.array([8 * 14.0 ** (-1 / 6), 8.0, 6.0, 7.5 * 4.5 ** (-1 / 6)])
    for f in (oc.newsnr, consolidate.newsnr):
        got = f(snr, r)
        assert got.dtype == np.float64
        np.testing.assert_allclose(got, exp, rtol=4e-16, atol=0)


def test_rank_in_window_matches_brute():
    rng = np.random.default_rng(3)
    for trial in range(200):
        n = int(rng.integers(0, 300))
        wt = rng.integers(0, int(rng.integers(1, 12)), n)
        stat = np.round(rng.uniform(4, 7, n), 1)            # many ties
        K = int(rng.integers(1, 6))
        rank = oc.rank_in_window(wt, stat)[0]
        np.testing.assert_array_equal(rank < K, oc.keep_loudest_brute(wt, stat, K), err_msg=str(trial))


def _brute_consolidate(c, gates, thr_ns, interval, num, exempt_snr):
    """fp_filter.py:629-654 written as a loop: cut, veto, then per window the num loudest non-exempt by newsnr,
    earlier first on ties."""
    ns = oc.newsnr(c['snr'], c['chi'].astype(np.float64) / np.maximum(c['dof'], 1))
    keep = np.zeros(len(ns), bool)
    wins = {}
    for i in range(len(ns)):
        ex = exempt_snr is not None and float(c['snr'][i]) >= float(np.float32(exempt_snr))
        ok = ns[i] >= (thr_ns or 0) or ex
        ok = ok and all(abs(c['t'][i] - g) > w + tp + 0.5 for g, w, tp in gates)
        if not ok:
            continue
        if ex or not interval:
            keep[i] = True
            continue
        w = int(round((c['t'][i] - GPS) * SR) / (interval * SR))
        wins.setdefault(w, []).append((-ns[i], i))
    for lst in wins.values():
        for _, i in sorted(lst)[:num]:
            keep[i] = True
    return keep


def _random_peaks(rng, n):
    snr = np.round(rng.uniform(3.9, 7.5, n), 2).astype(np.float32)
    ex = rng.random(n) < 0.5
    return dict(snr=snr, chi=np.where(ex, rng.uniform(0, 60, n), 0).astype(np.float32),
                dof=np.where(ex, 30, 0).astype(np.int32),
                t=GPS + rng.integers(0, 40 * 2048, n) / SR)


def test_consolidate_oracle_vs_brute():
    rng = np.random.default_rng(5)
    for trial in range(120):
        c = _random_peaks(rng, int(rng.integers(0, 600)))
        gates = [(GPS + rng.uniform(0, 40), 0.25, 0.25) for _ in range(int(rng.integers(0, 3)))]
        kw = dict(newsnr_threshold=[4.0, None][trial % 3 == 0], interval=[1.072, None][trial % 5 == 0],
                  num=int(rng.integers(1, 5)), exempt_snr=[5.0, None][trial % 2])
        keep, _ = oc.consolidate(c, gates, GPS, SR, **kw)
        exp = _brute_consolidate(c, gates, kw['newsnr_threshold'], kw['interval'], kw['num'], kw['exempt_snr'])
        np.testing.assert_array_equal(keep, exp, err_msg=str(trial))


def test_prekeep_union_property():
    """The reference's tests/test_prekeep.py: pre_keep on every part, consolidate after every part, gives the
    triggers of consolidate alone (300 trials, rng(1): random parts, rounded SNRs with many ties, gates, exempt
    peaks, keep-loudest on and off, floors)."""
    rng = np.random.default_rng(1)
    thr, exact_snr, margin = 4.0, 5.5, 0.04
    seg_start = [0, 1636352]
    fails = 0
    for trial in range(300):
        kw = dict(newsnr_threshold=4.0, interval=[1.072, None][trial % 7 == 0], num=int(rng.integers(1, 6)),
                  exempt_snr=[5.0, None][trial % 2])
        gates = [(GPS + rng.uniform(0, 1500), 0.25, 0.25) for _ in range(int(rng.integers(0, 3)))]
        runs = {'full': dict(parts=[], floor=None), 'pre': dict(parts=[], floor=None)}
        for part in range(int(rng.integers(1, 6))):
            s = int(rng.integers(0, 2)); n = int(rng.integers(0, 3000))
            samp = rng.integers(0, int(rng.choice([20, 800])) * 2048, n)
            apx = np.round(rng.uniform(3.84, 7, n), 2).astype(np.float32)
            ex = apx >= np.float32(exact_snr * (1 - margin)); rb = ~ex & (apx > np.float32(thr))
            chi_ex = rng.uniform(0, 3, n).astype(np.float32) * 20
            kp = oc.pre_keep(samp, apx, GPS, seg_start[s], SR, gates, thr, exact_snr, margin,
                             floor=runs['pre']['floor'], **kw)
            for name, keep in (('full', np.ones(n, bool)), ('pre', kp)):
                r = runs[name]
                for sel, isex in ((rb & keep, False), (ex & keep, True)):
                    k = int(sel.sum())
                    r['parts'].append(dict(tid=np.nonzero(sel)[0] + 10000 * part,
                                           t=GPS + (seg_start[s] + samp[sel]) / SR, snr=apx[sel],
                                           chi=chi_ex[sel] if isex else np.zeros(k, np.float32),
                                           dof=np.full(k, 10 if isex else 0, np.int32)))
                c = {k: np.concatenate([p[k] for p in r['parts']]) for k in r['parts'][0]}
                keep_c, floor = oc.consolidate(c, gates, GPS, SR, want_floor=(name == 'pre'), **kw)
                r['parts'] = [{k: v[keep_c] for k, v in c.items()}]
                if name == 'pre' and floor is not None:
                    r['floor'] = floor
        key = lambda r: sorted(zip(r['parts'][0]['tid'].tolist(), r['parts'][0]['t'].tolist()))
        fails += key(runs['full']) != key(runs['pre'])
    assert fails == 0


def test_prekeep_oracle_drops_something():
    """The property above is not vacuous: with a small num the pre-consolidation drops most rebuilt candidates."""
    rng = np.random.default_rng(2)
    n = 3000
    samp = rng.integers(0, 20 * 2048, n)
    apx = np.round(rng.uniform(3.84, 7, n), 2).astype(np.float32)
    kp = oc.pre_keep(samp, apx, GPS, 0, SR, [], 4.0, 5.5, 0.04, newsnr_threshold=4.0, interval=1.072, num=2)
    rb = (apx > 4.0) & (apx < np.float32(5.28))
    assert kp[rb].sum() <= 2 * 20 and kp[apx >= np.float32(5.28)].all()


def test_host_consolidate_matches_oracle():
    import consolidate
    import fp_filter
    rng = np.random.default_rng(9)
    for trial in range(60):
        tr = fp_filter.Triggers()
        parts = []
        for _ in range(int(rng.integers(0, 4))):
            c = _random_peaks(rng, int(rng.integers(0, 400)))
            n = len(c['snr'])
            c.update(tid=rng.integers(0, 500, n), ph=np.zeros(n, np.float32), sg=np.ones(n), ex=c['dof'] > 0,
                     samp=np.zeros(n, np.int64), seg=np.zeros(n, np.int64))
            tr.add(**c)
            parts.append(c)
        gates = [(GPS + rng.uniform(0, 40), 0.25, 0.25) for _ in range(int(rng.integers(0, 3)))]
        kl = consolidate.KeepLoudest(4.0, [1.072, None][trial % 5 == 0], int(rng.integers(1, 5)),
                                     [5.0, None][trial % 2])
        floor, dropped = consolidate.consolidate(tr, gates, GPS, SR, kl, want_floor=True)
        if not parts:
            assert floor is None and dropped == 0
            continue
        c = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
        keep, fl = oc.consolidate(c, gates, GPS, SR, kl.newsnr_threshold, kl.interval, kl.num, kl.exempt_snr,
                                  want_floor=True)
        assert dropped == int((~keep).sum())
        np.testing.assert_array_equal(tr.cat('tid', np.int64), c['tid'][keep])
        np.testing.assert_array_equal(tr.cat('t', np.float64), c['t'][keep])
        assert (fl is None) == (floor is None)
        if fl is not None:
            np.testing.assert_array_equal(floor, fl)


def test_prefix_consolidation_equals_whole():
    """Consolidating only the leading parts whose fields are final (later parts pending), then everything, keeps
    the same peaks as consolidating everything once."""
    import consolidate
    import fp_filter
    rng = np.random.default_rng(12)
    kl = consolidate.KeepLoudest(4.0, 1.072, 2, 5.0)
    for trial in range(40):
        a, b = fp_filter.Triggers(sg=True), fp_filter.Triggers(sg=True)
        for part in range(int(rng.integers(2, 6))):
            c = _random_peaks(rng, int(rng.integers(0, 300)))
            n = len(c['snr'])
            c.update(tid=rng.integers(0, 500, n), ph=np.zeros(n, np.float32), sg=np.ones(n), ex=c['dof'] > 0,
                     samp=np.zeros(n, np.int64), seg=np.zeros(n, np.int64))
            a.add(**c)
            b.add(**c)
            if part == 1:
                a.unready.append(a.parts['sgc'][-1])           # this part's sine-Gaussian values pending
                consolidate.consolidate(a, [], GPS, SR, kl)
                assert a.parts['sgc'][-1] is a.unready[0]     # left as it was, after the consolidated prefix
        a.unready = []
        consolidate.consolidate(a, [], GPS, SR, kl)
        consolidate.consolidate(b, [], GPS, SR, kl)
        assert sorted(zip(a.cat('tid', np.int64).tolist(), a.cat('t', np.float64).tolist())) == \
            sorted(zip(b.cat('tid', np.int64).tolist(), b.cat('t', np.float64).tolist()))


def test_candidate_cut_rounds_like_torch():
    """fp_filter.py:747 compares the float32 window maxima with the Python float (thr (1 - margin))^2: torch rounds
    the float64 square to float32. At thr 5.5, margin 0.04 that is 27.8784 (float32), not the square of the
    float32 5.28 (27.878403): the float32 value just above the cut is a candidate, the cut itself is not."""
    import fp_filter
    from oracle import core
    for thr, margin in ((5.5, 0.04), (4.0, 0.02), (6.0, 0.1), (5.0, 0.04)):
        cut = fp_filter.candidate_cut(thr, margin)
        for v in (np.nextafter(cut, np.float32(0)), cut, np.nextafter(cut, np.float32(100))):
            ref = bool(torch.tensor([v], dtype=torch.float32) > (thr * (1 - margin)) ** 2)
            assert bool(np.float32(v) > cut) == ref, (thr, margin, v)
            vw = np.array([[v]], np.float32)
            assert (len(core.candidates(vw, np.zeros((1, 1), np.int32), thr, margin, 0)[0]) == 1) == ref
    assert fp_filter.candidate_cut(5.5, 0.04) == np.float32(27.8784)


@pytest.mark.parametrize('seed', [0, 1, 2])
def test_prekeep_window_tables_cover_every_sample(seed):
    """PreKeep.tables: every analysed sample's keep-loudest window (as the oracle computes it) lies in the segment's
    [wt0, wt0 + nwin), and the segments' window ranges are consecutive in the batch."""
    import consolidate
    rng = np.random.default_rng(seed)
    kl = consolidate.KeepLoudest(4.0, [1.072, 0.5, 3.3][seed], 3, None)
    pk = consolidate.PreKeep(kl, 4.0, 5.5, 0.04, SR, 2048)
    segs, col0 = [], 0
    for _ in range(3):
        nw = int(rng.integers(1, 40))
        gps = GPS + float(rng.integers(0, 1000))
        segs.append((int(rng.integers(0, 1 << 21)), gps, col0, nw, [], None))
        col0 += nw
    segi, segd, gates, fls, nwin = pk.tables(segs)
    assert segi[:, 4].tolist() == np.r_[0, np.cumsum(segi[:-1, 5])].tolist() and nwin == segi[:, 5].sum()
    for (sbase, gps, _, nw, _, _), row in zip(segs, segi):
        t = gps + (sbase + np.arange(nw * 2048)) / SR
        wt = oc.window_of(t, gps, SR, kl.interval)
        assert wt.min() == row[1] and wt.max() == row[1] + row[5] - 1
