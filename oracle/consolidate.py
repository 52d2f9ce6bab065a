"""CPU restatement of the reference's trigger consolidation (keep-loudest per interval) and of its pre-consolidation
of the GEMM candidates. TEST INFRASTRUCTURE ONLY: see oracle/__init__.py.

  consolidate  fp_filter.py:629-663 (pycbc EventManager.consolidate_events per piece: gate veto, newsnr cut,
               keep-loudest per interval, the K-th loudest "floor" of every full window)
  pre_keep     fp_filter.py:665-695 (the same cut, veto and ranking on one tile and segment's candidates)
  newsnr       pycbc.events.ranking.newsnr of PyCBC 2.8.3 (q = 6, n = 2), which fp_filter.py:250 imports and
               fp_filter.py:639, 930 call; PyCBC is not in this image, so the published formula is restated:
               nsnr = snr * (0.5 (1 + rchisq^3))^(-1/6) where rchisq > 1, float64.
  keep_loudest_brute  an independent per-window loop (Python sort of (-stat, position)), the check that pins the
               ranking of the two functions above on small inputs.

The property that pins the pair (the reference's tests/test_prekeep.py): applying pre_keep to every part and then
consolidate after every part leaves the same triggers as consolidate alone, for random parts, ties, gates, exempt
peaks and floors.
"""
from __future__ import annotations

import numpy as np


def newsnr(snr, rchisq):
    """pycbc.events.ranking.newsnr (PyCBC 2.8.3) with q = 6, n = 2: float64 snr, multiplied by
    (0.5 (1 + rchisq^3))^(-1/6) where rchisq > 1."""
    ns = np.array(snr, ndmin=1, dtype=np.float64)
    r = np.array(rchisq, ndmin=1, dtype=np.float64)
    i = np.nonzero(r > 1.0)[0]
    ns[i] *= (0.5 * (1.0 + r[i] ** 3.0)) ** (-1.0 / 6.0)
    return ns


def window_of(t, gps_start, sr, interval):
    """fp_filter.py:648-649, 685: the keep-loudest window of end time t, int64(round((t - gps_start) sr) / (interval
    sr)) in float64 (round half to even, truncation)."""
    ti = np.round((np.asarray(t, np.float64) - gps_start) * sr)
    return (ti / (interval * sr)).astype(np.int64)


def rank_in_window(wt, stat):
    """Rank of every entry among the entries of its window, by stat descending, ties by position (the stable
    lexsort of fp_filter.py:650-652)."""
    od = np.lexsort((-np.asarray(stat), wt))
    wt_s = wt[od]
    first = np.r_[0, np.nonzero(np.diff(wt_s))[0] + 1] if len(od) else np.zeros(0, np.int64)
    rank = np.empty(len(od), np.int64)
    rank[od] = np.arange(len(od)) - np.repeat(first, np.diff(np.r_[first, len(od)]))
    return rank, od, wt_s, first


def keep_loudest_brute(wt, stat, K):
    """Pure-Python keep-loudest: in every window the K entries with the largest stat, the earlier entry first on
    ties. Returns a boolean mask."""
    keep = np.zeros(len(wt), bool)
    by = {}
    for i, (w, s) in enumerate(zip(np.asarray(wt).tolist(), np.asarray(stat).tolist())):
        by.setdefault(w, []).append((-s, i))
    for lst in by.values():
        for _, i in sorted(lst)[:K]:
            keep[i] = True
    return keep


def consolidate(c, gates, gps_start, sr, newsnr_threshold=None, interval=None, num=None, exempt_snr=None,
                want_floor=False):
    """fp_filter.py:629-663 on one piece's concatenated peaks c (dict of arrays: snr float32, chi float32, dof int32,
    t float64): returns (keep mask, floor or None). ns = newsnr(snr, chi / max(dof, 1)); exempt = snr >= exempt_snr
    (float32 compare); keep = ns >= newsnr_threshold or exempt; the gate veto |t - g_t| > w + taper + 0.5; with an
    interval, the `num` loudest by ns per window among the kept non-exempt peaks, plus the kept exempt ones. The
    floor (want_floor): per window up to the last used one, the num-th loudest ns of every window with at least num
    ranked peaks, -inf elsewhere."""
    snr = np.asarray(c['snr'], np.float32)
    ns = newsnr(snr, np.asarray(c['chi'], np.float32) / np.maximum(np.asarray(c['dof'], np.int32), 1))
    exempt = snr >= np.float32(exempt_snr) if exempt_snr is not None else np.zeros(len(ns), bool)
    keep = (ns >= (newsnr_threshold or 0)) | exempt
    t = np.asarray(c['t'], np.float64)
    for g_t, g_w, g_taper in gates:
        keep &= np.abs(t - g_t) > g_w + g_taper + 0.5
    floor = None
    if interval:
        idx = np.nonzero(keep & ~exempt)[0]
        wt = window_of(t[idx], gps_start, sr, interval)
        rank, od, wt_s, first = rank_in_window(wt, ns[idx])
        gated = keep.copy()
        keep = np.zeros(len(keep), bool)
        keep[idx[rank < num]] = True
        keep |= gated & exempt
        if want_floor:
            full = np.diff(np.r_[first, len(od)]) >= num
            floor = np.full(int(wt_s[-1]) + 1 if len(wt_s) else 1, -np.inf)
            floor[wt_s[first[full]]] = ns[idx[od[first[full] + num - 1]]]
    return keep, floor


def pre_keep(samp_g, apx, gps_start, seg_start, sr, gates, thr, exact_snr, margin, newsnr_threshold=None,
             interval=None, num=None, exempt_snr=None, floor=None):
    """fp_filter.py:665-695 on one tile and segment's candidates: samp_g (int64, a0 + window sample) and apx (float32
    rebuilt SNR). Exact-band candidates (apx >= exact_snr (1 - margin)) are all kept; a rebuilt one (apx > thr) is
    kept if it passes the newsnr cut (its newsnr is its SNR) or is exempt, the gate veto at t = gps_start +
    (seg_start + samp_g) / sr, the floor (apx > floor[window], windows past the floor's end pass) and the
    keep-loudest ranking among this part's non-exempt survivors. Float32 thresholds as torch compares them."""
    apx = np.asarray(apx, np.float32)
    ex = apx >= np.float32(exact_snr * (1 - margin))
    ok = ~ex & (apx > np.float32(thr))
    t = gps_start + (int(seg_start) + np.asarray(samp_g, np.int64)).astype(np.float64) / sr
    exempt = ok & (apx >= np.float32(exempt_snr)) if exempt_snr is not None else np.zeros(len(ok), bool)
    ok &= (apx >= np.float32(newsnr_threshold or 0)) | exempt
    for g_t, g_w, g_taper in gates:
        ok &= np.abs(t - g_t) > g_w + g_taper + 0.5
    idx = np.nonzero(ok & ~exempt)[0]
    if interval and len(idx):
        wt = window_of(t[idx], gps_start, sr, interval)
        if floor is not None:
            fl = np.asarray(floor, np.float64)
            bar = np.where(wt < len(fl), fl[np.minimum(wt, len(fl) - 1)], -np.inf)
            above = apx[idx].astype(np.float64) > bar
            ok[idx[~above]] = False
            idx, wt = idx[above], wt[above]
        rank, _, _, _ = rank_in_window(wt, apx[idx])
        top = np.zeros(len(ok), bool)
        top[idx[rank < num]] = True
        ok = top | (ok & exempt)
    return ex | ok
