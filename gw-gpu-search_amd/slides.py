"""Time slides over two detectors' trigger lists (the reference's coinc_rank.find_pairs, coinc_rank.py:100-138):
every (H1 trigger, L1 trigger, slide k) of the same template with |tH - (tL + k step) mod P| <= window, |k| <= K.
The per-detector SNRs are reused: a slide is only a shift of one list. Host code builds the sorted keys exactly as
the reference does (L copies at -P, 0, +P, key = tid * big + t); the range searches and the pair output run on the
GPU (csrc/slides.hip), in the reference's order (H index, then sorted L position)."""
from __future__ import annotations

import numpy as np
import torch

import _dev
import _lib


def find_pairs(H, L, P, step, K, window, stat_fn=None, min_stat=-np.inf):
    """H, L: dicts with 'x' (times in [0, P), float64) and 'tid'. Returns (h_idx, l_idx, k, stat) as numpy arrays;
    stat = stat_fn(h_idx, l_idx, dt) with dt = (tL + k step) - tH, or dt itself without stat_fn; pairs with
    stat < min_stat are dropped."""
    reach = K * step + window
    if 2 * reach >= P:
        raise SystemExit(f'--max-slides {K} x --slide-step {step} covers the whole span ({P:.0f} s); lower it')
    lx = np.asarray(L['x'], np.float64)
    lt = np.concatenate([lx - P, lx, lx + P])
    li = np.tile(np.arange(len(lx)), 3)
    lk = np.tile(np.asarray(L['tid']), 3)
    big = 4.0 * P + 10 * reach
    lkey = lk * big + lt
    o = np.argsort(lkey, kind='stable')
    lkey, li, lt = lkey[o], li[o], lt[o]
    hx = np.asarray(H['x'], np.float64)
    hkey = np.asarray(H['tid']) * big + hx
    nh, nl = len(hx), len(lkey)
    d = lambda a: _dev.as_dev(np.ascontiguousarray(a))
    hk, hxd, lkd, ltd, lid = d(hkey), d(hx), d(lkey), d(lt), d(li.astype(np.int32))
    counts = torch.empty(nh + 1, dtype=torch.int64, device='cuda')
    total = counts[nh:]
    _lib.call('gwgs_slide_count', hk.data_ptr(), hxd.data_ptr(), nh, lkd.data_ptr(), ltd.data_ptr(), lid.data_ptr(),
              nl, float(reach), float(step), int(K), float(window), counts.data_ptr(), total.data_ptr(),
              _dev.stream())
    n = int(total.item())
    oh = torch.empty(max(n, 1), dtype=torch.int32, device='cuda'); ol = torch.empty_like(oh)
    ok = torch.empty_like(oh); odt = torch.empty(max(n, 1), dtype=torch.float64, device='cuda')
    if n:
        _lib.call('gwgs_slide_write', hk.data_ptr(), hxd.data_ptr(), nh, lkd.data_ptr(), ltd.data_ptr(),
                  lid.data_ptr(), nl, float(reach), float(step), int(K), float(window), counts.data_ptr(),
                  oh.data_ptr(), ol.data_ptr(), ok.data_ptr(), odt.data_ptr(), _dev.stream())
    hh, ll, kk, dt = (x[:n].cpu().numpy() for x in (oh, ol, ok, odt))
    hh, ll, kk = hh.astype(np.int64), ll.astype(np.int64), kk.astype(np.int64)
    st = stat_fn(hh, ll, dt) if stat_fn is not None else dt
    m = st >= min_stat
    return hh[m], ll[m], kk[m], st[m]
