"""The batched form: detector chunks of uneven length (segment gaps), both detectors and both banks, filtered as a
few large launches instead of one pass per chunk.

Chunks are conditioned one by one (their PSDs are their own, fp_filter.py's prep per dataset) and grouped by
segment grid: every chunk on the 512 s grid (tlen 2^20) and every chunk on the 128 s grid (tlen 2^18) form one
Filter each, whose segments are taken `segment_batch` at a time per basis-transform batch and GEMM launch (the item
table of csrc/gemm_peaks.hip covers several segments of uneven analysed length in one launch). Bank products (dense
templates and float64 SVD bases) are built once per grid, since Vh and H live on the grid's frequency bins; the SVD
metric of a grid is sqrt(mean IP) over its segments (fp_filter.py:360-364)."""
from __future__ import annotations

import numpy as np
import torch

import fp_filter
import workloads


class BatchedSearch:
    """chunks: [(Stream, Conditioning or None)] as workloads.batched returns them; banks: synth banks; x_dev: the
    chunks' float32 strain on the GPU (kept resident; each step conditions copies of it)."""

    def __init__(self, chunks, banks, x_dev, opts: dict):
        self.chunks = [(st, c, x) for (st, c), x in zip(chunks, x_dev) if c is not None]
        self.banks = banks
        self.opts = opts
        self.grids = sorted({c.seg_len for _, c, _ in self.chunks})
        ds = self.condition()
        self.plans = {}
        for g in self.grids:
            IPall, sw, isw = fp_filter.svd_metric(ds[g])
            first = next(iter(opts.values()))
            self.plans[g] = [fp_filter.BankPlan(b.m1, b.m2, b.e0, ds[g][0].flen, ds[g][0].df, first, sw, isw,
                                                name=f'{b.name}@{g}s') for b in banks]

    def condition(self):
        """Condition every chunk (a fresh copy of its strain); datasets grouped by segment grid."""
        out = {g: [] for g in self.grids}
        for st, c, x in self.chunks:
            out[c.seg_len] += workloads.condition_streams([st], c, [x.clone()])
        return out

    def step(self, prec, self_check=False):
        """One pass over every chunk with both banks. Returns ({grid: datasets}, {grid: [per bank [per dataset
        trigger dict]]}, [Filter])."""
        ds = self.condition()
        o = self.opts[prec]
        if self_check:
            o = fp_filter.Options(**{**o.__dict__, 'self_check': True})
        outs, filters = {}, []
        for g in self.grids:
            f = fp_filter.Filter(ds[g], o)
            outs[g] = [[t.to_dict(gates=d.gates) for t, d in zip(tr, ds[g])] for tr in f.run_banks(self.plans[g])]
            filters.append(f)
        return ds, outs, filters

    def analysed_seconds(self, ds, cw=2048):
        return sum(workloads.analysed_seconds(ds[g], cw) for g in self.grids)

    def ranks(self):
        return {p.name: p.ranks for g in self.grids for p in self.plans[g]}


def grid_summary(chunks):
    """(chunks per grid, analysed seconds estimate) for a chunk list, without touching the GPU."""
    cnt = {}
    for _, c in chunks:
        if c is not None:
            cnt[c.seg_len] = cnt.get(c.seg_len, 0) + 1
    return cnt
