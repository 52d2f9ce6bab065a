"""The batched form: detector chunks of uneven length (segment gaps), both detectors and both banks, filtered as a
few large launches instead of one pass per chunk.

Chunks are grouped by segment grid: every chunk on the 512 s grid (tlen 2^20) and every chunk on the 128 s grid
(tlen 2^18) of a step form one Filter each, whose segments are taken `segment_batch` at a time per basis-transform
batch and GEMM launch (the item table of csrc/gemm_peaks.hip covers several segments of uneven analysed length in
one launch). Each chunk's PSD is its own (fp_filter.py's prep per dataset). Bank products (dense templates and
float64 SVD bases) are built once per grid, since Vh and H live on the grid's frequency bins; the SVD metric of a
grid is sqrt(mean IP) over the segments of the reference chunks (fp_filter.py:360-364 takes it over the group being
filtered; it sets only how well the basis compresses, never the SNR: conj(h sw) (D / sw) = conj(h) D)."""
from __future__ import annotations

import numpy as np

import fp_filter
import workloads


class BatchedSearch:
    """chunks: [(Stream, Conditioning or None)] as workloads.batched / batched_device return them; banks: synth banks;
    x_dev: the chunks' float32 strain on the GPU (kept resident; each step conditions copies of it); opts: {form:
    fp_filter.Options}; ref: indices of the chunks whose conditioned data give each grid's SVD metric (default: all
    chunks)."""

    def __init__(self, chunks, banks, x_dev, opts: dict, ref=None):
        self.chunks = [(st, c, x) for (st, c), x in zip(chunks, x_dev)]
        self.banks = banks
        self.opts = opts
        self.grids = sorted({c.seg_len for _, c, _ in self.chunks if c is not None})
        ds = self.condition(range(len(self.chunks)) if ref is None else ref)
        missing = [g for g in self.grids if not ds[g]]
        assert not missing, f'no reference chunk on the {missing} s grid(s)'
        self.plans = {}
        first = next(iter(opts.values()))
        for g in self.grids:
            IPall, sw, isw = fp_filter.svd_metric(ds[g])
            self.plans[g] = [fp_filter.BankPlan(b.m1, b.m2, b.e0, ds[g][0].flen, ds[g][0].df, first, sw, isw,
                                                name=f'{b.name}@{g}s') for b in banks]

    def condition(self, sel):
        """Condition the selected chunks (a fresh copy of each strain); datasets grouped by segment grid, in chunk
        order."""
        out = {g: [] for g in self.grids}
        for i in sel:
            st, c, x = self.chunks[i]
            if c is not None:
                out[c.seg_len] += workloads.condition_streams([st], c, [x.clone()])
        return out

    def step(self, prec, sel=None, self_check=False):
        """One pass over the selected chunks (default: all) with both banks. Returns ({grid: datasets}, {grid: [per
        bank [per dataset trigger dict]]}, [Filter])."""
        ds = self.condition(range(len(self.chunks)) if sel is None else sel)
        o = self.opts[prec]
        if self_check:
            o = fp_filter.Options(**{**o.__dict__, 'self_check': True})
        outs, filters = {}, []
        for g in self.grids:
            if not ds[g]:
                outs[g] = [[] for _ in self.banks]
                continue
            f = fp_filter.Filter(ds[g], o)
            outs[g] = [[t.to_dict(gates=d.gates) for t, d in zip(tr, ds[g])] for tr in f.run_banks(self.plans[g])]
            filters.append(f)
        return ds, outs, filters

    def analysed_seconds(self, ds, cw=2048):
        return sum(workloads.analysed_seconds(ds[g], cw) for g in self.grids)

    def ranks(self):
        return {p.name: p.ranks for g in self.grids for p in self.plans[g]}


def slices(n_chunks, ndet, nslice):
    """The chunk indices of each of nslice equal slices of a batched workload (chunks ordered chunk-major,
    detector-minor, as workloads.batched makes them): slice k holds chunks k n / nslice .. (k + 1) n / nslice - 1 of
    every detector."""
    per = n_chunks // nslice
    assert per * nslice == n_chunks, 'the slices must cover the chunks exactly'
    return [list(range(k * per * ndet, (k + 1) * per * ndet)) for k in range(nslice)]


def grid_summary(chunks):
    """(chunks per grid, analysed seconds estimate) for a chunk list, without touching the GPU."""
    cnt = {}
    for _, c in chunks:
        if c is not None:
            cnt[c.seg_len] = cnt.get(c.seg_len, 0) + 1
    return cnt
