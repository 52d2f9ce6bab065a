"""The batched form: detector chunks of uneven length (segment gaps), both detectors and both banks, filtered as a
few large launches instead of one pass per chunk.

Chunks are grouped by segment grid (workloads.chunk_conditioning: 512 s, 256 s and 128 s segments, tlen 2^20, 2^19
and 2^18 at 2048 Hz): the chunks of a step on one grid form one Filter, whose segments are taken in batches of up to
`segment_batch` segments and `batch_samples` analysed samples per basis-transform batch (one launch of each
transform pass per 16 segments) and GEMM launch (the item table of csrc/gemm_peaks.hip covers several segments of
uneven analysed length in one launch). Each chunk's PSD is its own (fp_filter.py's prep per dataset). Bank products
(dense templates and float64 SVD bases) are built once per grid, since Vh and H live on the grid's frequency bins;
the SVD metric of a grid is sqrt(mean IP) over the segments of the reference chunks (fp_filter.py:360-364 takes it
over the group being filtered; it sets only how well the basis compresses, never the SNR: conj(h sw) (D / sw) =
conj(h) D)."""
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
        """Condition the selected chunks; datasets grouped by segment grid, in chunk order. Chunks without autogating
        are conditioned together per grid (gpu_prep.condition_many: a few batched launches, no host synchronisation;
        the strain is read, not modified); a chunk with autogating is gated on a fresh copy of its strain and
        conditioned alone."""
        import gpu_prep
        out = {g: [] for g in self.grids}
        batch = {g: [] for g in self.grids}
        for i in sel:
            st, c, x = self.chunks[i]
            if c is None:
                continue
            if c.autogating_threshold is not None:
                out[c.seg_len] += workloads.condition_streams([st], c, [x.clone()])
            else:
                batch[c.seg_len].append((st, c, x))
        for g, items in batch.items():
            if not items:
                continue
            c0 = items[0][1]
            assert all((c.start_pad, c.end_pad, c.psd_seg_len, c.psd_stride, c.psd_inv_len) ==
                       (c0.start_pad, c0.end_pad, c0.psd_seg_len, c0.psd_stride, c0.psd_inv_len) for _, c, _ in items)
            out[g] += gpu_prep.condition_many([x for _, _, x in items], items[0][0].sr,
                                              [st.gps_start for st, _, _ in items], g, c0.start_pad, c0.end_pad,
                                              c0.psd_seg_len, c0.psd_stride, [c.psd_nseg for _, c, _ in items],
                                              c0.psd_inv_len, workloads.FLOW, tags=[st.ifo for st, _, _ in items])
        return out

    def _jobs(self, prec, sels, self_check, rec):
        """The (Filter, plans) of every (slice, grid) in order, set up lazily: a slice is conditioned when the
        pipeline reaches it, while the previous slice's last batches are still in flight."""
        o = self.opts[prec]
        if self_check:
            o = fp_filter.Options(**{**o.__dict__, 'self_check': True})
        for k, sel in enumerate(sels):
            ds = self.condition(sel)
            for g in self.grids:
                if ds[g]:
                    f = fp_filter.Filter(ds[g], o)
                    rec.append((k, g, ds[g], f))
                    yield f, self.plans[g]

    def run(self, prec, sels, self_check=False):
        """Several slices (lists of chunk indices) in one software pipeline with both banks. Returns per slice
        ({grid: datasets}, {grid: [per bank [per dataset trigger dict]]}, [Filter])."""
        rec = []
        res = fp_filter.run_jobs(self._jobs(prec, sels, self_check, rec))
        out = [({g: [] for g in self.grids}, {g: [[] for _ in self.banks] for g in self.grids}, []) for _ in sels]
        for (k, g, dsg, f), tr in zip(rec, res):
            out[k][0][g] = dsg
            out[k][1][g] = [[t.to_dict(gates=d.gates) for t, d in zip(tb, dsg)] for tb in tr]
            out[k][2].append(f)
        return out

    def step(self, prec, sel=None, self_check=False):
        """One pass over the selected chunks (default: all) with both banks. Returns ({grid: datasets}, {grid: [per
        bank [per dataset trigger dict]]}, [Filter])."""
        return self.run(prec, [range(len(self.chunks)) if sel is None else sel], self_check)[0]

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
