"""Matched filter of template banks against conditioned detector data by template compression, on one MI355X
(the reference's fp_filter.py run_group, :203-960, for its hot path).

Per tile of templates sorted by chirp time (once per run, BankPlan):
  - the sw-whitened templates are compressed by a float64 SVD to r complex basis filters such that every template
    keeps >= 1 - eps of its power (svd_tile, fp_filter.py:576-593);
Per tile and data segment (Filter.run: launch_batch, decode_batch, collect_batch):
  - the r basis filters are correlated with the segment by the inverse transform of
    q[j, k] = conj(Vh[j, k]) D[k] / sw[k] (basis_series, fp_filter.py:597-613): a fused four-step FFT of two HIP
    passes that forms q on load and writes the GEMM's time-major operand (fp32, bf16 or fp8) directly;
  - one fused MFMA GEMM rebuilds every template's SNR and keeps, per cluster window, only the max |z|^2 and its
    sample (fp_gemm.py:14-44 with fp_filter.py:745-746 fused into its epilogue), in fp32, bf16 or fp8 operands with
    fp32 accumulation;
  - windows whose rebuilt SNR is above thr (1 - margin) are candidates (fp_filter.py:747); the rebuilt complex SNR
    at each gives the rebuilt phase (:759-764); candidates below exact_snr (1 - margin) keep their rebuilt values
    (:768-781), the rest go to the exact stage: edges and chi-squared sums over the exact template in fp32 with
    float64 finalization (:783-853);
  - with opts.consolidate, every dataset's peaks are consolidated after each tile (keep-loudest per interval,
    fp_filter.py:629-663), and with opts.pre_consolidate the candidates that consolidation would drop are removed
    on the device before the candidate list is read back (fp_filter.py:665-695).
So snr, coa_phase and chisq of exact peaks are exact; the compressed filter chooses which windows to look at and the
sample within the window.

Trigger fields and their order are those of the reference's peak files (fp_filter.py:910-937); Triggers.to_dict()
returns them (HDF5 writing is outside this path).
"""
from __future__ import annotations

import ctypes
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np
import torch

import _dev
import _lib
import consolidate
import fp_gemm
import peak_kernel
import sgchisq
import templates as tmpl
from exact_stage import ExactStage, PeakRequest, SegmentData, TemplateTile

CH = 1 << 16          # frequency chunk for the Gram and basis products
SERIES_SEG = 16       # segments per launch of the fused basis transform (its kernels' limit)
BN = fp_gemm.BN
PRECISIONS = {'fp32': 0, 'bf16': 1, 'fp8': 2, 'fp16': 3}
OPERAND_DTYPES = {0: torch.float32, 1: torch.bfloat16, 2: torch.uint8, 3: torch.float16}
# Options.alt_second_pass by precision: where the list epilogue costs a large share of the GEMM (fp32: +25%, fp8:
# +50% over the window-max epilogue on MI355X) a second pass over the hot windows pays while they are a small
# fraction; for bf16 / fp16 (+9%) the first pass keeps the lists.
SECOND_PASS_DEFAULT = {'fp32': 0.08, 'fp8': 0.2, 'bf16': 0.0, 'fp16': 0.0}


@dataclass
class Options:
    snr_threshold: float = 5.5
    cluster_window: float = 1.0           # s
    chisq_bins: int = 16
    low_frequency_cutoff: float = 15.0
    tile: int = 512
    eps: float = 1e-4                      # worst lost power allowed by the SVD (0.9999 kept)
    margin: float = 0.04
    exact_snr: float = 0.0                 # 0: exact values for every candidate
    precision: str = 'fp32'                # GEMM operands: fp32 (plain form), bf16, or fp8 e4m3 with scales
    gemm_cfg: tuple | None = None          # (tiles_per_block,), from fp_gemm.pick_cfg
    self_check: bool = True                # the exact stage's first-call check
    timing: bool = False                   # synchronize around stages to fill the timing dict
    segment_batch: int = 32                # at most this many segments per basis-transform batch and GEMM launch
    batch_samples: int = 1 << 22           # and at most this many analysed samples per batch (a batch's operand Zt
                                           # takes batch_samples x 2 rp x its element size; one segment always fits)
    depth: int = 2                         # batches enqueued ahead of the one being decoded
    graphs: bool = False                   # capture each tile's norms and each batch's stage A in a HIP graph on
                                           # the first run and replay them on later runs (same segmentation; new
                                           # data through rebind)
    refine: int = 0                        # exact stage re-centres each peak on the best exact |z| within +-refine
                                           # samples inside its window, climbing on from a best on the range's edge
                                           # (0: the reference's rebuilt-argmax sample)
    alternates: int = 3                    # with refine: up to this many other local maxima of each candidate's
                                           # window evaluated too (those with |z| >= (1 - alt_margin) its maximum),
                                           # the largest exact |z| kept (0 to 3)
    alt_margin: float = 0.01
    alt_sep: int = 10                      # local maxima closer than alt_sep + 1 samples count once (the larger)
    alt_second_pass: float = None          # with alternates: the GEMM keeps only window maxima, and a second GEMM
                                           # pass over the windows holding exact-band candidates makes their
                                           # local-maximum lists, while such windows are under this fraction of a
                                           # batch's (measured on the batches decoded so far); above it the lists
                                           # are kept in the first pass; 0: always in the first pass (the lists are
                                           # the same either way); None: SECOND_PASS_DEFAULT by precision
    consolidate: object = None             # consolidate.KeepLoudest: consolidate every dataset's peaks after each
                                           # tile (None: keep every peak)
    pre_consolidate: bool = False          # also drop on the device the candidates consolidation would drop
    sgchisq_snr_threshold: float = None    # sine-Gaussian chi-squared of exact peaks with newsnr >= this, for the
                                           # templates BankPlan(sg_locations=...) gives tiles (None: off)
    gate: object = None                    # Gate: filter a tile on a segment only near base-search triggers
    overlap: bool = False                  # basis transforms on their own stream, overlapping the previous batch's
                                           # GEMM (off with graphs: each batch is captured on one stream)
    templates: str = 'closed'              # 'closed': the closed-form chirps on every bin; 'interp': the same
                                           # chirps at coarse nodes, interpolated (fp_filter.py --templates interp)
    exact_stage: str = 'dense'             # 'regen': the exact stage regenerates the interpolated templates from
                                           # their nodes instead of reading H (--exact-stage regen; needs interp)
    n_chirp: int = 6000                    # nodes uniform in f^(-5/3) and in log f (--n-chirp, --n-log)
    n_log: int = 1500


@dataclass
class Dataset:
    """One detector's conditioned data, as gpu_prep.prep returns it: D (nseg, flen) complex64 and IP (nseg, flen)
    float32 on the GPU, zero below kmin; per segment the analysed sample range and the segment start sample."""
    tag: str
    D: torch.Tensor
    IP: torch.Tensor
    flen: int
    tlen: int
    df: float
    sr: float
    gps_start: float
    analyze: list
    seg_start: list
    gates: list = field(default_factory=list)
    first: list = None         # with overlap-save chunks: the first chunk of each original segment (the peak file's
                               # sigmasq columns, fp_filter.py:878); None: every segment


@dataclass
class TileBasis:
    ids: np.ndarray            # bank rows of this tile, in tile order
    H: torch.Tensor            # (n, flen) complex64
    kmax: np.ndarray
    r: int
    rp: int                    # r padded to a multiple of 16
    A: torch.Tensor            # (n, r) complex64: z = A @ basis series
    Vc: torch.Tensor           # (rp, flen) complex64: conj(Vh) / sw, zero rows >= r and columns outside [kmin, kend)
    worst: float
    stage: ExactStage = None
    kend: int = 0              # max template kmax of the tile: Vc is zero from here on


def svd_tile(H, sw, isw, kmin, kend, eps):
    """fp_filter.py:576-593 with the Gram, eigendecomposition and projection in complex128 on the GPU: normalised
    Gram of the sw-weighted rows over [kmin, kend); eigh; worst lost power after k + 1 vectors; r = #(worst > eps)
    + 1. Returns r, rp, A (n, r) complex64, Vc (rp, flen) complex64 = conj(Vh) / sw, and worst[r - 1]."""
    n, flen = H.shape
    G = torch.zeros((n, n), dtype=torch.complex128, device=H.device)
    for c0 in range(kmin, kend, CH):
        c1 = min(kend, c0 + CH)
        hc = H[:, c0:c1].to(torch.complex128) * sw[c0:c1]
        G += hc @ hc.conj().T
    nrm = torch.sqrt(G.diagonal().real.clamp(min=1e-30))
    lam, W = torch.linalg.eigh(G / (nrm[:, None] * nrm[None, :]))
    lam, W = lam.flip(0).clamp(min=0), W.flip(1)
    worst = (1 - ((W.abs() ** 2) * lam[None, :]).cumsum(1)).max(0).values
    r = min(int((worst > eps).sum().item()) + 1, n)
    Wr = W[:, :r]
    s_r = torch.sqrt(lam[:r].clamp(min=1e-12 * float(lam[0])))
    Bm = (Wr.conj().T / s_r[:, None]) / nrm[None, :]
    rp = -(-r // 16) * 16
    Vc = torch.zeros((rp, flen), dtype=torch.complex64, device=H.device)
    for c0 in range(kmin, kend, CH):
        c1 = min(kend, c0 + CH)
        vh = Bm @ (H[:, c0:c1].to(torch.complex128) * sw[c0:c1])           # (r, chunk)
        Vc[:r, c0:c1] = (vh.conj() * isw[c0:c1]).to(torch.complex64)
    A = ((Wr * s_r[None, :]).conj() * nrm[:, None]).to(torch.complex64).contiguous()
    return r, rp, A, Vc, float(worst[r - 1].item())


class BankPlan:
    """Bank-side products, made once per run and resident in HBM: tiles sorted by chirp time, dense templates,
    SVD bases. sw: the SVD metric sqrt(mean IP) (fp_filter.py:363); it sets how well the basis compresses, never the
    SNR definition (conj(h sw) (D / sw) = conj(h) D)."""

    def __init__(self, m1, m2, e0, flen, df, opts: Options, sw, isw, name='bank', sg_locations=None):
        self.name, self.opts = name, opts
        self.flen, self.df = flen, df
        self.kmin = int(opts.low_frequency_cutoff / df)
        self.m1, self.m2 = np.asarray(m1, float), np.asarray(m2, float)
        self.e0 = np.zeros(len(self.m1)) if e0 is None else np.asarray(e0, float)
        mt = self.m1 + self.m2
        self.mchirp = mt * (self.m1 * self.m2 / mt ** 2) ** 0.6
        # --sgchisq-locations: per template its sine-Gaussian tile list [(q, offset Hz)] or None
        self.sg_locs = sgchisq.parse_locations(sg_locations, self.m1, self.m2) if sg_locations else None
        fl = opts.low_frequency_cutoff
        tau0 = 5.0 / (256.0 * np.pi * fl * (self.m1 * self.m2 / (self.m1 + self.m2) ** 2)) * \
            (np.pi * (self.m1 + self.m2) * 4.925490947641267e-06 * fl) ** (-5.0 / 3.0)
        self.tiles = []
        self.T = dict(templates_s=0.0, svd_s=0.0)
        assert opts.templates in ('closed', 'interp') and opts.exact_stage in ('dense', 'regen')
        assert opts.exact_stage == 'dense' or opts.templates == 'interp', 'the regen exact stage needs interp templates'
        self.grid = None
        if opts.templates == 'interp':
            # one node grid for the bank, over [flow, the bank's highest ISCO frequency] (fp_filter.py:518-524)
            fisco = 1.0 / (6.0 ** 1.5 * np.pi * mt * 4.925490947641267e-06)
            kend_b = min(int(fisco.max() / df), flen - 1)
            self.grid = tmpl.Grid(tmpl.coarse_grid(fl, kend_b * df, opts.n_chirp, opts.n_log), df, self.kmin, kend_b)
        for ids in tmpl.tiles(tau0, opts.tile):
            t0 = time.perf_counter()
            coarse = None
            if self.grid is not None:
                amp, coef = tmpl.nodes(self.m1[ids], self.m2[ids], self.e0[ids], self.grid)
                H = tmpl.interp_templates(amp, coef, self.grid, flen)
                kmax = tmpl.interp_end(self.m1[ids], self.m2[ids], self.grid)
                if opts.exact_stage == 'regen':
                    from exact_stage import Coarse
                    coarse = Coarse(self.grid, amp, coef)
            else:
                H, kmax = tmpl.generate(self.m1[ids], self.m2[ids], self.e0[ids], flen, df, fl)
            torch.cuda.synchronize(); t1 = time.perf_counter()
            kend = int(kmax.max())
            r, rp, A, Vc, worst = svd_tile(H, sw, isw, self.kmin, kend, opts.eps)
            torch.cuda.synchronize(); t2 = time.perf_counter()
            if 2 * rp > _lib.load().gwgs_gemm_max_k(PRECISIONS[opts.precision]):
                raise AssertionError(f'tile rank {r} exceeds the GEMM kernel limit; use a smaller --tile')
            nb = np.full(len(ids), opts.chisq_bins, np.int64)
            stage = ExactStage(TemplateTile(H, nb, self.kmin, kmax, coarse), self_check=opts.self_check)
            self.tiles.append(TileBasis(ids, H, kmax, r, rp, A, Vc, worst, stage, kend))
            self.T['templates_s'] += t1 - t0; self.T['svd_s'] += t2 - t1

    @property
    def ranks(self):
        return [t.r for t in self.tiles]


@dataclass
class Gate:
    """--gate / --gate-mc-ratio / --gate-window (fp_filter.py:146-149, 424-445, 724-756): base-search triggers
    (detector name per trigger, end time in GPS seconds, chirp mass). A tile runs on a segment only if a trigger of
    that segment's detector lies within `window` seconds of the segment's analysed range with chirp mass inside the
    tile's range widened by `mc_ratio`, and only peaks within `window` of such a trigger are kept."""
    ifo: np.ndarray
    time: np.ndarray
    mchirp: np.ndarray
    mc_ratio: float = 1.25
    window: float = 1.0

    def per_segment(self, m: Dataset):
        """fp_filter.py:430-438: per segment of m, the samples (relative to the segment start, sorted) and chirp
        masses of this detector's triggers within the window of its analysed range."""
        sel = np.asarray(self.ifo).astype(str) == m.tag.split('-')[0]
        t, mc = np.asarray(self.time, np.float64)[sel], np.asarray(self.mchirp, np.float64)[sel]
        o = np.argsort(t, kind='stable')
        gs = np.round((t[o] - m.gps_start) * m.sr).astype(np.int64)
        mc = mc[o]
        w = int(self.window * m.sr)
        out = []
        for s, (a0, a1) in enumerate(m.analyze):
            c = gs - m.seg_start[s]
            i0, i1 = np.searchsorted(c, a0 - w), np.searchsorted(c, a1 + w, side='right')
            out.append((c[i0:i1], mc[i0:i1]))
        return out


def near(samp, gs, w):
    """fp_filter.py:441-443: samples within w of a sorted sample list gs."""
    j = np.searchsorted(gs, samp)
    return ((np.abs(samp - gs[np.clip(j - 1, 0, len(gs) - 1)]) <= w) |
            (np.abs(samp - gs[np.clip(j, 0, len(gs) - 1)]) <= w))


class Triggers:
    """Peaks of one dataset and bank, accumulated over tiles and segments (fp_filter.py:323, 772-853); sgc / sgd:
    the sine-Gaussian chi-squared (1 where not computed) and where it was computed."""
    KEYS = ('tid', 't', 'snr', 'ph', 'chi', 'dof', 'sg', 'ex', 'samp', 'seg', 'sgc', 'sgd')

    def __init__(self, sg=False):
        self.parts = {k: [] for k in self.KEYS}
        self.sg = sg                       # the peak file carries sg_chisq and sg_chisq_computed
        self.unready = []                  # sgc arrays of parts whose sine-Gaussian values are still on the device
        self.settled = 0                   # 1 when the first part is a consolidation's result
        self.sigmasq = None                # (bank templates, first segments) float64: each template's sigma^2

    def ready_parts(self):
        """Number of leading parts whose fields are final (consolidation works on these)."""
        if not self.unready:
            return len(self.parts['snr'])
        ids = {id(a) for a in self.unready}
        return next(i for i, a in enumerate(self.parts['sgc']) if id(a) in ids)

    def add(self, **kw):
        n = len(kw['snr'])
        kw.setdefault('sgc', np.ones(n, np.float32))
        kw.setdefault('sgd', np.zeros(n, bool))
        for k in self.KEYS:
            self.parts[k].append(np.asarray(kw[k]))

    def cat(self, k, dt):
        assert not self.unready, 'sine-Gaussian values still pending'
        return np.concatenate(self.parts[k]).astype(dt) if self.parts[k] else np.zeros(0, dt)

    def to_dict(self, gates=(), chisq=True):
        """The peak-file fields (fp_filter.py:910-937): gate veto |t - g_t| > w + taper + 0.5, sorted by
        (template_id, end_time); newsnr as PyCBC's (fp_filter.py:930)."""
        tid, t, snr = self.cat('tid', np.int64), self.cat('t', np.float64), self.cat('snr', np.float32)
        keep = np.ones(len(t), bool)
        for g_t, g_w, g_taper in gates:
            keep &= np.abs(t - g_t) > g_w + g_taper + 0.5
        sel = np.nonzero(keep)[0][_order_by_template_time(tid[keep], t[keep])]
        out = dict(template_id=tid[sel], snr=snr[sel], end_time=t[sel],
                   coa_phase=self.cat('ph', np.float32)[sel], peak_sigmasq=self.cat('sg', np.float64)[sel],
                   exact=self.cat('ex', bool)[sel], sample=self.cat('samp', np.int64)[sel],
                   segment=self.cat('seg', np.int64)[sel])
        if chisq:
            chi, dof = self.cat('chi', np.float32)[sel], self.cat('dof', np.int32)[sel]
            out.update(chisq=chi, chisq_dof=dof, newsnr=consolidate.newsnr(out['snr'], chi / np.maximum(dof, 1)))
        if self.sg:
            out.update(sg_chisq=self.cat('sgc', np.float32)[sel], sg_chisq_computed=self.cat('sgd', bool)[sel])
        return out


class _SharedValues:
    """One exact-stage launch over several segments: result() waits once and keeps the PeakValues."""

    def __init__(self, pv):
        self.pv, self.v = pv, None

    def result(self):
        if self.v is None:
            self.v = self.pv.result()
        return self.v


class _PartValues:
    """The peaks [off, off + n) of a shared exact-stage result (one segment's). alt: (candidate i in [0, n), request
    j) pairs, the same candidate's requests at its window's alternates; each candidate keeps, of its own request and
    its alternates', the one with the largest exact |z| (the smaller sample on equal |z|, so the choice does not
    depend on which peak the GEMM's rounding made the window maximum)."""

    def __init__(self, shared, off, n, alt=None):
        self.shared, self.off, self.n, self.alt = shared, off, n, alt

    def result(self):
        from exact_stage import PeakValues
        v = self.shared.result()
        j = np.arange(self.off, self.off + self.n)
        if self.alt is not None and len(self.alt[0]):
            ci, aj = self.alt
            cand, req = np.r_[np.arange(self.n), ci], np.r_[j, aj]
            o = np.lexsort((v.samples[req], -np.abs(v.z[req].astype(np.complex128)), cand))
            first = np.r_[True, cand[o][1:] != cand[o][:-1]]
            j = req[o][first]
        return PeakValues(v.z[j], v.pw[j], v.fpeak[j], v.samples[j])


def candidate_cut(thr, margin):
    """The candidate cut on |z|^2 of fp_filter.py:747, vw > (thr (1 - margin))^2: the square in float64, then
    float32, as torch compares a float32 tensor with a Python float."""
    return np.float32((thr * (1 - margin)) ** 2)


def _order_by_template_time(tid, t):
    """The permutation np.lexsort((t, tid)) returns (stable), found as one float64 argsort (introsort) of the key
    tid M + (t - min t), M = 2 span + 2, about ten times as fast.  Rounding keeps the key monotone in (tid, t), and a
    trigger list holds each (tid, t) once, so the sorted order is the lexsort's unless two distinct (tid, t) round to
    one key or a pair repeats; the result is checked in one pass (keys sorted, equal pairs in index order) and the
    lexsort is taken when it fails."""
    if len(t) < 2:
        return np.arange(len(t), dtype=np.int64)
    t0, span = float(t.min()), float(t.max() - t.min())
    m = 2.0 * span + 2.0
    if tid.min() >= 0 and float(tid.max()) * m + m < 2.0 ** 40:
        o = np.argsort(tid * m + (t - t0))
        a, b = tid[o], t[o]
        da, db = np.diff(a), np.diff(b)
        if np.all((da > 0) | ((da == 0) & ((db > 0) | ((db == 0) & (np.diff(o) > 0))))):
            return o
    return np.lexsort((t, tid))


def _fp8_scales(pw):
    """Per-basis fp8 scales from (segments, rp) powers sum_k |q_jk|^2: Z is multiplied by 1 / rms and A by rms, with
    rms = sqrt of the mean over segments (0 rows stay 0)."""
    acc = pw[0].clone()
    for i in range(1, pw.shape[0]):          # elementwise (no reduction kernel: the stage may be graph-captured)
        acc += pw[i]
    rms = torch.sqrt(acc / pw.shape[0]).to(torch.float32)
    zcs = torch.where(rms > 0, 1.0 / rms.clamp(min=1e-30), torch.zeros_like(rms)).contiguous()
    return zcs, rms.contiguous()


def run_jobs(jobs):
    """Filter runs one after another in one software pipeline: jobs is an iterable of (Filter, plans), which may be a
    generator, so that a job is set up (its data conditioned, its Filter made) when the pipeline reaches it, while
    the previous job's last batches are still in flight. Batches are launched `depth` ahead of the one being decoded
    (the depth of the job being launched), decoded by their own Filter and finished one behind. Returns, per job, per
    plan, one Triggers per dataset."""
    done = []
    launched, fin = deque(), deque()
    for f, plans in jobs:
        outs = f._begin(plans)
        done.append((f, plans, outs))
        for cur in f._batches(plans, outs):
            if f.opts.timing:
                f.decode_batch(cur); f.finish_batch(cur)
                continue
            launched.append(cur)
            while len(launched) > f.opts.depth:
                st = launched.popleft()
                fin.append(st['filter'].decode_batch(st))
            while len(fin) > 1:
                st = fin.popleft()
                st['filter'].finish_batch(st)
    while launched:
        st = launched.popleft()
        fin.append(st['filter'].decode_batch(st))
    while fin:
        st = fin.popleft()
        st['filter'].finish_batch(st)
    for f, plans, outs in done:
        f._end(plans, outs)
    return [outs for _, _, outs in done]


def svd_metric(datasets):
    """fp_filter.py:360-364: IPall (all segments, flen), sw = sqrt(mean IPall) in float64, isw = 1/sw (0 where 0)."""
    IPall = torch.cat([m.IP for m in datasets], 0)
    sw = torch.sqrt(IPall.double().mean(0))
    isw = torch.where(sw > 0, 1.0 / sw.clamp(min=1e-30), torch.zeros_like(sw))
    return IPall, sw, isw


class Filter:
    """One bank against a batch of datasets (all with the same flen, tlen, sr)."""

    def __init__(self, datasets, opts: Options):
        _dev.require_gpu()
        self.opts = opts
        self.data = datasets
        m0 = datasets[0]
        assert all(m.flen == m0.flen and m.tlen == m0.tlen and m.sr == m0.sr for m in datasets)
        self.flen, self.tlen, self.df, self.sr = m0.flen, m0.tlen, m0.df, m0.sr
        self.kmin = int(opts.low_frequency_cutoff / self.df)
        self.cw = int(opts.cluster_window * self.sr)
        assert self.cw % BN == 0, 'cluster window must be a multiple of %d samples (GEMM output block)' % BN
        self.IPall = torch.cat([m.IP for m in datasets], 0)
        off = 0
        self.seg0 = []
        for m in datasets:
            self.seg0.append(off); off += len(m.analyze)
        self._q = None
        self._work = None
        self._bufs = None                  # graph mode: depth + 1 rotating buffer sets
        self._graphs = {}                  # graph mode: captured tile norms and stage-A batches
        self.T = dict(norm_s=0.0, basis_fft_s=0.0, gemm_s=0.0, rebuilt_z_s=0.0, refine_chisq_s=0.0, n_candidates=0,
                      n_exact_candidates=0, n_rebuilt_peaks=0, n_chisq_points=0, consolidate_s=0.0,
                      n_consolidated_out=0, sg_chisq_s=0.0, n_sg_chisq_points=0)
        self.rel = []
        self.code = PRECISIONS[opts.precision]
        assert not opts.pre_consolidate or (opts.consolidate is not None and not opts.graphs), \
            'pre_consolidate needs consolidate, and the candidate tables change between runs (no graphs)'
        # fp_filter.py:160: the gate drops peaks after the candidates are ranked, so the pre-consolidation's ranking
        # would not be the consolidation's; gated batches differ from run to run (no graphs)
        assert opts.gate is None or not (opts.pre_consolidate or opts.graphs), 'gate excludes pre_consolidate, graphs'
        self._prekeep = (consolidate.PreKeep(opts.consolidate, opts.snr_threshold, opts.exact_snr, opts.margin,
                                             self.sr, self.cw) if opts.pre_consolidate else None)
        self._floors = {}                  # (plan index, dataset index): the floor consolidation left (device)
        self._sg_pending = []              # (Triggers, sgc, sgd, [(indices, sgchisq.PendingSG)])
        self._cons_pool = None             # one worker thread for the consolidations (fp_filter.py:623, 888)
        self._cons_jobs = []               # (Triggers, plan index, dataset index, parts covered, future)
        self._ysc = None                   # bf16 / fp8 forms: the transform's fp16 intermediate scale per segment
        self._hot = None                   # running fraction of a batch's windows holding exact-band candidates
        self._y16_update()

    def _nsets(self):
        """Rotation sets of graph mode: a batch's buffers are reused depth + 1 batches later (_alloc_rotation)."""
        return max(1, self.opts.depth) + 1

    def _reset_run_state(self):
        """State one run hands to the next must not leak into it: the floors consolidation left (they would drop
        the next run's candidates at or below another run's floors), pending consolidation jobs and sine-Gaussian
        launches (both are drained by the end of every run)."""
        assert not self._cons_jobs and not self._sg_pending, 'a previous run did not finish'
        self._floors = {}

    def _overlap(self):
        return self.opts.overlap and not self.opts.graphs

    def _sync(self):
        if self.opts.timing:
            torch.cuda.synchronize()
        return time.perf_counter()

    def _y16_update(self):
        """Scale of the basis transform's fp16 intermediate per segment, one value for every basis row, in rows of
        self._ysc (segments of all datasets in order; opts.tile + 16 >= rp columns): the basis rows are unit vectors
        vh_j over the band (q_jk = conj(vh_jk) D_k / sw_k), so their power sum_k |q_jk|^2 is a weighted mean of
        |D_k / sw_k|^2 and the intermediate's rms is about sqrt(c / n1) with c the segment's mean of |D_k / sw_k|^2
        over [kmin, flen) (sw = sqrt(mean IP), as the SVD metric). The scale is 1 / (64 rms): fp16 then holds values
        from 4e-3 to 4e6 times the rms as normal numbers, far beyond what one row's spread or a loud segment needs,
        and its rounding stays relative (2^-11). Computed when data is bound, outside any captured graph."""
        if self.code == 0:
            return
        isw2 = torch.where(self.IPall.double().mean(0) > 0, 1.0 / self.IPall.double().mean(0).clamp(min=1e-300),
                           torch.zeros((), dtype=torch.float64, device=self.IPall.device))[self.kmin:]
        n1 = 1 << ((int(self.tlen).bit_length() - 1) // 2)
        Dall = self.data[0].D if len(self.data) == 1 else torch.cat([m.D for m in self.data], 0)
        c = ((Dall[:, self.kmin:].abs().double() ** 2) * isw2).mean(1)
        v = torch.where(c > 0, torch.sqrt(n1 / c.clamp(min=1e-300)) / 64.0, torch.ones_like(c))
        v = v.clamp(1e-30, 1e30).to(torch.float32)
        if self._ysc is None:
            self._ysc = torch.empty((len(v), self.opts.tile + 16), dtype=torch.float32, device='cuda')
        self._ysc.copy_(v[:, None].expand_as(self._ysc))

    def _qbuf(self, nb, rp):
        """(nb, rp, tlen) complex64 transform input whose columns >= flen stay zero (fp_filter.py:595): allocated
        zeroed once and only its [:flen] columns are ever written."""
        if self._q is None or self._q.shape[0] < nb or self._q.shape[1] < rp:
            self._q = None
            self._q = torch.zeros((nb, rp, self.tlen), dtype=torch.complex64, device='cuda')
        return self._q[:nb, :rp]

    def _ones(self, n):
        if getattr(self, '_ones_buf', None) is None or self._ones_buf.numel() < n:
            self._ones_buf = torch.ones(max(n, 1024), dtype=torch.float32, device='cuda')
        return self._ones_buf[:n]

    def _workbuf(self, nbytes):
        """Workspace of the fused basis transform (the column-pass output), grown on demand and reused."""
        if self._work is None or self._work.numel() < nbytes:
            self._work = None
            self._work = torch.empty(nbytes, dtype=torch.uint8, device='cuda')
        return self._work

    def basis_operand(self, tb: TileBasis, segs, Tns, row0, Zt=None):
        """Zt (sum Tn, 2 rp) in the GEMM's operand type: the basis series of every segment of the batch over its
        analysed span, stacked by rows (fp_filter.py:597-613). Power-of-two segment lengths 2^13..2^22 take the
        fused four-step transform (csrc/basis_fft.hip): no transform input is materialised and the output is written
        time-major directly; other lengths take one batched hipFFT transform and the LDS transpose (pack_zt).
        For fp8, the per-basis scale 1 / rms (rms from the batch's mean sum_k |q_jk|^2, Parseval) is applied to Z
        and returned for A."""
        rp, K = tb.rp, 2 * tb.rp
        nb = len(segs)
        dt = OPERAND_DTYPES[self.code]
        lib = _lib.load()
        ws = lib.gwgs_basis_series_workspace(tb.r, self.tlen)
        zcs = acs = None
        if Zt is None:
            Zt = torch.empty((int(row0[-1]), K), dtype=dt, device='cuda')
        if ws > 0:
            # the transform's intermediate: complex64 for the fp32 form; two fp16 scaled to unit rms per basis row
            # for the bf16 / fp8 forms (half the bytes; its 11-bit rounding is well below theirs)
            ytype = 0 if self.code == 0 else 1
            if self.code in (2, 3):
                pw = torch.zeros((nb, rp), dtype=torch.float64, device='cuda')
                for i, (di, s, gs, a0, nw) in enumerate(segs):
                    _lib.call('gwgs_vd_power', tb.Vc.data_ptr(), self.flen, tb.r, self.data[di].D[s].data_ptr(),
                              self.kmin, tb.kend, pw[i].data_ptr(), _dev.stream())
                zcs, acs = _fp8_scales(pw)
            assert not ytype or rp <= self._ysc.shape[1], 'tile larger than this Filter\'s opts.tile'
            esz = Zt.element_size()
            # every segment of the batch in one launch of each pass, up to SERIES_SEG at a time (each with its own
            # intermediate in the workspace)
            wseg = tb.r * self.tlen * (4 if ytype else 8)       # one segment's intermediate
            for c0 in range(0, nb, SERIES_SEG):
                ch = range(c0, min(nb, c0 + SERIES_SEG))
                work = self._workbuf(wseg * len(ch))
                dptr = (ctypes.c_void_p * len(ch))(*[self.data[segs[i][0]].D[segs[i][1]].data_ptr() for i in ch])
                yptr = (ctypes.c_void_p * len(ch))(*[self._ysc[self.seg0[segs[i][0]] + segs[i][1]].data_ptr()
                                                     if ytype else None for i in ch])
                a0s = (ctypes.c_long * len(ch))(*[segs[i][3] for i in ch])
                tns = (ctypes.c_long * len(ch))(*[Tns[i] for i in ch])
                zrs = (ctypes.c_long * len(ch))(*[int(row0[i]) for i in ch])
                e0 = _dev.METER.start()
                _lib.call('gwgs_basis_series_batch', self.code, tb.Vc.data_ptr(), self.flen, tb.r, rp, len(ch), dptr,
                          self.kmin, tb.kend, self.tlen, a0s, tns, zrs, _dev.ptr(zcs), ytype, yptr, work.data_ptr(),
                          Zt.data_ptr(), K, _dev.stream())
                # HBM bytes: Vc rows (once per launch) and D over [kmin, kend), Y written and read, Zt written
                _dev.METER.stop('basis_series', e0, 8 * tb.r * (tb.kend - self.kmin)
                                + sum(8 * (tb.kend - self.kmin) + (8 if ytype else 16) * tb.r * self.tlen
                                      + Tns[i] * K * esz for i in ch))
            return Zt, acs
        q = self._qbuf(nb, rp)
        for i, (di, s, gs, a0, nw) in enumerate(segs):
            _lib.call('gwgs_form_q', tb.Vc.data_ptr(), rp, self.flen, self.data[di].D[s].data_ptr(), self.tlen,
                      q[i].data_ptr(), _dev.stream())
        Zc = torch.fft.ifft(q, dim=2, norm='forward')                 # (nb, rp, tlen): basis series
        if self.code in (2, 3):
            pw = torch.empty((nb, rp), dtype=torch.float64, device='cuda')
            for i in range(nb):
                _lib.call('gwgs_basis_power', q[i].data_ptr(), self.tlen, rp, 0, self.flen, pw[i].data_ptr(),
                          _dev.stream())
            zcs, acs = _fp8_scales(pw)
        for i, (di, s, gs, a0, nw) in enumerate(segs):
            _lib.call('gwgs_pack_zt', self.code, Zc[i].data_ptr(), self.tlen, rp, a0, Tns[i], Tns[i], _dev.ptr(zcs),
                      Zt[row0[i]].data_ptr(), _dev.stream())
        del Zc
        return Zt, acs

    def run(self, plan: BankPlan):
        """Filter every tile of the plan against every segment of every dataset. Returns one Triggers per dataset."""
        return self.run_banks([plan])[0]

    def run_banks(self, plans):
        """Every tile of every plan (bank) against every segment; returns, per plan, one Triggers per dataset.

        Software pipeline on one stream (run_jobs): the transforms, GEMM and candidate selection of the next `depth`
        batches are enqueued before a batch's candidates are read back, and a batch's exact stage is enqueued before
        the previous batch's results are collected, so the GPU has queued work while the host decodes and finalises,
        across tiles and banks. Copies to the host run on a side stream after the events they depend on. With
        opts.timing every stage is synchronised instead."""
        return run_jobs([(self, plans)])[0]

    def _begin(self, plans):
        """Per-run set-up: the state one run must not hand to the next, the output lists, the segment list."""
        self._reset_run_state()
        if self.IPall is None:                 # released at the end of a previous run's launches
            self.IPall = torch.cat([m.IP for m in self.data], 0)
        if self._overlap():
            # the transform stream sees everything enqueued before the run (conditioning, the fp16 scales)
            _dev.basis_stream().wait_stream(torch.cuda.current_stream())
        sg_on = self.opts.sgchisq_snr_threshold is not None
        return [[Triggers(sg=sg_on) for _ in self.data] for _ in plans]

    def _batches(self, plans, outs):
        """Launch every batch of the run in order (tile norms at each tile's start), yielding each batch's state."""
        segs = []
        for di, m in enumerate(self.data):
            for s, (a0, a1) in enumerate(m.analyze):
                nw = (a1 - a0) // self.cw
                if nw > 0:
                    segs.append((di, s, self.seg0[di] + s, a0, nw))
        gate = self.opts.gate
        gsegs = [gate.per_segment(m) for m in self.data] if gate is not None else None
        if self.opts.graphs and self._bufs is None:
            self._alloc_rotation(plans, segs)
        ordinal = 0
        for pi, plan in enumerate(plans):
            for ti, tb in enumerate(plan.tiles):
                tsegs, gsel = segs, None
                if gate is not None:
                    # fp_filter.py:726-739: the segments where a trigger's chirp mass fits this tile
                    mc = plan.mchirp[tb.ids]
                    lo, hi = mc.min() / gate.mc_ratio, mc.max() * gate.mc_ratio
                    gsel = {}
                    for sg_ in segs:
                        gsm, gmc = gsegs[sg_[0]][sg_[1]]
                        g_ = gsm[(gmc >= lo) & (gmc <= hi)]
                        if len(g_):
                            gsel[(sg_[0], sg_[1])] = g_
                    tsegs = [sg_ for sg_ in segs if (sg_[0], sg_[1]) in gsel]
                    self.T['gate_runs'] = self.T.get('gate_runs', 0) + len(tsegs)
                    self.T['gate_skips'] = self.T.get('gate_skips', 0) + len(segs) - len(tsegs)
                    if not tsegs:
                        continue
                t0 = self._sync()
                sig, norm = self.tile_norms(tb, (id(plan), ti))
                tile_host = _dev.Fetch(sig, norm)
                t1 = self._sync(); self.T['norm_s'] += t1 - t0
                for b0, b1 in self._split(tsegs):
                    cur = self.launch_batch(tb, tsegs[b0:b1], norm, tile_host, key=(id(plan), ti, b0),
                                            ordinal=ordinal, pi=pi)
                    ordinal += 1
                    cur['outs'], cur['pi'], cur['last'], cur['sg_locs'] = outs[pi], pi, b1 >= len(tsegs), plan.sg_locs
                    cur['gate'] = gsel
                    cur['filter'] = self
                    cur['ntmp'] = len(plan.m1)
                    yield cur
        if not self.opts.graphs:
            # every transform and norm of the run is enqueued: the workspaces go back to the allocator (reused in
            # stream order by the next job), so a pipeline of many jobs holds one set at a time
            self._work = self._q = None
            self.IPall = None

    def _end(self, plans, outs):
        """After the run's last batch: the pending sine-Gaussian values and the final consolidations."""
        self._resolve_sg()
        if self.opts.consolidate is not None:
            # the parts added after each dataset's last consolidation (sine-Gaussian values pending then, or found
            # while it ran): one more pass, which leaves an already consolidated list as it is
            self._apply_consolidated(block=True)
            if any(len(tr.parts['snr']) > tr.settled for o_ in outs for tr in o_):
                for pi in range(len(plans)):
                    self._consolidate(pi, outs[pi])
                self._apply_consolidated(block=True)
        if self._cons_pool is not None:
            self._cons_pool.shutdown(wait=True)
            self._cons_pool = None

    def finish_batch(self, st):
        """The sine-Gaussian values of earlier batches (their kernels ran right after this batch's exact stage, which
        collect_batch waits for anyway), the consolidations that have finished, collect_batch, and after a tile's last
        batch the consolidation of the plan's datasets (fp_filter.py:886-888) over their parts with final fields, on
        a worker thread as the reference's --pre-consolidate does: its floors feed the pre-consolidation of batches
        launched after it is applied. Parts still waiting for sine-Gaussian values, or found while a consolidation
        runs, are consolidated later; applying the consolidation to any prefix of the peak list leaves the final list
        unchanged."""
        self._resolve_sg()
        self._apply_consolidated(block=False)
        self.collect_batch(st)
        if st.get('last'):
            self._record_sigmasq(st)
            self._consolidate(st['pi'], st['outs'])

    def _record_sigmasq(self, st):
        """The tile's sigma^2 of every template in each dataset's first segments into the dataset's Triggers (the
        peak file's 2-D sigmasq, fp_filter.py:878)."""
        sig_h, _ = st['tile_host'].get()
        tb = st['tb']
        ntmp = int(st['ntmp'])
        for di, (m, tr) in enumerate(zip(self.data, st['outs'])):
            first = m.first if m.first is not None else list(range(len(m.analyze)))
            if tr.sigmasq is None:
                tr.sigmasq = np.zeros((ntmp, len(first)))
            tr.sigmasq[tb.ids] = sig_h[:, self.seg0[di] + np.asarray(first, np.int64)]

    def _consolidate(self, pi, outs):
        """Start the consolidation of every dataset's finished parts; a list whose previous consolidation is still
        running is left for the next one (the end of the run consolidates whatever remains)."""
        o = self.opts
        if o.consolidate is None:
            return
        self._apply_consolidated(block=False)
        if self._cons_pool is None:
            self._cons_pool = ThreadPoolExecutor(1)
        busy = {id(job[0]) for job in self._cons_jobs}
        for di, m in enumerate(self.data):
            if id(outs[di]) in busy:
                continue
            parts, n = consolidate.snapshot(outs[di])
            if n == 0:
                continue
            fut = self._cons_pool.submit(consolidate.consolidate_parts, parts, list(m.gates), m.gps_start, m.sr,
                                         o.consolidate, o.pre_consolidate)
            self._cons_jobs.append((outs[di], pi, di, n, fut))

    def _apply_consolidated(self, block):
        """Put finished consolidations in place (all of them, waiting, with block) and take their floors."""
        t0 = time.perf_counter()
        for job in [j for j in self._cons_jobs if block or j[4].done()]:
            self._cons_jobs.remove(job)
            tr, pi, di, n, fut = job
            kept, floor, dropped = fut.result()
            consolidate.replace_prefix(tr, n, kept)
            self.T['n_consolidated_out'] += dropped
            if floor is not None:
                self._floors[(pi, di)] = _dev.to_dev(floor)
        self.T['consolidate_s'] += time.perf_counter() - t0

    def _resolve_sg(self):
        """Wait for every enqueued sine-Gaussian launch and write its values into its parts."""
        if not self._sg_pending:
            return
        t0 = time.perf_counter()
        for tr, sgc, sgd, pend in self._sg_pending:
            for gi, pv in pend:
                v, ok = pv.result()
                sgc[gi[ok]] = v[ok]
                sgd[gi[ok]] = True
                self.T['n_sg_chisq_points'] += int(ok.sum())
            tr.unready = [a for a in tr.unready if a is not sgc]
        self._sg_pending = []
        self.T['sg_chisq_s'] += time.perf_counter() - t0

    def tile_norms(self, tb: TileBasis, key):
        """sigma^2 and the norm of every template of the tile in every segment (fp_filter.py:710-714); with graphs
        captured once per tile and replayed."""
        work = 8 * len(tb.ids) * (tb.kend - self.kmin) + 4 * self.IPall.shape[0] * (tb.kend - self.kmin)
        if not self.opts.graphs:
            e0 = _dev.METER.start()
            sig, norm = tmpl.sigmasq(tb.H, self.IPall, self.df, self.kmin, tb.kend)
            _dev.METER.stop('sigmasq', e0, work)
            return sig, norm
        rec = self._graphs.get(('tile',) + key)
        if rec is None:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=self._capture_stream()):
                sig, norm = tmpl.sigmasq(tb.H, self.IPall, self.df, self.kmin, tb.kend)
            rec = self._graphs[('tile',) + key] = dict(graph=g, sig=sig, norm=norm)
        rec['graph'].replay()
        return rec['sig'], rec['norm']

    def _capture_stream(self):
        if getattr(self, '_cap', None) is None:
            self._cap = torch.cuda.Stream()
        return self._cap

    def _batch_geometry(self, tb: TileBasis, segs):
        """Host-side layout of one batch: rows of each segment in Zt, output columns, the GEMM item table and the
        segments' indices into the norm table (device copies made here, outside any capture)."""
        Tns = [nw * self.cw for (_, _, _, _, nw) in segs]
        row0 = np.concatenate([[0], np.cumsum(Tns)]).astype(np.int64)
        nws = [nw for (_, _, _, _, nw) in segs]
        col0 = np.concatenate([[0], np.cumsum(nws)]).astype(np.int64)
        per = self.cw // BN
        items = fp_gemm.items_table([(int(row0[i]), Tns[i] // BN, int(col0[i]), i, Tns[i]) for i in range(len(segs))],
                                    fp_gemm.tiles_per_item(self.opts.gemm_cfg, per))
        gsl = _dev.to_dev(np.array([gs for (_, _, gs, _, _) in segs], np.int64))
        o = self.opts
        hseg = (fp_gemm.hot_segments(row0, col0, Tns) if o.refine > 0 and o.alternates > 0 and self._second_frac() > 0
                else None)
        return dict(Tns=Tns, row0=row0, nwsum=sum(nws), col0=col0, items=items, gsl=gsl, K=2 * tb.rp, per=per,
                    hseg=hseg)

    def _split(self, segs):
        """Consecutive batches [b0, b1) of the segment list: up to opts.segment_batch segments and opts.batch_samples
        analysed samples each (a longer segment alone)."""
        out, b0, tot = [], 0, 0
        for i, sg in enumerate(segs):
            t = sg[4] * self.cw
            if i > b0 and (i - b0 >= max(1, self.opts.segment_batch) or tot + t > self.opts.batch_samples):
                out.append((b0, i))
                b0, tot = i, 0
            tot += t
        if b0 < len(segs):
            out.append((b0, len(segs)))
        return out

    def _alloc_rotation(self, plans, segs):
        """depth + 1 buffer sets, each sized for the largest batch of the run, for the captured stage-A graphs:
        batch i of a run writes set i mod (depth + 1). Invariant: run_banks decodes batch i (its candidate list read
        back and awaited, its rebuilt phases enqueued on the compute stream) once batches i + 1 .. i + depth are
        launched and before batch i + depth + 1 is, so every reader of set i mod (depth + 1) is complete or ordered
        before the next batch that overwrites it."""
        zt = a = w = nmax = 0
        for plan in plans:
            for tb in plan.tiles:
                n, K = len(tb.ids), 2 * tb.rp
                for b0, b1 in self._split(segs):
                    sg = segs[b0:b1]
                    T = sum(nw * self.cw for (_, _, _, _, nw) in sg)
                    nw = sum(x[4] for x in sg)
                    zt, a, w, nmax = max(zt, T * K), max(a, n * K), max(w, n * nw), max(nmax, n)
        nwork = int(_lib.load().gwgs_select_workspace(w)) + 1
        dt = OPERAND_DTYPES[self.code]
        self._bufs = [dict(zt=torch.empty(zt, dtype=dt, device='cuda'), ar=torch.empty(a, dtype=dt, device='cuda'),
                           ai=torch.empty(a, dtype=dt, device='cuda'),
                           vw=torch.empty(w, dtype=torch.float32, device='cuda'),
                           iw=torch.empty(w, dtype=torch.int32, device='cuda'),
                           sel=(torch.empty(w, dtype=torch.int32, device='cuda'),
                                torch.empty(w, dtype=torch.int32, device='cuda'),
                                torch.empty(w, dtype=torch.float32, device='cuda'),
                                torch.empty(nwork, dtype=torch.int32, device='cuda')),
                           **(dict(vl=torch.empty(w * fp_gemm.LM, dtype=torch.float32, device='cuda'),
                                   il=torch.empty(w * fp_gemm.LM, dtype=torch.int32, device='cuda'),
                                   alt=torch.empty(w * max(self.opts.alternates, 1), dtype=torch.int32,
                                                   device='cuda'))
                              if self.opts.refine > 0 and self.opts.alternates > 0 else {}))
                      for _ in range(self._nsets())]
        ws = max((_lib.load().gwgs_basis_series_workspace(tb.r, self.tlen) for plan in plans for tb in plan.tiles),
                 default=0)
        if ws > 0:
            # the batched transform's intermediates for up to SERIES_SEG segments, so that no capture allocates
            self._workbuf(ws * min(max(b1 - b0 for b0, b1 in self._split(segs)), SERIES_SEG))
        self._ones(nmax)

    def _batch_device(self, tb: TileBasis, segs, norm, geo, bufs=None, pk=None):
        """All device work of stage A (no host reads, no host-to-device copies but the pre-consolidation's tables):
        basis series, template operands, row scales, the GEMM with the window epilogue, the pre-consolidation (pk:
        its per-segment tables, or None) and the candidate compaction. bufs: a rotation set to write into (graph
        capture) or None (fresh allocations)."""
        o = self.opts
        n, rp, K = len(tb.ids), tb.rp, geo['K']
        dt = OPERAND_DTYPES[self.code]
        T, nwsum = int(geo['row0'][-1]), geo['nwsum']
        LM = fp_gemm.LM
        lists = o.refine > 0 and o.alternates > 0
        assert 0 <= o.alternates < LM
        # the lists in a second pass over the candidates' windows (below) while those are sparse
        second = (lists and bufs is None and geo['hseg'] is not None and
                  (self._hot is None or self._hot < self._second_frac()))
        lists = lists and not second
        if bufs is None:
            Zt = Ar = Ai = vw = iw = sel = None
            if lists:
                vl = torch.empty((n, nwsum * LM), dtype=torch.float32, device='cuda')
                il = torch.empty((n, nwsum * LM), dtype=torch.int32, device='cuda')
        else:
            Zt = bufs['zt'][:T * K].view(T, K)
            Ar, Ai = bufs['ar'][:n * K].view(n, K), bufs['ai'][:n * K].view(n, K)
            vw, iw = bufs['vw'][:n * nwsum].view(n, nwsum), bufs['iw'][:n * nwsum].view(n, nwsum)
            sel = bufs['sel']
            if lists:
                vl = bufs['vl'][:n * nwsum * LM].view(n, nwsum * LM)
                il = bufs['il'][:n * nwsum * LM].view(n, nwsum * LM)
        t0 = self._sync()
        if self._overlap() and bufs is None:
            # on the transform stream; the GEMM below waits for this batch's transforms only, so the next batch's
            # transforms can run beside this batch's GEMM. Zt (and the fp8 scales) come from that stream's pool and
            # are marked as read on the compute stream.
            main = torch.cuda.current_stream()
            with torch.cuda.stream(_dev.basis_stream()):
                Zt, acs = self.basis_operand(tb, segs, geo['Tns'], geo['row0'], Zt)
                done = _dev.mark()
            main.wait_event(done)
            Zt.record_stream(main)
            if acs is not None:
                acs.record_stream(main)
        else:
            Zt, acs = self.basis_operand(tb, segs, geo['Tns'], geo['row0'], Zt)
        t1 = self._sync(); self.T['basis_fft_s'] += t1 - t0
        if Ar is None:
            Ar = torch.empty((n, K), dtype=dt, device='cuda'); Ai = torch.empty_like(Ar)
            vw = torch.empty((n, nwsum), dtype=torch.float32, device='cuda')
            iw = torch.empty((n, nwsum), dtype=torch.int32, device='cuda')
        ones = self._ones(n)
        rows2 = (norm[:, geo['gsl']] ** 2).T.to(torch.float32)         # (nb, n): norm^2 per (segment, template)
        if self.code in (2, 3):
            # per-basis scales (unit-rms Z, rms on A) and a per-template scale to the operand range, whose inverse
            # square the GEMM applies to |z|^2
            pws = torch.empty(n, dtype=torch.float32, device='cuda')
            _lib.call('gwgs_stack_a_fp8' if self.code == 2 else 'gwgs_stack_a_f16', tb.A.data_ptr(), n, tb.r,
                      ones.data_ptr(), acs.data_ptr(), K, Ar.data_ptr(), Ai.data_ptr(), pws.data_ptr(), _dev.stream())
            rows2 = rows2 * pws[None, :]
        else:
            _lib.call('gwgs_stack_a', self.code, tb.A.data_ptr(), n, tb.r, ones.data_ptr(), None, n, K,
                      Ar.data_ptr(), Ai.data_ptr(), _dev.stream())
        rows2 = rows2.contiguous()
        e0 = _dev.METER.start()
        ratio = float(np.float32((1 - o.alt_margin) ** 2))
        cut = candidate_cut(o.snr_threshold, o.margin)
        if lists:
            # with each window's largest local maxima that can be alternates: within the margin of the window's
            # maximum, in windows that reach the candidate cut (the gate a hair lower than the alternates' ratio and
            # the floor than the cut, so that rounding cannot drop one)
            fp_gemm.gemm_peaks_multi_lm(self.code, Zt, K, Ar, Ai, K, n, K, geo['items'], True, geo['per'],
                                        o.alt_sep, o.gemm_cfg, vw, iw, vl, il, rows2, gate=ratio * (1 - 1e-5),
                                        floor=float(cut) * (1 - 1e-5))
        else:
            fp_gemm.gemm_peaks_multi(self.code, Zt, K, Ar, Ai, K, n, K, geo['items'], 1 | fp_gemm.STACKED,
                                     geo['per'], o.gemm_cfg, vw, iw, rows2)
        # algorithmic flops: 4 n K_real Tn with K_real = 2 r (padding not counted)
        _dev.METER.stop('gemm_' + o.precision, e0, 4.0 * n * (2 * tb.r) * float(sum(geo['Tns'])))
        # the operand Zt read once: the launch's algorithmic HBM bytes
        _dev.METER.note('gemm_' + o.precision + '_zt_bytes', float(sum(geo['Tns'])) * K * Zt.element_size())
        if pk is not None:
            e0 = _dev.METER.start()
            self._prekeep.launch(vw, iw, cut, pk)
            # algorithmic bytes: one read of the window maxima (the candidates' samples and drops are a few percent)
            _dev.METER.stop('pre_keep', e0, 4.0 * vw.numel())
        idx, samp, val, cnt = peak_kernel.select_launch(vw, iw, float(cut), sel)
        alt = None
        if lists or second:
            na, w = o.alternates, idx.numel()
            alt = (bufs['alt'][:w * na] if bufs is not None else
                   torch.empty(w * na, dtype=torch.int32, device='cuda')).view(w, na)
            xcut = float(np.float32((o.exact_snr * (1 - o.margin)) ** 2))
            if second:
                # the local-maximum lists of only the windows that hold exact-band candidates: the same GEMM kernel
                # in list mode over one-window items built on the device from the candidate list (the same tiles
                # and arithmetic as a first-pass list, so the same lists); the window maxima it writes again go to
                # scratch
                e0 = _dev.METER.start()
                flags = torch.empty(nwsum, dtype=torch.int32, device='cuda')
                items2 = torch.empty(nwsum * fp_gemm.ITEM.itemsize, dtype=torch.uint8, device='cuda')
                nit = torch.empty(1, dtype=torch.int32, device='cuda')
                fp_gemm.hot_items(idx, val, cnt, nwsum, xcut, geo['hseg'], geo['per'], flags, items2, nit)
                vl = torch.empty((n, nwsum * LM), dtype=torch.float32, device='cuda')
                il = torch.empty((n, nwsum * LM), dtype=torch.int32, device='cuda')
                fp_gemm.gemm_peaks_multi_lm_dyn(self.code, Zt, K, Ar, Ai, K, n, K, items2, nit, nwsum, True,
                                                geo['per'], o.alt_sep, o.gemm_cfg, torch.empty_like(vw),
                                                torch.empty_like(iw), vl, il, rows2, gate=ratio * (1 - 1e-5),
                                                floor=float(cut) * (1 - 1e-5))
                _dev.METER.stop('second_pass', e0, 0.0)
            # every exact-band candidate's alternates: its window's other local maxima within the margin of its
            # window maximum, farther than the re-centring half width from it
            _lib.call('gwgs_window_alternates', vl.data_ptr(), il.data_ptr(), vl.stride(0), LM, idx.data_ptr(),
                      samp.data_ptr(), val.data_ptr(), cnt.data_ptr(), w, nwsum, self.cw, xcut, ratio,
                      max(int(o.refine), int(o.alt_sep)), na, alt.data_ptr(), _dev.stream())
        t2 = self._sync(); self.T['gemm_s'] += t2 - t1
        return dict(Zt=Zt, Ar=Ar, Ai=Ai, rows2=rows2, vw=vw, iw=iw, sel=(idx, samp, val, cnt), alt=alt, second=second)

    def launch_batch(self, tb: TileBasis, segs, norm, tile_host, key=None, ordinal=0, pi=0):
        """Stage A of one tile against a batch of segments, enqueued only: the basis series of every segment
        (basis_operand), one GEMM launch over all their windows (an item table of (segment, run of tiles)) with each
        segment's template norm^2 applied to |z|^2 as a per-(segment, template) row scale, and the candidate list
        (fp_filter.py:741-757) compacted on the device. With opts.graphs the stage is captured in a HIP graph on
        the first run (into rotation set ordinal mod 3) and replayed afterwards. Returns the batch state."""
        if self.opts.graphs:
            rec = self._graphs.get(key)
            if rec is None:
                geo = self._batch_geometry(tb, segs)
                g = torch.cuda.CUDAGraph()
                torch.cuda.synchronize()
                with torch.cuda.graph(g, stream=self._capture_stream()):
                    dev = self._batch_device(tb, segs, norm, geo, self._bufs[ordinal % self._nsets()])
                rec = self._graphs[key] = dict(graph=g, geo=geo, dev=dev)
            rec['graph'].replay()
            geo, dev = rec['geo'], rec['dev']
        else:
            geo = self._batch_geometry(tb, segs)
            pk = None
            if self._prekeep is not None:
                pk = [(self.data[di].seg_start[s] + a0, self.data[di].gps_start, int(geo['col0'][i]), nw,
                       self.data[di].gates, self._floors.get((pi, di)))
                      for i, (di, s, gs, a0, nw) in enumerate(segs)]
            dev = self._batch_device(tb, segs, norm, geo, pk=pk)
        idx, samp, val, cnt = dev['sel']
        sel = peak_kernel.Selection(idx, samp, val, cnt, _dev.mark(), extra=dev.get('alt'))
        return dict(tb=tb, segs=segs, Zt=dev['Zt'], Ar=dev['Ar'], Ai=dev['Ai'], rows2=dev['rows2'],
                    nwsum=geo['nwsum'], col0=geo['col0'], row0=geo['row0'], K=geo['K'], sel=sel,
                    tile_host=tile_host, second=dev.get('second', False), Tns=geo['Tns'], per=geo['per'])

    def rebind(self, datasets):
        """New data with the same segmentation: D and IP are copied into this Filter's tensors (their addresses are
        what captured graphs read); gates and start times are taken from the new datasets."""
        assert len(datasets) == len(self.data)
        for m, d in zip(self.data, datasets):
            assert m.D.shape == d.D.shape and m.tlen == d.tlen and list(m.analyze) == list(d.analyze)
            m.D.copy_(d.D)
            m.IP.copy_(d.IP)
            m.gates, m.gps_start, m.seg_start = list(d.gates), d.gps_start, list(d.seg_start)
        if self.IPall is None:
            self.IPall = torch.cat([m.IP for m in self.data], 0)
        else:
            torch.cat([m.IP for m in self.data], 0, out=self.IPall)
        self._reset_run_state()
        self._y16_update()
        self.T = {k: 0 if isinstance(v, int) else 0.0 for k, v in self.T.items()}
        self.rel = []

    def decode_batch(self, st):
        """Stage B1: read the batch's candidate list back (waiting only for its own stage A), split it by segment,
        and enqueue the rebuilt phases (fp_filter.py:759-764) and each segment's exact stage."""
        t0 = time.perf_counter()
        res = st['sel'].result()
        g, il, val = res[:3]
        # (candidates, LM - 1) window offsets of each candidate's alternates, -1 past the last
        offa = res[3] if len(res) > 3 else None
        if st.get('second'):
            self._hot_update(st, g, val)
        st['pend'] = []
        self.T['n_candidates'] += len(g)
        if len(g) == 0:
            return st
        tb, segs, nwsum, col0, row0, K = st['tb'], st['segs'], st['nwsum'], st['col0'], st['row0'], st['K']
        n = len(tb.ids)
        m_ = (g // nwsum).astype(np.int64)
        c_ = g % nwsum
        si = np.searchsorted(col0, c_, side='right') - 1                # batch-local segment of each candidate
        il = il.astype(np.int64)
        approx = np.sqrt(val)
        rows_d, tcol, sidx = _dev.to_dev_many(m_.astype(np.int32), (row0[si] + il).astype(np.int32),
                                              (si * n + m_).astype(np.int32))
        zreb_d = torch.empty(len(g), dtype=torch.complex64, device='cuda')
        _lib.call('gwgs_rebuilt_z', self.code, st['Zt'].data_ptr(), K, st['Ar'].data_ptr(), st['Ai'].data_ptr(), K,
                  K, rows_d.data_ptr(), tcol.data_ptr(), sidx.data_ptr(), len(g), st['rows2'].data_ptr(),
                  zreb_d.data_ptr(), _dev.stream())
        st['zreb'] = _dev.Fetch(zreb_d)
        o = self.opts
        xparts = []                           # (pend index, tile rows, samples, window starts) of exact candidates
        for i, (di, s, gs, a0, nw) in enumerate(segs):
            sel = np.flatnonzero(si == i)
            if not len(sel):
                continue
            m = self.data[di]
            if st.get('gate') is not None:
                # fp_filter.py:754-757: only peaks within the window of a selected trigger
                k_ = near(a0 + il[sel], st['gate'][(di, s)], int(o.gate.window * m.sr))
                self.T['gate_peaks_dropped'] = self.T.get('gate_peaks_dropped', 0) + int((~k_).sum())
                sel = sel[k_]
                if not len(sel):
                    continue
            tb_, samp, ax = m_[sel], a0 + il[sel], approx[sel]
            ex = ax >= o.exact_snr * (1 - o.margin)
            if ex.any():
                sx = samp[ex]
                # the cluster window of each candidate bounds its re-centring (a0 + window index * cw)
                wlo = a0 + (sx - a0) // self.cw * self.cw
                s2 = offa[sel][ex] if offa is not None else np.full((len(sx), 1), -1)
                xparts.append((len(st['pend']), tb_[ex], sx, wlo, s2))
            st['pend'].append([di, s, gs, sel, tb_, samp, ax, ex, None])
        if xparts:
            # the exact stage of every segment of the batch in one edges launch and one sums launch
            pend = st['pend']
            seglist = [SegmentData(self.data[pend[k][0]].D[pend[k][1]], self.data[pend[k][0]].IP[pend[k][1]],
                                   self.df, self.tlen) for k, _, _, _, _ in xparts]
            rows = np.concatenate([x[1] for x in xparts])
            smp = np.concatenate([x[2] for x in xparts])
            sidx = np.repeat(np.arange(len(xparts)), [len(x[1]) for x in xparts])
            if o.refine > 0:
                wlo = np.concatenate([x[3] for x in xparts])
                s2 = np.concatenate([x[4] for x in xparts])
                # one more request per alternate: every request re-centred in its window, each candidate keeping
                # the largest exact |z| (_PartValues)
                ci, ai = np.nonzero(s2 >= 0)
                req_rows, req_smp = np.r_[rows, rows[ci]], np.r_[smp, wlo[ci] + s2[ci, ai]]
                req_lo = np.r_[wlo, wlo[ci]]
                pv = tb.stage.evaluate_async_multi(seglist, np.r_[sidx, sidx[ci]], PeakRequest(req_rows, req_smp),
                                                   refine=o.refine, lo=req_lo, hi=req_lo + self.cw)
            else:
                pv = tb.stage.evaluate_async_multi(seglist, sidx, PeakRequest(rows, smp))
                ci = np.zeros(0, np.int64)
            aj = len(rows) + np.arange(len(ci))
            self.T['n_alternates'] = self.T.get('n_alternates', 0) + len(ci)
            shared = _SharedValues(pv)
            off = 0
            for k, r_, _, _, _ in xparts:
                m_ = (ci >= off) & (ci < off + len(r_))
                pend[k][8] = _PartValues(shared, off, len(r_), (ci[m_] - off, aj[m_]))
                off += len(r_)
        # the batch's device buffers stay referenced by st until collect_batch
        self.T['rebuilt_z_s'] += time.perf_counter() - t0
        return st

    def _second_frac(self):
        """The hot-window fraction below which the local-maximum lists come from a second GEMM pass."""
        f = self.opts.alt_second_pass
        return SECOND_PASS_DEFAULT[self.opts.precision] if f is None else f

    def _hot_update(self, st, g, val):
        """The running fraction of windows holding exact-band candidates, for a batch whose local-maximum lists came
        from the second pass: while it stays below _second_frac() the next batches take that pass too, else
        the lists return to the first GEMM."""
        o = self.opts
        xcut = np.float32((o.exact_snr * (1 - o.margin)) ** 2)
        cols = np.unique(g[val >= xcut] % st['nwsum']) if len(g) else np.zeros(0, np.int64)
        frac = len(cols) / max(st['nwsum'], 1)
        self._hot = frac if self._hot is None else 0.5 * (self._hot + frac)
        self.T['second_pass_windows'] = self.T.get('second_pass_windows', 0) + len(cols)

    def _sg_values(self, locs, m, s, tid, samples, snr, chi, dof, fpeak):
        """fp_filter.py:854-883: the sine-Gaussian chi-squared of exact peaks with newsnr >= the threshold whose
        template has a tile list, one launch per distinct tile list (sgchisq.sg_chisq_launch on the segment's D and
        IP, the template's kmin as the first tile's lower bound), enqueued only. Returns (values float32, 1 where not
        computed; computed mask; [(indices, PendingSG)]): _resolve_sg fills the first two in place."""
        n = len(snr)
        sgc, sgd = np.ones(n, np.float32), np.zeros(n, bool)
        thr = self.opts.sgchisq_snr_threshold
        if thr is None or locs is None or n == 0:
            return sgc, sgd, []
        t0 = time.perf_counter()
        ns = consolidate.newsnr(snr, chi / np.maximum(dof, 1))
        groups = {}
        for i in np.nonzero(ns >= thr)[0]:
            if locs[tid[i]] is not None:
                groups.setdefault(tuple(locs[tid[i]]), []).append(i)
        pend = []
        for vals, gi in groups.items():
            gi = np.array(gi)
            pend.append((gi, sgchisq.sg_chisq_launch(m.D[s], m.IP[s], samples[gi], fpeak[gi], list(vals),
                                                     np.full(len(gi), self.kmin), self.df, self.tlen)))
        self.T['sg_chisq_s'] += time.perf_counter() - t0
        return sgc, sgd, pend

    def collect_batch(self, st):
        """Stage B2: wait for the batch's exact values and finalise its triggers (fp_filter.py:768-853)."""
        t0 = time.perf_counter()
        if not st['pend']:
            return
        outs = st['outs']
        zreb, = st['zreb'].get()
        sig_h, norm_h = st['tile_host'].get()
        tb = st['tb']
        o = self.opts
        thr = o.snr_threshold
        for di, s, gs, sel, tb_, samp, ax, ex, pv in st['pend']:
            m = self.data[di]
            out = outs[di]
            tstamp = lambda sm: m.gps_start + (m.seg_start[s] + sm) / m.sr
            rb = ~ex & (ax > thr)
            if rb.any():
                nrb = int(rb.sum())
                out.add(tid=tb.ids[tb_[rb]], t=tstamp(samp[rb]), snr=ax[rb],
                        ph=np.angle(zreb[sel][rb]).astype(np.float32), chi=np.zeros(nrb, np.float32),
                        dof=np.zeros(nrb, np.int32), sg=sig_h[tb_[rb], gs], ex=np.zeros(nrb, bool), samp=samp[rb],
                        seg=np.full(nrb, gs), sgc=np.ones(nrb, np.float32), sgd=np.zeros(nrb, bool))
                self.T['n_rebuilt_peaks'] += nrb
            tbx = tb_[ex]
            self.T['n_exact_candidates'] += len(tbx)
            if pv is None:
                continue
            v = pv.result()
            sx = v.samples
            nr = norm_h[tbx, gs]
            zx = v.z.astype(np.complex128)
            snr = np.abs(zx) * nr
            self.rel.append(ax[ex] / np.maximum(snr, 1e-30))
            k = snr > thr
            nbin = tb.stage.tile.nbins[tbx]
            chi = ((nbin * v.pw.astype(np.float64) - np.abs(zx) ** 2) * nr ** 2).astype(np.float32)
            dof = (2 * nbin - 2).astype(np.int32)
            sgc, sgd, pend = self._sg_values(st.get('sg_locs'), m, s, tb.ids[tbx[k]], sx[k], snr[k], chi[k],
                                             dof[k], v.fpeak[k])
            out.add(tid=tb.ids[tbx[k]], t=tstamp(sx[k]), snr=snr[k].astype(np.float32),
                    ph=np.angle(v.z[k]).astype(np.float32), chi=chi[k], dof=dof[k], sg=sig_h[tbx[k], gs],
                    ex=np.ones(int(k.sum()), bool), samp=sx[k], seg=np.full(int(k.sum()), gs), sgc=sgc, sgd=sgd)
            if pend:
                out.unready.append(sgc)
                self._sg_pending.append((out, sgc, sgd, pend))
            self.T['n_chisq_points'] += int(k.sum())
        self.T['refine_chisq_s'] += time.perf_counter() - t0


# ---------------------------------------------------------------------------------------------------------------------
# Peak files and the command line (fp_filter.py:47-160, 910-977)
# ---------------------------------------------------------------------------------------------------------------------

PEAK_FIELDS = ('template_id', 'template_hash', 'snr', 'end_time', 'sigmasq', 'coa_phase', 'peak_sigmasq', 'exact',
               'chisq', 'chisq_dof', 'newsnr', 'sg_chisq', 'sg_chisq_computed')


def peak_file_arrays(fields, template_hash, sigmasq, gates, provenance_json='{}', config_md5=''):
    """The datasets and attributes of one peak file (fp_filter.py:910-937) from Triggers.to_dict fields (gate veto
    and (template_id, end_time) order already applied): 1-D fields of equal length, template_hash from the bank,
    the 2-D sigmasq (templates x first segments, float64), attrs n_gates / provenance / config_md5 and the gates
    group time / width / taper. Keys are the HDF5 paths ('attrs/...' for attributes)."""
    out = {'template_id': np.asarray(fields['template_id'], np.int64),
           'template_hash': np.asarray(template_hash)[np.asarray(fields['template_id'], np.int64)],
           'snr': fields['snr'], 'end_time': fields['end_time'],
           'sigmasq': np.asarray(sigmasq if sigmasq is not None else np.zeros((len(template_hash), 0)), np.float64),
           'coa_phase': fields['coa_phase'], 'peak_sigmasq': fields['peak_sigmasq'], 'exact': fields['exact']}
    for k in ('chisq', 'chisq_dof', 'newsnr', 'sg_chisq', 'sg_chisq_computed'):
        if k in fields:
            out[k] = fields[k]
    gt = np.array(list(gates), np.float64).reshape(-1, 3)
    out.update({'attrs/n_gates': np.int64(len(gt)), 'attrs/provenance': np.array(provenance_json),
                'attrs/config_md5': np.array(config_md5), 'gates/time': gt[:, 0], 'gates/width': gt[:, 1],
                'gates/taper': gt[:, 2]})
    return out


def write_peaks(path, arrays):
    """Write a peak file atomically (fp_filter.py:918-933): to path + '.part', then renamed, so a reader never sees a
    partial file. h5py is not available here: the file is a numpy .npz archive whose keys are the HDF5 dataset paths
    (read back with read_peaks)."""
    import os
    tmp = path + '.part'
    try:
        with open(tmp, 'wb') as f:
            np.savez(f, **arrays)
    except BaseException:
        if os.path.exists(tmp):
            os.remove(tmp)
        raise
    os.replace(tmp, path)


def read_peaks(path):
    """A peak file written by write_peaks: dict of its datasets (attributes under 'attrs/...')."""
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def template_hashes(m1, m2, e0=None):
    """Stable 64-bit template hashes for banks that carry none (the reference takes PyCBC's bank table column): the
    first 8 bytes of md5 over the float64 parameters."""
    import hashlib
    e0 = np.zeros(len(m1)) if e0 is None else np.asarray(e0, np.float64)
    out = np.empty(len(m1), np.int64)
    for i, v in enumerate(np.stack([np.asarray(m1, np.float64), np.asarray(m2, np.float64), e0], 1)):
        out[i] = np.frombuffer(hashlib.md5(v.tobytes()).digest()[:8], np.int64)[0]
    return out


def load_bank(path, approximant, max_templates=None):
    """A template bank file (numpy .npz: mass1, mass2, optional template_hash, optional ecc / e0) for an approximant
    this path generates: the closed-form 0PN+1PN stationary-phase chirp ('SPAtmplt', 'TaylorF2'), with the leading
    eccentric phase term for an eccentric approximant ('SEOBNRv5EHM_ecc', 'TaylorF2Ecc'). Returns (m1, m2, e0,
    template_hash). Other approximants need PyCBC's waveform generators, which are not available here."""
    ecc_names = ('SEOBNRv5EHM_ecc', 'TaylorF2Ecc')
    if approximant not in ('SPAtmplt', 'TaylorF2') + ecc_names:
        raise NotImplementedError(f'approximant {approximant}: only the closed-form SPA chirps are generated here')
    with np.load(path, allow_pickle=False) as z:
        m1, m2 = np.asarray(z['mass1'], float), np.asarray(z['mass2'], float)
        e0 = None
        for k in ('ecc', 'e0', 'ecc_b'):
            if k in z.files and approximant in ecc_names:
                e0 = np.asarray(z[k], float)
        th = np.asarray(z['template_hash']) if 'template_hash' in z.files else None
    if max_templates is not None:
        m1, m2 = m1[:max_templates], m2[:max_templates]
        e0 = None if e0 is None else e0[:max_templates]
        th = None if th is None else th[:max_templates]
    if th is None:
        th = template_hashes(m1, m2, e0)
    return m1, m2, e0, th


def prep_one(job):
    """fp_filter.py:53-97: condition one dataset and save D and IP to an .npz, returning its metadata. job = (tag,
    argv, outdir, cache, wait_s). With a cache directory a finished result is reused (key: the hash of the full argv;
    written under a temporary name and renamed). The conditioning is gpu_prep.prep (the GPU stages in this process;
    the reference runs PyCBC on CPU workers here)."""
    import hashlib, json, os
    import gpu_prep
    tag, argv, outdir, cache, wait_s = job
    key = hashlib.sha1(' '.join(argv).encode()).hexdigest()[:16]
    if cache:
        cj = os.path.join(cache, f'{tag}-{key}.json')
        t_wait = time.time() + wait_s
        while wait_s and not os.path.exists(cj) and time.time() < t_wait:
            time.sleep(5)
        if os.path.exists(cj):
            meta = json.load(open(cj))
            meta['prep_s'] = 0.0
            meta['cached'] = True
            return meta
    t0 = time.perf_counter()
    g = gpu_prep.prep(argv)
    os.makedirs(cache or outdir, exist_ok=True)
    fn = os.path.join(cache or outdir, f'{tag}-{key}.npz')
    with open(fn + '.part', 'wb') as f:
        np.savez(f, D=g['D'].cpu().numpy(), IP=g['IP'].cpu().numpy())
    os.replace(fn + '.part', fn)
    meta = dict(tag=tag, file=fn, flen=g['flen'], tlen=g['tlen'], df=g['df'], sr=g['sr'], gps_start=g['gps_start'],
                analyze=[tuple(map(int, a)) for a in g['analyze']], seg_start=[int(x) for x in g['seg_start']],
                gates=[tuple(map(float, x)) for x in g['gates']], prep_s=time.perf_counter() - t0,
                injection_file=gpu_prep.options(argv).injection_file, cached=False)
    if cache:
        json.dump(meta, open(cj + '.part', 'w'))
        os.replace(cj + '.part', cj)
    return meta


def _chisq_bins(spec):
    spec = str(spec).strip('"\'')
    try:
        return int(spec)
    except ValueError:
        raise NotImplementedError(f'--chisq-bins {spec!r}: only a fixed bin count is supported (an expression needs '
                                  f'PyCBC)') from None


def main(argv=None):
    """fp_filter.py:100-160 and run_group: the command line of the reference's filter.

    fp_filter.py --datasets datasets.txt --bank BANK.npz --approximant A --argv inspiral_argv.txt --out OUTDIR
        [--tile 1024] [--eps 1e-3] [--margin 0.04] [--exact-snr S] [--consolidate] [--pre-consolidate] ...
        -- [pycbc_inspiral options]

    datasets.txt: one 'IFO GPS_START GPS_END STORE [TRIG_START_OFFSET]' per line; writes OUTDIR/peaks/IFO-START-END.npz
    per dataset (OUTDIR/peaks_b2/ for --bank2) and OUTDIR/timing.json. --templates interp (the default) takes the
    closed-form chirps at coarse nodes, interpolated; --templates pycbc takes them closed-form on every bin (PyCBC's
    generators are absent); --exact-stage regen regenerates the interpolated templates in the exact stage. Flags the
    reference accepts for choices this path does not have (edge algorithms, kernel variants, worker counts) are
    accepted and ignored; --queue DIR filters the DIR/*.job groups one after another."""
    import argparse, json, os, shlex, sys
    import consolidate as cons
    import gpu_prep
    import window_filter
    from multi_filter import pycbc_parser, setopt, dataset_argv
    import provenance
    args = list(sys.argv[1:] if argv is None else argv)
    split = args.index('--') if '--' in args else len(args)
    ap = argparse.ArgumentParser(prog='fp_filter')
    ap.add_argument('--datasets'); ap.add_argument('--bank', required=True)
    ap.add_argument('--approximant', required=True); ap.add_argument('--argv', required=True)
    ap.add_argument('--out'); ap.add_argument('--queue')
    ap.add_argument('--pre-consolidate', action='store_true')
    ap.add_argument('--tile', type=int, default=1024); ap.add_argument('--eps', type=float, default=1e-3)
    ap.add_argument('--margin', type=float, default=0.04)
    ap.add_argument('--templates', choices=['interp', 'pycbc'], default='interp')
    ap.add_argument('--prep-procs', type=int, default=8); ap.add_argument('--gen-threads', type=int, default=8)
    ap.add_argument('--n-chirp', type=int, default=6000); ap.add_argument('--n-log', type=int, default=1500)
    ap.add_argument('--max-templates', type=int)
    ap.add_argument('--prep-cache'); ap.add_argument('--wait-prep', type=float, default=0)
    ap.add_argument('--gemm-cfg', help='tiles_per_block[,variant] for the GEMM launch; skips the tuning sweep')
    ap.add_argument('--prep-only', action='store_true')
    ap.add_argument('--prep', choices=['cpu', 'gpu'], default='cpu',
                    help='cpu: prep_one per dataset through an .npz (reused with --prep-cache); gpu: in memory')
    ap.add_argument('--edges', choices=['blocked', 'full', 'fused'], default='blocked')
    ap.add_argument('--exact-snr', type=float, default=0)
    ap.add_argument('--consolidate', action='store_true')
    ap.add_argument('--consolidate-exempt-snr', type=float, default=None)
    ap.add_argument('--exact-stage', choices=['dense', 'regen'], default='dense')
    ap.add_argument('--fuse-q', action='store_true'); ap.add_argument('--sg-kernel', action='store_true')
    ap.add_argument('--sg-batch', action='store_true'); ap.add_argument('--legacy-qzero', action='store_true')
    ap.add_argument('--chunk'); ap.add_argument('--chunk2'); ap.add_argument('--gate')
    ap.add_argument('--gate-mc-ratio', type=float, default=1.25); ap.add_argument('--gate-window', type=float, default=1.0)
    ap.add_argument('--bank2'); ap.add_argument('--approximant2', default='IMRPhenomD'); ap.add_argument('--eps2', type=float)
    ap.add_argument('--history')
    ap.add_argument('--precision', choices=list(PRECISIONS), default='bf16', help='GEMM operand type (this path)')
    ap.add_argument('--refine', type=int, default=0, help='exact-stage re-centring half width (0: the reference rule)')
    a = ap.parse_args(args[:split])
    assert not a.gate or a.chunk, '--gate needs --chunk'
    base = shlex.split(open(a.argv).read()) + args[split + 1:]
    base = setopt(setopt(base, '--bank-file', a.bank), '--approximant', a.approximant)
    assert bool(a.queue) != bool(a.datasets and a.out), 'give --datasets and --out, or --queue'
    assert not a.pre_consolidate or (a.consolidate and not a.gate), '--pre-consolidate needs --consolidate and no --gate'
    if a.prep_cache:
        os.makedirs(a.prep_cache, exist_ok=True)

    def ds_argv(i, s_, e_, st, *trig_start):
        av = dataset_argv(base, i, int(s_), int(e_), st)
        return setopt(av, '--trig-start-time', str(int(s_) + int(trig_start[0]))) if trig_start else av

    def read_jobs(ds_file, out):
        ds = [ln.split() for ln in open(ds_file) if ln.strip() and not ln.startswith('#')]
        return [(f'{d[0]}-{d[1]}-{d[2]}', ds_argv(*d), os.path.join(out, 'prep'), a.prep_cache, a.wait_prep) for d in ds]

    def run_group(datasets_file, out):
        T = dict(prep_s=0.0, templates_s=0.0, svd_s=0.0)
        t_start = time.perf_counter()
        jobs = read_jobs(datasets_file, out)
        if a.prep_only:
            assert a.prep_cache, '--prep-only needs --prep-cache'
            for j in jobs:
                prep_one(j)
            return
        for d in ('peaks', 'prep'):
            os.makedirs(os.path.join(out, d), exist_ok=True)
        data = []
        t0 = time.perf_counter()
        for tag, av, pdir, cache, wait in jobs:
            if a.prep == 'gpu':
                g = gpu_prep.prep(av)
                D, IP, meta = g['D'], g['IP'], g
            else:
                meta = prep_one((tag, av, pdir, cache, wait))
                with np.load(meta['file']) as z:
                    D, IP = torch.from_numpy(z['D']).cuda(), torch.from_numpy(z['IP']).cuda()
                if not a.prep_cache:
                    os.remove(meta['file'])
            data.append(Dataset(tag, D, IP, int(meta['flen']), int(meta['tlen']), float(meta['df']),
                                float(meta['sr']), float(meta['gps_start']), [tuple(x) for x in meta['analyze']],
                                list(meta['seg_start']), [tuple(x) for x in meta['gates']]))
        T['prep_s'] = time.perf_counter() - t0
        opt, _ = pycbc_parser().parse_known_args(base)
        nbins = _chisq_bins(opt.chisq_bins)
        sgp = argparse.ArgumentParser(add_help=False)
        sgp.add_argument('--sgchisq-snr-threshold', type=float); sgp.add_argument('--sgchisq-locations', nargs='+')
        sgo, _ = sgp.parse_known_args(base)
        kl = cons.KeepLoudest.from_pycbc(opt, a.consolidate_exempt_snr) if a.consolidate else None
        gate = None
        if a.gate:
            with np.load(a.gate, allow_pickle=False) as z:
                gate = Gate(np.asarray(z['ifo']).astype(str), z['time'], z['mchirp'], a.gate_mc_ratio, a.gate_window)
        prov = provenance.record(dict(bank=a.bank, bank2=a.bank2, datasets=datasets_file, argv_file=a.argv,
                                      gate=a.gate), argv=['fp_filter'] + args + ['[resolved]'] + base,
                                 dq_files=[datasets_file],
                                 dq_flags=[x for i, x in enumerate(base) if x.startswith('--autogating')
                                           or (i and base[i - 1].startswith('--autogating'))])
        T['provenance'] = prov
        provenance.log(a.history, 'fp_filter', 'start', out, prov)
        banks = [(a.bank, a.approximant, 'peaks', a.eps, a.chunk)]
        if a.bank2:
            banks.append((a.bank2, a.approximant2, 'peaks_b2', a.eps2 or a.eps, a.chunk2))
        T['n_peaks'] = {}
        whole = data
        for bank_i, (path, appx, pdir, eps_b, chunk) in enumerate(banks):
            ds = window_filter.to_chunks(whole, chunk)[0] if chunk else whole
            m1, m2, e0, th = load_bank(path, appx, a.max_templates)
            o = Options(snr_threshold=opt.snr_threshold, cluster_window=opt.cluster_window, chisq_bins=nbins,
                        low_frequency_cutoff=opt.low_frequency_cutoff, tile=a.tile, eps=eps_b, margin=a.margin,
                        exact_snr=a.exact_snr, precision=a.precision, refine=a.refine,
                        templates='interp' if a.templates == 'interp' else 'closed', exact_stage=a.exact_stage,
                        n_chirp=a.n_chirp, n_log=a.n_log,
                        gemm_cfg=tuple(int(v) for v in a.gemm_cfg.split(',')) if a.gemm_cfg else None,
                        consolidate=kl, pre_consolidate=a.pre_consolidate,
                        sgchisq_snr_threshold=sgo.sgchisq_snr_threshold if sgo.sgchisq_locations else None,
                        gate=gate if chunk else None)
            IPall, sw, isw = svd_metric(ds)
            plan = BankPlan(m1, m2, e0, ds[0].flen, ds[0].df, o, sw, isw, name=os.path.basename(path),
                            sg_locations=sgo.sgchisq_locations if o.sgchisq_snr_threshold is not None else None)
            T['templates_s'] += plan.T['templates_s']; T['svd_s'] += plan.T['svd_s']
            f = Filter(ds, o)
            trs = f.run(plan)
            for k, v in f.T.items():
                if isinstance(v, (int, float)):
                    T[k + ('' if bank_i == 0 else f'_b{bank_i + 1}')] = v
            os.makedirs(os.path.join(out, pdir), exist_ok=True)
            for d, tr in zip(ds, trs):
                fields = tr.to_dict(gates=d.gates, chisq=nbins > 0)
                arr = peak_file_arrays(fields, th, tr.sigmasq, d.gates, json.dumps(prov, sort_keys=True),
                                       prov['config_md5'])
                write_peaks(os.path.join(out, pdir, d.tag + '.npz'), arr)
                T['n_peaks'][d.tag + ('' if bank_i == 0 else f'_b{bank_i + 1}')] = int(len(fields['snr']))
            T.setdefault('ranks', {})[os.path.basename(path)] = plan.ranks
        T['total_s'] = time.perf_counter() - t_start
        T['data_ks'] = sum((a1 - a0) / m.sr for m in data for a0, a1 in m.analyze) / 1000
        json.dump(T, open(os.path.join(out, 'timing.json'), 'w'), indent=1, default=float)
        provenance.log(a.history, 'fp_filter', 'done', out, prov, n_peaks=sum(T['n_peaks'].values()),
                       total_s=T['total_s'])
        return T

    if not a.queue:
        return run_group(a.datasets, a.out)
    seen = set()
    while True:
        js = sorted(f_ for f_ in os.listdir(a.queue) if f_.endswith('.job') and f_ not in seen)
        if not js:
            if os.path.exists(os.path.join(a.queue, 'END')):
                return None
            time.sleep(2)
            continue
        seen.add(js[0])
        ds_file, out = open(os.path.join(a.queue, js[0])).read().split()[:2]
        run_group(ds_file, out)


if __name__ == '__main__':
    import os as _os, sys as _sys
    _sys.path.insert(0, _os.path.dirname(_os.path.abspath(__file__)))
    main()
