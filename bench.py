#!/usr/bin/env python
"""Throughput of the filtering path on one MI355X: template-seconds filtered per GPU-second.

Headline workload (the batched configuration, workloads.batched_device): per detector 200 chunks of 64-1024 s in 16
s steps (segment gaps), 2 detectors at 2048 Hz of synthetic coloured noise with 5 lines, drawn on the GPU; both
banks, 4,612 circular + 1,268 eccentric = 5,880 templates; chunks on 512 s, 256 s and 128 s segment grids. The job
is cut into 20 slices of 10 chunks per detector: one step filters one slice (conditioning of every chunk from its
raw strain resident in HBM: float64 median-Welch PSD, truncation, whitening; template norms; per tile the basis
transforms and the fused GEMM + window max; candidate selection; the exact stage with +-3-sample re-centring: edges,
chi-squared sums, float64 finalization), so the 20 default steps filter every chunk once. Warm-up steps run on
separate chunks. The bank products (templates, float64 SVD bases) are built once before timing, as a search builds
them once per run; config.value_with_setup counts them, the autotuner and the SVD metric's conditioning into the
whole job (not the synthetic data's generation, which stands in for reading strain), and config.setup_parts breaks
that down.

value = 5,880 x analysed seconds of the timed slices (both detectors) / their wall time, for the BF16 form. The
FP32 plain form (and the FP8 form) are timed on the first slices beside it, the trigger lists compared with the
BF16 form's, and every exact trigger's SNR of the loudest templates compared with a float64 direct matched filter of
the same conditioned data (torch complex128 FFTs). The production workload (2 detectors x 512 s with lines,
glitches and 20 injections) follows as a secondary object with its reports (consolidation at the production cuts,
overlap-save chunks, refinement, time slides).

Prints one JSON line. Progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, 'gw-gpu-search_amd')
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = 'template·s filtered per GPU-s at 5,880 templates, BF16 vs FP32 at SNR err≤1e-3'
PEAK = {'bf16': 2500.0, 'fp32': 157.3, 'fp8': 5000.0}      # dense MFMA TFLOP/s, MI355X


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def direct_snr_check(datasets, plans, outs, max_templates=24):
    """Relative error of reported exact SNRs against a float64 direct frequency-domain matched filter on the same D
    and IP: |IFFT(conj(h) D)| 4 df / sigma in complex128, for the exact triggers of up to max_templates templates
    per (dataset, bank), the loudest first."""
    import numpy as np
    import torch
    errs = []
    for bi, plan in enumerate(plans):
        tile_of = {}
        for tb in plan.tiles:
            for r, tid in enumerate(tb.ids):
                tile_of[int(tid)] = (tb, r)
        for di, d in enumerate(datasets):
            tr = outs[bi][di]
            ex = tr['exact']
            if not ex.any():
                continue
            tids = tr['template_id'][ex]; samp = tr['sample'][ex]; snr = tr['snr'][ex]; seg = tr['segment'][ex]
            order = np.argsort(-snr)
            series = {}
            for i in order:
                t = int(tids[i])
                s = int(seg[i]) - sum(len(x.analyze) for x in datasets[:di])
                if (t, s) not in series:
                    if len(series) >= max_templates:
                        continue
                    tb, r = tile_of[t]
                    h = tb.H[r].to(torch.complex128)
                    D = d.D[s].to(torch.complex128); ip = d.IP[s].to(torch.float64)
                    sig = 4 * d.df * float((h.abs() ** 2 * ip).sum())
                    q = torch.zeros(d.tlen, dtype=torch.complex128, device='cuda')
                    q[:d.flen] = torch.conj(h) * D
                    series[(t, s)] = ((torch.fft.ifft(q, norm='forward').abs() * (4 * d.df / np.sqrt(sig)))
                                      .cpu().numpy())
                ref = series[(t, s)][int(samp[i])]
                errs.append(abs(float(snr[i]) - ref) / ref)
    return (float(max(errs)) if errs else None), len(errs)


def compare_triggers(a, b):
    """Exact above-threshold trigger lists of two forms: (template, sample) sets; (template, 1 s window) sets; for
    windows in both, the fraction at the same sample, the largest sample offset and the largest relative SNR gap."""
    import numpy as np
    ea, eb = a['exact'], b['exact']
    ka = set(zip(a['template_id'][ea].tolist(), a['sample'][ea].tolist()))
    kb = set(zip(b['template_id'][eb].tolist(), b['sample'][eb].tolist()))
    wa = {(int(t), int(s) // 2048): (int(s), float(v)) for t, s, v in
          zip(a['template_id'][ea], a['sample'][ea], a['snr'][ea])}
    wb = {(int(t), int(s) // 2048): (int(s), float(v)) for t, s, v in
          zip(b['template_id'][eb], b['sample'][eb], b['snr'][eb])}
    both = set(wa) & set(wb)
    moved = [k for k in both if wa[k][0] != wb[k][0]]
    return dict(n_a=len(ka), n_b=len(kb), only_a=len(ka - kb), only_b=len(kb - ka),
                windows_only_a=len(set(wa) - set(wb)), windows_only_b=len(set(wb) - set(wa)),
                same_sample=(len(both) - len(moved)) / max(len(both), 1),
                max_sample_offset=max([abs(wa[k][0] - wb[k][0]) for k in moved], default=0),
                max_snr_rel_gap_moved=max([abs(wa[k][1] - wb[k][1]) / wa[k][1] for k in moved], default=0.0))


def measured_traffic():
    """HBM bytes per GEMM launch of the headline workload from the committed rocprofv3 --pmc summary (FETCH_SIZE +
    WRITE_SIZE, scaled by a plain copy of known size), or None when the summary is absent."""
    import re
    p = os.path.join(ROOT, 'profiles', 'hbm_traffic_batched_pmc.md')
    if not os.path.exists(p):
        return None
    m = re.search(r'traffic_bytes_per_launch \(gemm_peaks, bf16\) = ([0-9.eE+]+)', open(p).read())
    return float(m.group(1)) if m else None


def cpu_baseline(x, sr, gps_start, cond, bank, label, ntemp=512):
    """The oracle's plain method (numpy / scipy.fft / OpenBLAS float32) on a bounded sample: one detector stretch x
    (float32 strain, host), the first `ntemp` circular templates in chirp-time order as one tile; conditioning plus
    the per-segment work (basis transforms, GEMM and window maxima, exact stage) is timed, the tile's templates and SVD
    are not (as on the GPU, where the bank products are made once per run)."""
    import numpy as np
    from oracle import core, pipeline
    cores = min(len(os.sched_getaffinity(0)), int(os.environ.get('OMP_NUM_THREADS', '16')))
    ids = pipeline.tiles_by_tau0(bank.m1, bank.m2, 15.0, ntemp)[0]
    t0 = time.perf_counter()
    d = core.condition(x, sr, gps_start, cond.seg_len, cond.start_pad, cond.end_pad, cond.psd_seg_len,
                       cond.psd_stride, cond.psd_nseg, cond.psd_inv_len, 15.0)
    t_cond = time.perf_counter() - t0
    T = {}
    pipeline.filter_plain([d], bank.m1[ids], bank.m2[ids], None, tile=ntemp, dtype=np.float32, workers=cores,
                          timing=T)
    t = t_cond + T['basis_fft_s'] + T['gemm_s'] + T['exact_s']
    an = sum((a1 - a0) // 2048 * 2048 / sr for a0, a1 in d['analyze'])
    return dict(value=ntemp * an / t, unit='template*s/CPU-s', cores=cores, kind='port',
                per_core=ntemp * an / t / cores,
                sample=f'{label}: {an:.0f} analysed s, {ntemp} circular templates as one tile, float32 plain method '
                       f'(oracle/pipeline.py), conditioning + per-segment work timed ({t:.1f} s: conditioning '
                       f'{t_cond:.1f}, basis transforms {T["basis_fft_s"]:.1f}, GEMM and window maxima '
                       f'{T["gemm_s"]:.1f}, exact stage {T["exact_s"]:.1f}), templates/SVD excluded '
                       f'({T["templates_s"] + T["svd_s"]:.1f} s)')


def refinement_report(plans, streams, ds, outs, top, neighbours=8, steps=20):
    """The differentiable refinement on a step's triggers: the `top` loudest exact triggers refined off the bank
    (differentiable.refine_triggers, `steps` projected-gradient steps on the hull of `neighbours` templates), timed;
    the gain rho / rho_start (the compressed template's SNR before and after) over all of them, and per injection the
    loudest exact trigger within 0.1 s of its merger time (the injections' masses are drawn off the bank)."""
    import numpy as np
    import torch
    import differentiable
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = differentiable.refine_triggers(plans, ds, outs, top=top, neighbours=neighbours, steps=steps)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    gain = r['rho'] / np.maximum(r['rho_start'], 1e-30)
    inj = []
    for d, st in enumerate(streams):
        for j in st.injections:
            near = np.flatnonzero((r['dataset'] == d) & (np.abs(r['end_time'] - (st.gps_start + j.t0)) < 0.1))
            if len(near):
                k = near[np.argmax(r['snr'][near])]
                inj.append(dict(det=st.ifo, t0=round(j.t0, 3), m1=round(j.m1, 2), m2=round(j.m2, 2),
                                snr_optimal=round(j.snr, 2), exact_snr=round(float(r['snr'][k]), 4),
                                nearest=round(float(r['rho_start'][k]), 4), refined=round(float(r['rho'][k]), 4),
                                gain=float(gain[k]), delta_samples=round(float(r['delta'][k]), 3)))
    ig = np.array([x['gain'] for x in inj])
    return dict(n_refined=int(len(gain)), steps=steps, neighbours=neighbours, seconds=dt,
                triggers_per_s=len(gain) / max(dt, 1e-9),
                gain_median=float(np.median(gain)) if len(gain) else None,
                gain_max=float(gain.max()) if len(gain) else None,
                frac_improved=float((gain > 1 + 1e-6).mean()) if len(gain) else None,
                injections_found=len(inj), injection_gain_median=float(np.median(ig)) if len(ig) else None,
                injection_gain_max=float(ig.max()) if len(ig) else None, injections=inj)


def slides_report(ds, outs, names, K=500, step=0.1, window=0.012):
    """Time slides over the two detectors' trigger lists of one step, per bank (slides.find_pairs, the reference's
    coinc_rank.find_pairs with its defaults: 0.1 s step, 12 ms window): L1 shifted by k step for |k| <= K over the
    analysed span P, same template, ranked by the quadrature sum of newsnr. Times the pair search."""
    import numpy as np
    import torch
    import slides
    gps0 = min(d.gps_start for d in ds)
    P = max(d.gps_start + (d.seg_start[-1] + d.tlen) / d.sr for d in ds) - gps0
    res = []
    for b, name in enumerate(names):
        H, L = outs[b][0], outs[b][1]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        hh, ll, kk, stat = slides.find_pairs(dict(x=H['end_time'] - gps0, tid=H['template_id']),
                                             dict(x=L['end_time'] - gps0, tid=L['template_id']), P, step, K, window,
                                             stat_fn=lambda h, l, dt: np.sqrt(H['newsnr'][h] ** 2 +
                                                                              L['newsnr'][l] ** 2))
        dt = time.perf_counter() - t0
        res.append(dict(bank=name, slides=2 * K, span_s=P, triggers=(len(H['snr']), len(L['snr'])), pairs=len(hh),
                        zero_lag_pairs=int((kk == 0).sum()), loudest_zero_lag=float(stat[kk == 0].max())
                        if (kk == 0).any() else None, seconds=dt))
    return res


class ProductionWorkload:
    """2 detectors x 512 s at 2048 Hz (lines, glitches, 20 injections), both banks; one 512 s segment each."""

    def __init__(self, opts):
        import torch
        import fp_filter, workloads
        self.streams, self.banks, self.cond = workloads.production()
        self.x_dev = [torch.from_numpy(st.x).cuda() for st in self.streams]
        ds0 = workloads.condition_streams(self.streams, self.cond, [x.clone() for x in self.x_dev])
        IPall, sw, isw = fp_filter.svd_metric(ds0)
        self.opts = opts
        self.filters = {}
        # the production run's sine-Gaussian tiles (used only by forms with a sgchisq threshold)
        self.plans = [fp_filter.BankPlan(b.m1, b.m2, b.e0, ds0[0].flen, ds0[0].df, opts['bf16'], sw, isw,
                                         name=b.name, sg_locations=SG_LOCATIONS) for b in self.banks]

    def step(self, prec, check=False):
        """One pass: conditioning of both streams and the filtering of both banks."""
        import fp_filter, workloads
        ds = workloads.condition_streams(self.streams, self.cond, [x.clone() for x in self.x_dev])
        o = self.opts[prec]
        if o.graphs:
            # one Filter per form for the whole run: its captured graphs read its own D / IP, refreshed by rebind
            f = self.filters.get(prec)
            if f is None:
                f = self.filters[prec] = fp_filter.Filter(ds, o)
            else:
                f.rebind(ds)
            ds = f.data
        else:
            if check:
                o = fp_filter.Options(**{**o.__dict__, 'self_check': True})
            f = fp_filter.Filter(ds, o)
        outs = [[t.to_dict(gates=d.gates) for t, d in zip(tr, ds)] for tr in f.run_banks(self.plans)]
        return ds, [f], outs

    def ranks(self):
        return {p.name: p.ranks for p in self.plans}

    def describe(self, tile):
        return ('2 detectors x 512 s at 2048 Hz (lines, glitches, 20 injections), 4,612 circular + 1,268 eccentric '
                'SPA templates, tiles of %d, SVD worst lost power 1e-4, 16 chi-squared bins, SNR threshold 5.5, 1 s '
                'windows; BF16 GEMM, FP32 exact stage' % tile)

    def compare(self, a, b):
        return [compare_triggers(a[i][d], b[i][d]) for i in range(len(a)) for d in range(len(a[i]))]

    def direct_check(self, ds, outs):
        return direct_snr_check(ds, self.plans, outs)


class BatchedWorkload:
    """The batched configuration: per detector n_chunks chunks of 64-1024 s in 16 s steps (segment gaps), 2 detectors
    at 2048 Hz of noise with the production lines drawn on the GPU, grouped on segment grids (512, 256 and 128 s); both
    banks. The job is cut into nslice slices of n_chunks / nslice chunks per detector; one step filters one slice, so
    nslice steps filter every chunk once. Warm-up steps filter a separate slice of warm chunks (other lengths and
    noise), whose conditioned data also give each grid's SVD metric. Datasets and trigger lists of a step are
    flattened over grids in grid order."""

    def __init__(self, n_chunks, opts, nslice=20, n_warm=10, grids=None):
        import torch
        import batched, workloads
        t0 = time.perf_counter()
        self.grids = tuple(grids or workloads.GRIDS)
        chunks, x_dev, self.banks = workloads.batched_device(n_chunks, grids=self.grids)
        wch, wx, _ = workloads.batched_device(n_warm, seed=97, noise_seed=500000, n_circ=0, n_ecc=0, grids=self.grids)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        self.n_chunks, self.nslice = n_chunks, nslice
        self.slices = batched.slices(n_chunks, 2, nslice)
        self.warm = list(range(len(chunks), len(chunks) + len(wch)))
        self.search = batched.BatchedSearch(chunks + wch, self.banks, x_dev + wx, opts, ref=self.warm)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        plans = [p for g in self.search.grids for p in self.search.plans[g]]
        tm = sum(p.T['templates_s'] for p in plans)
        sv = sum(p.T['svd_s'] for p in plans)
        # synthetic data (not part of a search's set-up) apart from the bank products: templates and SVD bases of
        # every grid, and the rest of the search set-up (the reference chunks' conditioning for the SVD metric)
        self.setup = dict(data_s=t1 - t0, templates_s=tm, svd_s=sv, search_other_s=t2 - t1 - tm - sv)
        self._last_ds = None

    def _flat(self, ds, outs):
        g = self.search.grids
        return [d for x in g for d in ds[x]], [[t for x in g for t in outs[x][b]] for b in range(len(self.banks))]

    def step(self, prec, k=None, check=False):
        """One step: slice k (mod nslice), or the warm slice for k None. Returns (datasets, filters, trigger lists)
        flattened over grids."""
        sel = self.warm if k is None else self.slices[k % self.nslice]
        ds, outs, filters = self.search.step(prec, sel, self_check=check)
        self._last_ds = ds
        fd, fo = self._flat(ds, outs)
        return fd, filters, fo

    def run(self, prec, ks, check=False):
        """Slices ks in one pipeline (the next slice conditioned while the previous one's last batches run). Returns
        per slice (grouped datasets, filters, trigger lists flattened over grids)."""
        res = self.search.run(prec, [self.slices[k % self.nslice] for k in ks], self_check=check)
        return [(ds, filters, self._flat(ds, outs)[1]) for ds, outs, filters in res]

    def ranks(self):
        return self.search.ranks()

    def describe(self, tile):
        return ('batched: %d chunks per detector of 64-1024 s in 16 s steps, 2 detectors at 2048 Hz (noise with lines), '
                'grouped on %s s segment grids; 4,612 circular + 1,268 eccentric SPA templates, tiles of '
                '%d, SVD worst lost power 1e-4, 16 chi-squared bins, SNR threshold 5.5, 1 s windows; BF16 GEMM, FP32 '
                'exact stage; one step = one of %d slices of %d chunks per detector' %
                (self.n_chunks, ' / '.join(str(g) for g in sorted(self.grids, reverse=True)), tile, self.nslice,
                 self.n_chunks // self.nslice))

    def compare(self, a, b):
        # per bank, the lists of every dataset merged (keys are unique: each dataset has its own end times)
        return [compare_triggers(_merge(a[i]), _merge(b[i])) for i in range(len(a))]

    def direct_check(self, ds, outs, max_templates=12):
        """The float64 direct filter on a step's datasets (grouped by grid; default: the last step's), grid by grid
        (each grid's plans hold the dense templates on its frequency bins)."""
        ds = self._last_ds if ds is None else ds
        errs, n = [], 0
        off = 0
        for g in self.search.grids:
            nd = len(ds[g])
            if nd:
                part = [outs[b][off:off + nd] for b in range(len(outs))]
                e, k = direct_snr_check(ds[g], self.search.plans[g], part, max_templates=max_templates)
                if e is not None:
                    errs.append(e)
                n += k
            off += nd
        return (max(errs) if errs else None), n


SG_LOCATIONS = ['mtotal>30:20-15,20-30,20-45,20-60,20-75,20-90,20-105,20-120']


def chunked_report(wl, cfg, tile, hist_s=8.0, chunk_batch=(64,)):
    """Overlap-save chunks (fp_filter.py:324-357 --chunk / --chunk2): the circular bank's templates shorter than 6 s
    from 15 Hz against the production data as whole 512 s segments, as 16 s chunks (8 s history, 2 s future) and as
    the same chunks band-limited to 512 Hz (sampled at 1024 Hz), and both gated by the whole-segment search's
    triggers above SNR 12 (--gate: a tile runs on a chunk only near such a trigger of matching chirp mass). One warmup
    and one timed run each; reports step time, template*s per GPU-s over each form's own analysed span (the gated
    forms: over the span they would cover ungated) and the loudest trigger of each dataset."""
    import numpy as np
    import torch
    import fp_filter, window_filter, workloads
    b = wl.banks[0]
    sel = np.nonzero(b.tau0(15.0) < 6.0)[0]
    ds = workloads.condition_streams(wl.streams, wl.cond, [x.clone() for x in wl.x_dev])
    res = dict(templates=int(len(sel)))
    forms = [('whole', None, 4)] + [(f'chunks_b{b_}', '16,%g,2' % hist_s, b_) for b_ in chunk_batch] + \
        [('chunks_512hz', '16,%g,2,512' % hist_s, chunk_batch[-1]),
         ('chunks_gated', '16,%g,2' % hist_s, chunk_batch[-1]), ('chunks_512hz_gated', '16,%g,2,512' % hist_s,
                                                                 chunk_batch[-1])]
    gate = None
    mc_all = (b.m1 * b.m2) ** 0.6 / (b.m1 + b.m2) ** 0.2
    for name, spec, nb in forms:
        d = ds if spec is None else window_filter.to_chunks(ds, spec)[0]
        opts = fp_filter.Options(tile=tile, precision='bf16', gemm_cfg=cfg, self_check=False, segment_batch=nb,
                                 gate=gate if name.endswith('gated') else None)
        IPall, sw, isw = fp_filter.svd_metric(d)
        plan = fp_filter.BankPlan(b.m1[sel], b.m2[sel], None, d[0].flen, d[0].df, opts, sw, isw)
        fp_filter.Filter(d, opts).run(plan)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        outs = [t.to_dict(gates=x.gates) for t, x in zip(fp_filter.Filter(d, opts).run(plan), d)]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        an = workloads.analysed_seconds(d, int(d[0].sr))
        loud = []
        for o in outs:
            i = int(np.argmax(o['snr'])) if len(o['snr']) else None
            loud.append(None if i is None else dict(snr=float(o['snr'][i]), end_time=float(o['end_time'][i]),
                                                     template=int(sel[o['template_id'][i]])))
        if name == 'whole':
            # the gate of the gated forms: this search's triggers above SNR 12 (fp_filter.py:146-149)
            g = [(x.tag.split('-')[0], o['end_time'][k], mc_all[sel[o['template_id'][k]]])
                 for x, o in zip(d, outs) for k in np.nonzero(o['snr'] > 12)[0]]
            gate = fp_filter.Gate(np.array([x[0] for x in g]), np.array([x[1] for x in g]),
                                  np.array([x[2] for x in g]), mc_ratio=1.25, window=1.0)
            res['gate_triggers'] = len(g)
        res[name] = dict(spec=spec, segment_batch=nb, ms=dt * 1e3, analysed_s=an, value=len(sel) * an / dt,
                         segments=sum(len(x.analyze) for x in d), tlen=d[0].tlen, ranks=plan.ranks,
                         n_triggers=int(sum(len(o['snr']) for o in outs)), loudest=loud)
    return res


def consolidation_report(step, nsteps, work, outs_headline):
    """The production cuts (SNR threshold 4, exact band from 5.5, newsnr >= 4, the 30 loudest per 1.072 s window,
    exempt from SNR 5, consolidated after every tile: fp_filter.py:629-663; the sine-Gaussian chi-squared of exact
    peaks with newsnr >= 6 of templates with mtotal > 30, 8 tiles: ops/run_prod_o4a.sh:31) on the same step, with the
    consolidation on the host only and with the device pre-consolidation of the candidates (fp_filter.py:665-695).
    Reports step times, candidate counts, the pre-consolidation and sine-Gaussian kernels' rates and whether the two
    trigger lists are identical."""
    import numpy as np
    import torch
    import _dev
    res = {}
    outs = {}
    for name in ('cons', 'pre'):
        step(name)
        _dev.METER.reset(); _dev.METER.on = True
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(nsteps):
            _, fs, outs[name] = step(name)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / nsteps
        _dev.METER.on = False
        st = _dev.METER.stats()
        r = dict(ms_per_step=dt * 1e3, value=work / dt,
                 n_candidates=int(sum(f.T['n_candidates'] for f in fs)),
                 n_exact_candidates=int(sum(f.T['n_exact_candidates'] for f in fs)),
                 n_triggers=int(sum(len(o['template_id']) for ob in outs[name] for o in ob)),
                 consolidate_host_s=float(sum(f.T['consolidate_s'] for f in fs)),
                 n_sg_chisq_points=int(sum(f.T['n_sg_chisq_points'] for f in fs)),
                 sg_chisq_host_s=float(sum(f.T['sg_chisq_s'] for f in fs)))
        if 'pre_keep' in st:
            n_, t_, by = st['pre_keep']
            # latency-bound: five dependent launches over small arrays
            r['pre_keep_kernels'] = dict(achieved=by / t_ / 1e9, unit='GB/s', bound='latency', frac=None,
                                         launches=n_, avg_launch_ms=t_ / n_ * 1e3, bytes_per_launch=by / n_,
                                         share_of_step=t_ / nsteps / dt)
        if 'sg_tiles' in st:
            n_, t_, by = st['sg_tiles']
            # VALU-bound: its bytes are served from L2 / MALL, so GB/s over 8 TB/s would be no HBM fraction
            r['sg_tiles_kernel'] = dict(achieved=by / t_ / 1e9, unit='GB/s', bound='valu', frac=None,
                                        launches=n_, avg_launch_ms=t_ / n_ * 1e3, bytes_per_launch=by / n_,
                                        share_of_step=t_ / nsteps / dt)
        res[name] = r
    same = all(all(np.array_equal(x[k], y[k]) for k in x) for a_, b_ in zip(outs['cons'], outs['pre'])
               for x, y in zip(a_, b_))
    res['identical_triggers'] = bool(same)
    res['speedup_pre_vs_host_only'] = res['cons']['ms_per_step'] / res['pre']['ms_per_step']
    res['n_triggers_headline_cuts'] = int(sum(len(o['template_id']) for ob in outs_headline for o in ob))
    return res


def _merge(dicts):
    import numpy as np
    keys = dicts[0].keys()
    return {k: np.concatenate([d[k] for d in dicts]) for k in keys}


# what bounds each metered kernel (the counters' reading, profiles/hbm_traffic_batched_pmc.md): only for the
# HBM-bound ones is algorithmic bytes / time over 8 TB/s a roofline fraction; the exact stage's and edges' short
# launches are bound by their dependent chains, the sine-Gaussian tiles by the VALU (their D and IP stay in L2 / MALL),
# the second list pass by its few workgroups
KERNEL_BOUND = {'basis_series': 'hbm', 'sigmasq': 'hbm', 'edges': 'latency', 'chisq': 'latency', 'pre_keep': 'latency',
                'sg_tiles': 'valu', 'second_pass': 'latency'}


def _kernel_rates(ms, nsteps, dt):
    """Metered kernels' rates: name -> achieved GB/s on their algorithmic bytes, with the bound; frac (of 8 TB/s)
    only for the HBM-bound ones."""
    out = {}
    for name, (n_, t_, by) in ms.items():
        if not name.startswith('gemm'):
            bound = KERNEL_BOUND.get(name, 'hbm')
            out[name] = dict(achieved=by / t_ / 1e9, peak=8000.0, unit='GB/s', bound=bound,
                             frac=by / t_ / 1e9 / 8000.0 if bound == 'hbm' else None,
                             launches=n_, avg_launch_ms=t_ / n_ * 1e3, bytes_per_launch=by / n_,
                             share_of_step=t_ / nsteps / dt)
    return out


def _diff_totals(diffs):
    keys = ('n_a', 'n_b', 'only_a', 'only_b', 'windows_only_a', 'windows_only_b')
    tot = {k: int(sum(d[k] for d in diffs)) for k in keys}
    tot['same_sample_min'] = min(d['same_sample'] for d in diffs) if diffs else None
    return tot


def run_batched(a, opts, cfg, tuned, t_setup):
    """The headline: the batched configuration, one slice per step, BF16; FP32 and FP8 forms timed on the first
    a.fp32_slices slices; trigger lists compared; exact SNRs checked against the float64 direct filter."""
    import numpy as np
    import torch
    import _dev
    wl = BatchedWorkload(a.chunks, opts, nslice=a.slices, grids=[int(g) for g in a.grids.split(',')])
    torch.cuda.synchronize()
    # the search's set-up: autotuner, bank products (templates, SVD) and the SVD metric's conditioning; the
    # synthetic data's generation is not counted
    setup_s = time.perf_counter() - t_setup - wl.setup['data_s']
    log('setup', round(setup_s, 1), 's', {k: round(v, 2) for k, v in wl.setup.items()}, '; ranks', wl.ranks())
    ntemp = sum(len(b) for b in wl.banks)
    log('warmup')
    for _ in range(a.warmup):
        wl.step('bf16', None, check=True)
    meter = _dev.METER
    meter.reset(); meter.on = True
    torch.cuda.synchronize()
    outs_bf = {}
    t0 = time.perf_counter()
    res_steps = wl.run('bf16', list(range(a.steps)))
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    meter.on = False
    ms = meter.stats()
    an_s = sum(wl.search.analysed_seconds(ds) for ds, _, _ in res_steps)
    for k, (_, _, outs) in enumerate(res_steps):
        if k < a.fp32_slices:
            outs_bf[k] = outs
    log(f'{a.steps} steps in {total:.3f} s')
    del res_steps
    dt = total / a.steps
    value = ntemp * an_s / total
    n_l, t_k, fl = ms['gemm_bf16']
    nz, zb = meter.notes.get('gemm_bf16_zt_bytes', (0, 0.0))
    res = dict(value=value, ms_per_step=dt * 1e3, analysed_s=an_s, setup_s=setup_s, ms=ms, n_launches=n_l,
               zt_bytes_per_launch=zb / max(nz, 1),
               t_gemm=t_k, flops=fl, hbm=_kernel_rates(ms, a.steps, dt), ntemp=ntemp, wl=wl)
    F = min(a.fp32_slices, a.steps)
    forms = {}
    t_bf = None
    for prec in ('bf16', 'fp32', 'fp8'):
        if F <= 0 or (prec == 'fp8' and not a.fp8_steps):
            continue
        log(f'{prec} form on {F} slices')
        wl.step(prec, None)
        meter.reset(); meter.on = True
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        rs = wl.run(prec, list(range(F)))
        torch.cuda.synchronize()
        t_f = time.perf_counter() - t1
        meter.on = False
        if prec == 'bf16':                       # the same slices in the BF16 form, timed alike
            t_bf = t_f
            continue
        outs_f = {k: rs[k][2] for k in range(F)}
        st = meter.stats()
        n_, t_, f_ = st['gemm_' + prec]
        err_f = nchk = None
        if not a.no_check:
            err_f, nchk = wl.direct_check(rs[0][0], outs_f[0])
        forms[prec] = dict(seconds=t_f, bf16_seconds=t_bf, speedup_bf16=t_f / t_bf,
                           gemm_tflops=f_ / t_ / 1e12, gemm_frac=f_ / t_ / 1e12 / PEAK[prec],
                           gemm_avg_launch_ms=t_ / max(n_, 1) * 1e3, snr_rel_err_vs_fp64=err_f, n_checked=nchk,
                           outs=outs_f)
    err_bf = nchk_bf = None
    if not a.no_check and F > 0:
        ds0, _, o0 = wl.run('bf16', [0], check=True)[0]
        err_bf, nchk_bf = wl.direct_check(ds0, o0)
    res.update(forms=forms, err_bf=err_bf, nchk_bf=nchk_bf, outs_bf=outs_bf, F=F)
    return res


def run_production(a, opts, cfg):
    """The production workload (2 detectors x 512 s, lines, glitches, 20 injections; both banks), reported beside
    the headline: BF16 / FP32 / FP8 step times, the production cuts with consolidation, the overlap-save chunked forms,
    the refinement of the loudest triggers and time slides."""
    import torch
    import _dev, workloads
    t0 = time.perf_counter()
    wl = ProductionWorkload(opts)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    wl.step('bf16', check=True)
    ntemp = sum(len(b) for b in wl.banks)
    meter = _dev.METER
    res = dict(workload=wl.describe(a.tile), setup_s=setup, ranks=wl.ranks())
    outs = {}
    for prec, nst in (('bf16', a.production_steps), ('fp32', 1), ('fp8', a.fp8_steps)):
        if nst <= 0:
            continue
        wl.step(prec)
        meter.reset(); meter.on = True
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(nst):
            ds, _, outs[prec] = wl.step(prec)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t1) / nst
        meter.on = False
        st = meter.stats()
        n_, t_, f_ = st['gemm_' + prec]
        an_s = workloads.analysed_seconds(ds, 2048)
        res[prec] = dict(value=ntemp * an_s / dt, ms_per_step=dt * 1e3, steps=nst, gemm_tflops=f_ / t_ / 1e12,
                         gemm_frac=f_ / t_ / 1e12 / PEAK[prec], gemm_avg_launch_ms=t_ / max(n_, 1) * 1e3,
                         gemm_share_of_step=t_ / nst / dt)
        if prec == 'bf16':
            res['hbm_kernels'] = _kernel_rates(st, nst, dt)
        if not a.no_check:
            res[prec]['snr_rel_err_vs_fp64'], res[prec]['n_checked'] = wl.direct_check(ds, outs[prec])
    if 'fp32' in outs:
        res['speedup_bf16_vs_fp32'] = res['fp32']['ms_per_step'] / res['bf16']['ms_per_step']
        res['trigger_diff_bf16_vs_fp32'] = wl.compare(outs['bf16'], outs['fp32'])
        if 'fp8' in outs:
            res['trigger_diff_fp8_vs_fp32'] = wl.compare(outs['fp8'], outs['fp32'])
        # the reference's rule: exact values at each form's rebuilt argmax
        r0 = {p: wl.step(p + '_r0')[2] for p in ('bf16', 'fp32')}
        res['trigger_diff_bf16_vs_fp32_refine0'] = wl.compare(r0['bf16'], r0['fp32'])
    if a.consolidate_steps > 0:
        log('production cuts with consolidation')
        res['consolidation'] = consolidation_report(wl.step, a.consolidate_steps,
                                                    ntemp * workloads.analysed_seconds(ds, 2048), outs['bf16'])
    if a.chunked > 0:
        log('overlap-save chunks')
        res['chunked'] = chunked_report(wl, cfg, a.tile)
    if a.refine_top > 0:
        log('refinement of the loudest triggers')
        res['refinement'] = refinement_report(wl.plans, wl.streams, ds, outs['bf16'], a.refine_top)
    if a.slides > 0:
        log('time slides')
        res['time_slides'] = slides_report(ds, outs['bf16'], [b.name for b in wl.banks], K=a.slides)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20, help='timed steps; one step filters one slice of the batched '
                    'workload, so --slices steps filter every chunk once')
    ap.add_argument('--warmup', type=int, default=2, help='untimed steps on the warm chunks')
    ap.add_argument('--chunks', type=int, default=200, help='chunks per detector of the batched workload')
    ap.add_argument('--slices', type=int, default=20, help='slices of the batched workload (one per step)')
    ap.add_argument('--grids', default='512,256,128', help='segment lengths (s) of the batched workload\'s grids: a '
                    'chunk takes the longest that leaves it a full segment beyond the 64 s of pads')
    ap.add_argument('--fp32-slices', type=int, default=2, help='slices timed in the FP32 (and FP8) forms and compared '
                    'with the BF16 form')
    ap.add_argument('--fp8-steps', type=int, default=1, help='0: no FP8 form')
    ap.add_argument('--fp8-margin', type=float, default=0.08, help='candidate margin of the FP8 form')
    ap.add_argument('--fp8-alt-margin', type=float, default=0.04, help='alternates\' margin of the FP8 form')
    ap.add_argument('--tile', type=int, default=512)
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline leg')
    ap.add_argument('--no-check', action='store_true', help='skip the float64 direct-filter SNR check')
    ap.add_argument('--gemm-cfg', default=None, help='tiles_per_block[,variant]')
    ap.add_argument('--refine', type=int, default=3, help='exact-stage re-centring half width in samples (0: reference rule)')
    ap.add_argument('--alternates', type=int, default=3, help='with --refine: other local maxima per candidate window '
                    'evaluated exactly (0 to 3)')
    ap.add_argument('--templates', choices=['closed', 'interp'], default='closed',
                    help='interp: the chirps at coarse nodes, interpolated (the reference\'s --templates interp)')
    ap.add_argument('--exact-stage', choices=['dense', 'regen'], default='dense',
                    help='regen: the exact stage regenerates the interpolated templates instead of reading H')
    ap.add_argument('--alt-second-pass', type=float, default=None, help='with --alternates: local-maximum lists in '
                    'a second GEMM pass over the windows with exact-band candidates while those are under this '
                    'fraction of a batch\'s windows (0: always in the first pass; default: by precision, '
                    'fp_filter.SECOND_PASS_DEFAULT)')
    ap.add_argument('--segment-batch', type=int, default=32, help='at most this many segments per basis-transform '
                    'batch and GEMM launch')
    ap.add_argument('--batch-samples', type=int, default=1 << 22, help='and at most this many analysed samples')
    ap.add_argument('--overlap', type=int, default=0, help='1: basis transforms on their own stream beside the GEMM')
    ap.add_argument('--depth', type=int, default=2, help='batches enqueued ahead of the one being decoded')
    ap.add_argument('--production', type=int, default=1, help='also run the production workload (2 detectors x 512 s '
                    'with lines, glitches and injections) and its reports; 0: skip')
    ap.add_argument('--production-steps', type=int, default=3, help='timed BF16 steps of the production workload')
    ap.add_argument('--refine-top', type=int, default=10000, help='refine this many of the loudest exact triggers '
                    'off the bank (production workload; 0: skip)')
    ap.add_argument('--consolidate-steps', type=int, default=1,
                    help='timed steps of the production cuts (SNR threshold 4, exact band from 5.5, keep the 30 '
                    'loudest per 1.072 s, exempt from SNR 5, sine-Gaussian chi-squared from newsnr 6) with host '
                    'consolidation only and with the device pre-consolidation (production workload; 0: skip)')
    ap.add_argument('--chunked', type=int, default=1, help='run the overlap-save chunked forms beside whole segments '
                    '(production workload; 0: skip)')
    ap.add_argument('--slides', type=int, default=500, help='time slides |k| <= K over the two detectors\' '
                    'trigger lists (production workload; 0: skip)')
    a = ap.parse_args()
    if a.gpus != 1:
        print(f'bench.py runs on one GPU: there is no multi-GPU path here (asked for --gpus {a.gpus}); the '
              f'filtering work shards by (detector, chunk) or by template tile', file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import _lib
    if not os.path.exists(_lib.LIB_PATH):
        _lib.build()
    import fp_filter
    assert torch.cuda.is_available(), 'bench.py needs a GPU'
    t_setup = time.perf_counter()
    import fp_gemm
    tuned = None
    t_tune = time.perf_counter()
    if a.gemm_cfg:
        cfg = tuple(int(v) for v in a.gemm_cfg.split(','))
    else:
        # the autotuner on the filter's shape: a 512-template tile, K = 2 x 416, two 448 s segments of windows, in
        # the mode the BF16 form's first pass runs (with the local-maximum lists unless a second pass makes them)
        sp = fp_filter.SECOND_PASS_DEFAULT['bf16'] if a.alt_second_pass is None else a.alt_second_pass
        cfg, tuned = fp_gemm.pick_cfg(M=512, K=832, Tn=2 * 448 * 2048, dtype=torch.bfloat16,
                                      lists=a.refine > 0 and a.alternates > 0 and sp == 0)
        log('autotuned GEMM launch', cfg, {k: (round(v * 1e3, 3) if isinstance(v, float) else v)
                                          for k, v in tuned.items()})
    t_tune = time.perf_counter() - t_tune
    opts = {p: fp_filter.Options(tile=a.tile, precision=p, gemm_cfg=cfg, self_check=False, refine=a.refine,
                                 alternates=a.alternates, segment_batch=a.segment_batch,
                                 batch_samples=a.batch_samples, overlap=bool(a.overlap),
                                 alt_second_pass=a.alt_second_pass,
                                 depth=a.depth, templates=a.templates, exact_stage=a.exact_stage)
            for p in ('bf16', 'fp32', 'fp8')}
    # e4m3's 3-bit mantissa moves rebuilt values by a few percent: the fp8 form widens the candidate margin (the
    # reference's --margin, whose role is to absorb the compressed filter's error) and the alternates' margin
    opts['fp8'] = fp_filter.Options(**{**opts['fp8'].__dict__, 'margin': a.fp8_margin,
                                       'alt_margin': a.fp8_alt_margin})
    # the reference's rule (exact values at the rebuilt argmax), for the trigger comparison beside the refined one
    for p in ('bf16', 'fp32'):
        opts[p + '_r0'] = fp_filter.Options(**{**opts[p].__dict__, 'refine': 0})
    import consolidate
    kl = consolidate.KeepLoudest(newsnr_threshold=4.0, interval=1.072, num=30, exempt_snr=5.0)
    for name, pre in (('cons', False), ('pre', True)):
        opts[name] = fp_filter.Options(tile=a.tile, precision='bf16', gemm_cfg=cfg, self_check=False,
                                       refine=a.refine, alternates=a.alternates, snr_threshold=4.0, exact_snr=5.5,
                                       templates=a.templates, exact_stage=a.exact_stage, consolidate=kl,
                                       pre_consolidate=pre, sgchisq_snr_threshold=6.0)
    b = run_batched(a, opts, cfg, tuned, t_setup)
    wl = b['wl']
    F = b['F']
    diffs = {}
    for prec, f in b['forms'].items():
        diffs[prec] = [wl.compare(b['outs_bf'][k] if prec == 'fp32' else f['outs'][k],
                                  b['forms']['fp32']['outs'][k]) for k in range(F)] if 'fp32' in b['forms'] else None
    dflat = {p: [x for per in d for x in per] for p, d in diffs.items() if d is not None}
    fp32 = b['forms'].get('fp32')
    fp8 = b['forms'].get('fp8')
    err_max = max([e for e in (b['err_bf'],) if e is not None], default=None)
    # the whole job (every slice once) with the bank set-up (templates, SVD, autotune) counted: per-slice time
    # from the timed steps, scaled to all slices
    slices_s = b['ms_per_step'] / 1e3 * a.slices
    an_job = b['analysed_s'] * a.slices / a.steps
    res = dict(
        metric=METRIC, value=b['value'], unit='template*s/GPU-s', n_gpus=1, steps=a.steps, warmup=a.warmup,
        ms_per_step=b['ms_per_step'], higher_is_better=True, scaling='weak', vs_baseline=None, dtype='bf16',
        data='synthetic (noise drawn on the GPU, seeded)',
        config=dict(workload=wl.describe(a.tile), templates=b['ntemp'], analysed_s=b['analysed_s'],
                    grids=list(wl.grids),
                    slices=a.slices, tile=a.tile, segment_batch=a.segment_batch, batch_samples=a.batch_samples,
                    ranks=wl.ranks(),
                    setup_s=round(b['setup_s'], 1),
                    setup_parts={k: round(v, 2) for k, v in dict(b['wl'].setup, autotune_s=t_tune).items()},
                    value_with_setup=b['ntemp'] * an_job / (slices_s + b['setup_s']),
                    speedup_bf16_vs_fp32=None if fp32 is None else fp32['speedup_bf16'],
                    speedup_fp32_slices=F,
                    snr_rel_err_vs_fp64=err_max, snr_rel_err_n_checked=b['nchk_bf'],
                    snr_rel_err_vs_fp64_fp32=None if fp32 is None else fp32['snr_rel_err_vs_fp64'],
                    trigger_diff_bf16_vs_fp32=_diff_totals(dflat['fp32']) if 'fp32' in dflat else None,
                    gemm_cfg=list(cfg) if cfg else None,
                    autotune_ms={k: round(v * 1e3, 3) for k, v in (tuned or {}).items() if isinstance(v, float)}),
        roofline=dict(bound='mfma', achieved=b['flops'] / b['t_gemm'] / 1e12, peak=PEAK['bf16'], unit='TFLOP/s',
                      frac=b['flops'] / b['t_gemm'] / 1e12 / PEAK['bf16'], traffic=measured_traffic(),
                      kernel='gemm_peaks (bf16, fused |z|^2 + window max)', launches=b['n_launches'],
                      algorithmic_bytes_per_launch=b['zt_bytes_per_launch'],
                      avg_launch_ms=b['t_gemm'] / max(b['n_launches'], 1) * 1e3,
                      gemm_share_of_step=b['t_gemm'] / a.steps / (b['ms_per_step'] / 1e3)),
        hbm_kernels=b['hbm'],
        fp32=None if fp32 is None else {k: v for k, v in fp32.items() if k != 'outs'},
        fp8=None if fp8 is None else dict({k: v for k, v in fp8.items() if k != 'outs'},
                                          trigger_diff_fp8_vs_fp32=_diff_totals(dflat['fp8']) if 'fp8' in dflat
                                          else None),
        trigger_diff_bf16_vs_fp32=dflat.get('fp32'),
    )
    if a.production > 0:
        log('production workload')
        res['production'] = pr = run_production(a, opts, cfg)
        # the target's fields for the production workload, in config where the record keeps them
        res['config']['production'] = dict(
            bf16_ms_per_step=pr['bf16']['ms_per_step'], bf16_value=pr['bf16']['value'],
            speedup_bf16_vs_fp32=pr.get('speedup_bf16_vs_fp32'),
            snr_rel_err_vs_fp64=pr['bf16'].get('snr_rel_err_vs_fp64'),
            trigger_diff_bf16_vs_fp32=_diff_totals(pr['trigger_diff_bf16_vs_fp32'])
            if 'trigger_diff_bf16_vs_fp32' in pr else None,
            trigger_diff_bf16_vs_fp32_refine0=_diff_totals(pr['trigger_diff_bf16_vs_fp32_refine0'])
            if 'trigger_diff_bf16_vs_fp32_refine0' in pr else None)
    if not a.no_cpu:
        log('cpu baseline')
        import workloads
        # one chunk of the first slice on the 512 s grid, H1
        st, cond, x = next((st, c, x) for i in wl.slices[0] for st, c, x in [wl.search.chunks[i]]
                           if c is not None and c.seg_len == 512 and st.ifo == 'H1')
        res['cpu_baseline'] = cpu_baseline(x.cpu().numpy(), st.sr, st.gps_start, cond, wl.banks[0],
                                           f'H1 chunk of {x.numel() / st.sr:.0f} s from the first slice')
    else:
        res['cpu_baseline'] = None
    print(json.dumps(res, default=float), flush=True)


if __name__ == '__main__':
    main()
