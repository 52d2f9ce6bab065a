#!/usr/bin/env python
"""Throughput of the filtering path on one MI355X: template-seconds filtered per GPU-second.

Workload (workloads.production): 2 detectors x 512 s at 2048 Hz of synthetic coloured noise with lines, glitches and
20 injections; both banks, 4,612 circular + 1,268 eccentric = 5,880 templates. One step = one pass of the hot path
over the batch with the raw strain resident in HBM: autogating, conditioning (float64 median-Welch PSD, truncation,
whitening), template norms, per tile the basis transforms and the fused GEMM + window max, candidate selection, the
rebuilt phases and the exact stage (edges, chi-squared sums, float64 finalization). The bank products (templates,
float64 SVD bases) are built once before timing, as a search builds them once per run.

value = 5,880 x analysed seconds (both detectors) / seconds per step, for the BF16 form. The FP32 plain form of the
same step is timed beside it, the trigger lists of the two are compared, and every exact trigger's SNR is compared
with a float64 direct matched filter of the same conditioned data (computed with torch complex128 FFTs).

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
    """HBM bytes per GEMM launch from the committed rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, scaled by a
    plain copy of known size), or None when the summary is absent."""
    import re
    p = os.path.join(ROOT, 'profiles', 'hbm_traffic_c2_pmc.md')
    if not os.path.exists(p):
        return None
    m = re.search(r'traffic_bytes_per_launch \(gemm_peaks, bf16\) = ([0-9.eE+]+)', open(p).read())
    return float(m.group(1)) if m else None


def cpu_baseline(streams, banks, cond, ntemp=512):
    """The oracle's plain method (numpy / scipy.fft / OpenBLAS float32) on a bounded sample: H1 only, the first
    `ntemp` circular templates in chirp-time order as one tile; conditioning plus the per-segment work (basis
    transforms, GEMM and window maxima, exact stage) is timed, the tile's templates and SVD are not (as on the GPU)."""
    import numpy as np
    from oracle import core, pipeline
    cores = min(len(os.sched_getaffinity(0)), int(os.environ.get('OMP_NUM_THREADS', '16')))
    st = streams[0]
    bank = banks[0]
    ids = pipeline.tiles_by_tau0(bank.m1, bank.m2, 15.0, ntemp)[0]
    t0 = time.perf_counter()
    d = core.condition(st.x, st.sr, st.gps_start, cond.seg_len, cond.start_pad, cond.end_pad, cond.psd_seg_len,
                       cond.psd_stride, cond.psd_nseg, cond.psd_inv_len, 15.0)
    t_cond = time.perf_counter() - t0
    T = {}
    pipeline.filter_plain([d], bank.m1[ids], bank.m2[ids], None, tile=ntemp, dtype=np.float32, workers=cores,
                          timing=T)
    t = t_cond + T['basis_fft_s'] + T['gemm_s'] + T['exact_s']
    an = sum((a1 - a0) // 2048 * 2048 / st.sr for a0, a1 in d['analyze'])
    return dict(value=ntemp * an / t, unit='template*s/CPU-s', cores=cores, kind='port',
                sample=f'H1 512 s, {ntemp} circular templates as one tile, float32 plain method '
                       f'(oracle/pipeline.py), conditioning + per-segment work timed ({t:.1f} s), '
                       f'templates/SVD excluded ({T["templates_s"] + T["svd_s"]:.1f} s)')


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
    """Uneven chunks (64-1024 s in 16 s steps, 2 detectors, noise with lines), both banks, grouped by segment grid
    (batched.BatchedSearch). Datasets and trigger lists are flattened over grids in grid order."""

    def __init__(self, n_chunks, opts):
        import torch
        import batched, workloads
        chunks, self.banks = workloads.batched(n_chunks)
        x_dev = [torch.from_numpy(st.x).cuda() for st, _ in chunks]
        self.n_chunks = n_chunks
        self.search = batched.BatchedSearch(chunks, self.banks, x_dev, opts)
        self.plans = None

    def step(self, prec, check=False):
        ds, outs, filters = self.search.step(prec, self_check=check)
        g = self.search.grids
        self._last_ds = ds
        flat_ds = [d for k in g for d in ds[k]]
        flat = [[t for k in g for t in outs[k][b]] for b in range(len(self.banks))]
        return flat_ds, filters, flat

    def ranks(self):
        return self.search.ranks()

    def describe(self, tile):
        return ('batched: %d chunks per detector, 64-1024 s in 16 s steps, 2 detectors at 2048 Hz (noise with lines), '
                'grouped on 512 s and 128 s segment grids; 4,612 circular + 1,268 eccentric SPA templates, tiles of '
                '%d, SVD worst lost power 1e-4, 16 chi-squared bins, SNR threshold 5.5, 1 s windows; BF16 GEMM, FP32 '
                'exact stage' % (self.n_chunks, tile))

    def compare(self, a, b):
        # per bank, the lists of every dataset merged (keys are unique: each dataset has its own end times)
        out = []
        for i in range(len(a)):
            out.append(compare_triggers(_merge(a[i]), _merge(b[i])))
        return out

    def direct_check(self, ds, outs):
        # the float64 direct filter on the datasets of the first grid (the flat lists start with them; that grid's
        # plans hold the dense templates on its frequency bins)
        g = self.search.grids[0]
        nd = len(self._last_ds[g])
        part = [outs[b][:nd] for b in range(len(outs))]
        return direct_snr_check(self._last_ds[g], self.search.plans[g], part, max_templates=12)


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
            r['pre_keep_kernels'] = dict(achieved=by / t_ / 1e9, peak=8000.0, unit='GB/s', frac=by / t_ / 8e12,
                                         launches=n_, avg_launch_ms=t_ / n_ * 1e3, bytes_per_launch=by / n_,
                                         share_of_step=t_ / nsteps / dt)
        if 'sg_tiles' in st:
            n_, t_, by = st['sg_tiles']
            r['sg_tiles_kernel'] = dict(achieved=by / t_ / 1e9, peak=8000.0, unit='GB/s', frac=by / t_ / 8e12,
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--fp32-steps', type=int, default=1)
    ap.add_argument('--fp8-steps', type=int, default=1, help='timed steps of the FP8 form (production workload)')
    ap.add_argument('--fp8-margin', type=float, default=0.04, help='candidate margin of the FP8 form')
    ap.add_argument('--fp8-refine', type=int, default=3, help='exact-stage re-centring half width of the FP8 form')
    ap.add_argument('--tile', type=int, default=512)
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline leg')
    ap.add_argument('--no-check', action='store_true', help='skip the float64 direct-filter SNR check')
    ap.add_argument('--gemm-cfg', default=None, help='tiles_per_block[,variant]')
    ap.add_argument('--refine', type=int, default=3, help='exact-stage re-centring half width in samples (0: reference rule)')
    ap.add_argument('--timing', action='store_true', help='one extra step with stage timing (synchronised) to stderr')
    ap.add_argument('--workload', choices=('production', 'batched'), default='production',
                    help='production: 2 detectors x 512 s with lines, glitches and injections; batched: uneven chunks')
    ap.add_argument('--chunks', type=int, default=200, help='chunks per detector of the batched workload')
    ap.add_argument('--graphs', action='store_true', help='capture stage A of every batch in HIP graphs (replayed '
                    'from the second step on; production workload)')
    ap.add_argument('--refine-top', type=int, default=10000, help='refine this many of the loudest exact triggers '
                    'off the bank after the timed steps (production workload; 0: skip)')
    ap.add_argument('--consolidate-steps', type=int, default=1,
                    help='timed steps of the production cuts (SNR threshold 4, exact band from 5.5, keep the 30 '
                    'loudest per 1.072 s, exempt from SNR 5, sine-Gaussian chi-squared from newsnr 6) with host '
                    'consolidation only and with the device pre-consolidation (production workload; 0: skip)')
    ap.add_argument('--chunked', type=int, default=1, help='run the overlap-save chunked forms beside whole segments '
                    '(production workload; 0: skip)')
    ap.add_argument('--slides', type=int, default=500, help='time slides |k| <= K over the two detectors\' '
                    'trigger lists after the timed steps (production workload; 0: skip)')
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
    import _dev, fp_filter, workloads
    assert torch.cuda.is_available(), 'bench.py needs a GPU'
    t_setup = time.perf_counter()
    import fp_gemm
    tuned = None
    if a.gemm_cfg:
        cfg = tuple(int(v) for v in a.gemm_cfg.split(','))
    else:
        # the autotuner on the filter's shape: a 512-template tile, K = 2 x 416, two 448 s segments of windows
        cfg, tuned = fp_gemm.pick_cfg(M=512, K=832, Tn=2 * 448 * 2048, dtype=torch.bfloat16)
        log('autotuned GEMM launch', cfg, {k: (round(v * 1e3, 3) if isinstance(v, float) else v)
                                          for k, v in tuned.items()})
    opts = {p: fp_filter.Options(tile=a.tile, precision=p, gemm_cfg=cfg, self_check=False,
                                 refine=a.refine, graphs=a.graphs) for p in ('bf16', 'fp32', 'fp8')}
    opts['fp8'].margin, opts['fp8'].refine = a.fp8_margin, a.fp8_refine
    import consolidate
    kl = consolidate.KeepLoudest(newsnr_threshold=4.0, interval=1.072, num=30, exempt_snr=5.0)
    for name, pre in (('cons', False), ('pre', True)):
        opts[name] = fp_filter.Options(tile=a.tile, precision='bf16', gemm_cfg=cfg, self_check=False,
                                       refine=a.refine, snr_threshold=4.0, exact_snr=5.5, consolidate=kl,
                                       pre_consolidate=pre, sgchisq_snr_threshold=6.0)
    if a.workload == 'batched':
        wl = BatchedWorkload(a.chunks, opts)
    else:
        wl = ProductionWorkload(opts)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t_setup
    log('setup', round(setup_s, 1), 's; ranks', wl.ranks())
    banks = wl.banks
    cw = 2048
    step = wl.step

    log('warmup')
    for _ in range(a.warmup):
        ds, _, _ = step('bf16', check=True)
    if a.timing:
        for p in ('bf16', 'fp32'):
            opts[p].timing = True
            t0 = time.perf_counter()
            _, fs, _ = step(p)
            torch.cuda.synchronize()
            log(f'timing {p}: step {time.perf_counter() - t0:.3f} s', json.dumps(
                {k: round(sum(f.T[k] for f in fs), 4) for k in fs[0].T}))
            opts[p].timing = False
    an_s = workloads.analysed_seconds(ds, cw)
    ntemp = sum(len(b) for b in banks)
    plans = wl.plans
    meter = _dev.METER
    meter.reset(); meter.on = True
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(a.steps):
        ds, filters, outs_bf = step('bf16')
        log(f'step {k + 1}/{a.steps}')
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / a.steps
    ms = meter.stats()
    value = ntemp * an_s / dt
    meter.on = False
    def metered_plain(prec):
        # kernels inside replayed graphs carry no events: one more step of the same work without graphs, metered
        opts[prec].graphs = False
        meter.reset(); meter.on = True
        step(prec)
        st = meter.stats()
        meter.on = False
        opts[prec].graphs = a.graphs
        return st

    if a.graphs:
        ms = metered_plain('bf16')
    n_l, t_k, fl = ms['gemm_bf16']
    dt32 = outs_fp = None
    if a.fp32_steps > 0:
        log('fp32 form')
        step('fp32')
        meter.reset(); meter.on = True
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(a.fp32_steps):
            _, f32f, outs_fp = step('fp32')
        torch.cuda.synchronize()
        dt32 = (time.perf_counter() - t1) / a.fp32_steps
        meter.on = False
        n32, t32, fl32 = (metered_plain('fp32') if a.graphs else meter.stats())['gemm_fp32']
    dt8 = outs_f8 = None
    if a.fp8_steps > 0 and a.workload == 'production':
        # the FP8 e4m3 form of the same step (per-basis and per-template scales), reported beside the headline
        log('fp8 form')
        step('fp8')
        meter.reset(); meter.on = True
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        for _ in range(a.fp8_steps):
            _, _, outs_f8 = step('fp8')
        torch.cuda.synchronize()
        dt8 = (time.perf_counter() - t2) / a.fp8_steps
        meter.on = False
        n8, t8, fl8 = (metered_plain('fp8') if a.graphs else meter.stats())['gemm_fp8']
    hbm = {}
    msteps = 1 if a.graphs else a.steps
    for name, (n_, t_, by) in ms.items():
        if not name.startswith('gemm'):
            hbm[name] = dict(achieved=by / t_ / 1e9, peak=8000.0, unit='GB/s', frac=by / t_ / 1e9 / 8000.0,
                             launches=n_, avg_launch_ms=t_ / n_ * 1e3, bytes_per_launch=by / n_,
                             share_of_step=t_ / msteps / dt)
    diffs = None
    if outs_fp is not None:
        diffs = wl.compare(outs_bf, outs_fp)
    err = err32 = err8 = None
    nchk = 0
    if not a.no_check:
        err, nchk = wl.direct_check(ds, outs_bf)
        if outs_fp is not None:
            err32, _ = wl.direct_check(ds, outs_fp)
        if outs_f8 is not None:
            err8, _ = wl.direct_check(ds, outs_f8)
    res = dict(
        metric=METRIC, value=value, unit='template*s/GPU-s', n_gpus=1, steps=a.steps, warmup=a.warmup,
        ms_per_step=dt * 1e3, higher_is_better=True, scaling='weak', vs_baseline=None, dtype='bf16',
        data='synthetic',
        config=dict(workload=wl.describe(a.tile), templates=ntemp, analysed_s=an_s, tile=a.tile,
                    ranks=wl.ranks(), setup_s=round(setup_s, 1), gemm_cfg=list(cfg) if cfg else None,
                    autotune_ms={k: round(v * 1e3, 3) for k, v in (tuned or {}).items() if isinstance(v, float)}),
        roofline=dict(bound='mfma', achieved=fl / t_k / 1e12, peak=PEAK['bf16'], unit='TFLOP/s',
                      frac=fl / t_k / 1e12 / PEAK['bf16'], traffic=measured_traffic(),
                      kernel='gemm_peaks (bf16, fused |z|^2 + window max)', launches=n_l,
                      avg_launch_ms=t_k / max(n_l, 1) * 1e3, gemm_share_of_step=t_k / msteps / dt),
        fp32=None if dt32 is None else dict(
            value=ntemp * an_s / dt32, ms_per_step=dt32 * 1e3, gemm_tflops=fl32 / t32 / 1e12,
            gemm_frac=fl32 / t32 / 1e12 / PEAK['fp32'], gemm_share_of_step=t32 / dt32),
        hbm_kernels=hbm,
        speedup_bf16_vs_fp32=None if dt32 is None else dt32 / dt,
        snr_rel_err_vs_fp64=dict(bf16=err, fp32=err32, n_checked=nchk),
        trigger_diff_bf16_vs_fp32=diffs,
        fp8=None if dt8 is None else dict(
            value=ntemp * an_s / dt8, ms_per_step=dt8 * 1e3, gemm_tflops=fl8 / t8 / 1e12,
            gemm_frac=fl8 / t8 / 1e12 / PEAK['fp8'], gemm_avg_launch_ms=t8 / max(n8, 1) * 1e3,
            speedup_vs_fp32=None if dt32 is None else dt32 / dt8, snr_rel_err_vs_fp64=err8,
            trigger_diff_fp8_vs_fp32=None if outs_fp is None else wl.compare(outs_f8, outs_fp)),
    )
    if a.workload == 'production' and a.consolidate_steps > 0:
        log('production cuts with consolidation')
        res['consolidation'] = consolidation_report(step, a.consolidate_steps, ntemp * an_s, outs_bf)
    if a.workload == 'production' and a.chunked > 0:
        log('overlap-save chunks')
        res['chunked'] = chunked_report(wl, cfg, a.tile)
    if a.workload == 'production' and a.refine_top > 0:
        log('refinement of the loudest triggers')
        res['refinement'] = refinement_report(wl.plans, wl.streams, ds, outs_bf, a.refine_top)
    if a.workload == 'production' and a.slides > 0:
        log('time slides')
        res['time_slides'] = slides_report(ds, outs_bf, [b.name for b in banks], K=a.slides)
    if not a.no_cpu:
        log('cpu baseline')
        streams, banks_p, cond = workloads.production()
        res['cpu_baseline'] = cpu_baseline(streams, banks_p, cond)
    else:
        res['cpu_baseline'] = None
    print(json.dumps(res), flush=True)


if __name__ == '__main__':
    main()
