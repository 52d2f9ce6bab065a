"""Parity of the filter's remaining bands and configurations, through the C ABI:

- the rebuilt band (fp_filter.py:759-782): candidates between the SNR threshold and the exact band keep the
  compressed filter's SNR and phase, against the oracle's plain method and the float64 direct filter, on a bank the
  SVD compresses (ranks 44-58 of 64 at a worst lost power of 1e-4);
- the pipelined filter against its synchronised form, a second run of the same Filter and a run after rebind;
- the production shape (2 detectors x 512 s with lines, glitches and injections) on one 512-template tile of each
  bank, against the oracle's conditioning and plain method and the float64 direct filter;
- the batched form (uneven chunks on two segment grids, both banks) against one Filter per chunk and the float64
  direct filter;
- gpu_prep.condition against the reference's own gpu_stage2 (golden vectors)."""
import json
import os
import sys

import numpy as np
import pytest
import torch

from oracle import core, pipeline

pytestmark = pytest.mark.gpu

SR = 2048.0
SVD_SNR_BOUND = 0.05     # |rebuilt - exact| SNR: sqrt(eps) = 0.01 times the data's SNR (<= 5) along the residual


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope='module')
def dense_c1(gpu):
    """One detector, 64 s of seeded coloured noise, 192 templates of 20-60 Msun in three tiles of 64 (SVD ranks
    44-58 at eps 1e-4), two injections (SNR 12 and 25)."""
    import gpu_prep, synth
    bank = synth.circular_bank(192, seed=13, mmin=20, mmax=60)
    inj = [synth.Injection(30.0, float(bank.m1[5]), float(bank.m2[5]), 12.0, 0.3),
           synth.Injection(47.0, 33.3, 27.1, 25.0, 1.1)]
    st = synth.make_stream('H1', 64, SR, seed=1, injections=inj)
    ds = gpu_prep.condition(torch.from_numpy(st.x).cuda(), SR, st.gps_start, 64, 8, 8, 4, 2, 31, 4, 15.0, tag='H1')
    return bank, ds, st


def _host(ds):
    return dict(D=ds.D.cpu().numpy(), IP=ds.IP.cpu().numpy(), analyze=ds.analyze, seg_start=ds.seg_start,
                tlen=ds.tlen, df=ds.df, sr=ds.sr, gps_start=ds.gps_start)


def _direct(H, D, IP, df, tlen, tids, samps):
    """float64 direct matched filter (complex128 on the GPU) at (template row of H, sample): SNR and phase."""
    Hd = torch.as_tensor(H).cuda() if not isinstance(H, torch.Tensor) else H
    Dd = torch.as_tensor(D).cuda().to(torch.complex128)
    ipd = torch.as_tensor(IP).cuda().to(torch.float64)
    snr, ph, cache = np.empty(len(tids)), np.empty(len(tids)), {}
    flen = Dd.shape[0]
    for i, (t, s) in enumerate(zip(np.asarray(tids).tolist(), np.asarray(samps).tolist())):
        if t not in cache:
            h = Hd[t].to(torch.complex128)
            sig = 4 * df * float((h.abs() ** 2 * ipd).sum())
            q = torch.zeros(tlen, dtype=torch.complex128, device='cuda')
            q[:flen] = torch.conj(h) * Dd
            cache[t] = (torch.fft.ifft(q, norm='forward') * (4 * df / np.sqrt(sig))).cpu().numpy()
        snr[i], ph[i] = abs(cache[t][s]), np.angle(cache[t][s])
    return snr, ph


@pytest.mark.parametrize('precision', ['fp32', 'bf16'])
def test_rebuilt_band_vs_oracle(dense_c1, precision):
    """SNR threshold 4, exact band from 5.5 (margin 0.04: rebuilt SNR >= 5.28): the same (template, window) list as
    the oracle's plain method on the same conditioned data except within the SVD bound of a cut; rebuilt peaks carry
    the rebuilt SNR and phase (within the SVD bound of the float64 direct filter and, on tiles of equal rank, within
    fp32 / bf16 rounding of the oracle's rebuilt values), exact = False and chisq = dof = 0; exact peaks are within
    1e-4 of the direct filter with dof = 30."""
    import fp_filter
    bank, ds, _ = dense_c1
    thr, esnr, margin = 4.0, 5.5, 0.04
    _say('rebuilt band', precision)
    opts = fp_filter.Options(tile=64, precision=precision, snr_threshold=thr, exact_snr=esnr, margin=margin, eps=1e-4)
    IPall, sw, isw = fp_filter.svd_metric([ds])
    plan = fp_filter.BankPlan(bank.m1, bank.m2, None, ds.flen, ds.df, opts, sw, isw)
    trig = fp_filter.Filter([ds], opts).run(plan)[0].to_dict()
    h = _host(ds)
    swh = sw.cpu().numpy()
    res, _ = pipeline.filter_plain([h], bank.m1, bank.m2, None, thr=thr, margin=margin, exact_snr=esnr, tile=64,
                                   eps=1e-4, sw=swh)
    o = res[0]
    # oracle ranks per tile (same tiles: chirp-time order, 64 per tile)
    H, _ = pipeline.spa_bank(bank.m1, bank.m2, None, ds.flen, ds.df, 15.0)
    H = H.astype(np.complex64)
    kmin = int(15.0 / ds.df)
    same_rank = {}
    for tb, ids in zip(plan.tiles, pipeline.tiles_by_tau0(bank.m1, bank.m2, 15.0, 64)):
        assert np.array_equal(tb.ids, ids)
        r_o = core.svd_tile(H[ids], swh, kmin, 1e-4)[0]
        for t in ids:
            same_rank[int(t)] = r_o == tb.r
    ex = trig['exact']
    assert (~ex).sum() > 100 and ex.sum() > 20, (int((~ex).sum()), int(ex.sum()))
    assert np.all(trig['chisq'][~ex] == 0) and np.all(trig['chisq_dof'][~ex] == 0)
    assert np.all(trig['chisq_dof'][ex] == 30)
    np.testing.assert_array_equal(trig['newsnr'][~ex], trig['snr'][~ex].astype(np.float64))
    rb_cut = np.float32(esnr * (1 - margin))
    assert np.all(trig['snr'][~ex] < rb_cut) and np.all(trig['snr'][~ex] > thr)
    # against the float64 direct filter
    d_snr, d_ph = _direct(H, h['D'][0], h['IP'][0], ds.df, ds.tlen, trig['template_id'], trig['sample'])
    assert np.max(np.abs(trig['snr'][ex] - d_snr[ex]) / d_snr[ex]) < 1e-4
    assert np.max(np.abs(trig['snr'][~ex] - d_snr[~ex])) < SVD_SNR_BOUND
    dph = np.abs(np.angle(np.exp(1j * (trig['coa_phase'] - d_ph))))
    assert np.max(dph[~ex] * d_snr[~ex]) < SVD_SNR_BOUND + 0.02
    # against the oracle's lists
    kw = lambda t, s: (int(t), int(s) // 2048)
    kg = {kw(t, s): i for i, (t, s) in enumerate(zip(trig['template_id'], trig['sample']))}
    ko = {kw(t, s): i for i, (t, s) in enumerate(zip(o['tid'], o['samp']))}
    for k in set(kg) ^ set(ko):
        v = trig['snr'][kg[k]] if k in kg else o['snr'][ko[k]]
        assert abs(v - thr) < SVD_SNR_BOUND, (k, v)
    both = sorted(set(kg) & set(ko))
    same = [k for k in both if trig['sample'][kg[k]] == o['samp'][ko[k]]]
    assert len(same) >= {'fp32': 0.98, 'bf16': 0.7}[precision] * len(both)
    tol_eq = {'fp32': 2e-4, 'bf16': 1e-2}[precision]
    nrb = 0
    for k in same:
        i, j = kg[k], ko[k]
        if bool(trig['exact'][i]) != bool(o['exact'][j]):
            assert abs(trig['snr'][i] - rb_cut) < SVD_SNR_BOUND or abs(o['snr'][j] - rb_cut) < SVD_SNR_BOUND
            continue
        if trig['exact'][i]:
            assert abs(trig['snr'][i] - o['snr'][j]) / o['snr'][j] < 1e-4
            continue
        nrb += 1
        if same_rank[k[0]]:
            assert abs(trig['snr'][i] - o['snr'][j]) / o['snr'][j] < tol_eq, (k, trig['snr'][i], o['snr'][j])
            assert abs(np.angle(np.exp(1j * (trig['coa_phase'][i] - o['ph'][j])))) < 5 * tol_eq
        else:
            assert abs(trig['snr'][i] - o['snr'][j]) < 2 * SVD_SNR_BOUND
    assert nrb > 100


def test_pipelined_equals_synchronised(dense_c1):
    """run_banks as a software pipeline (two batches ahead, one segment per batch, consolidation on its worker
    thread, pre-consolidation and sine-Gaussian values one batch late) gives the same peak-file fields as the
    synchronised form (every stage awaited), as a second run of the same Filter (the floors of the first run are not
    carried over) and as a run after rebind to the same data."""
    import consolidate, fp_filter, gpu_prep, synth
    bank, ds, _ = dense_c1
    st2 = synth.make_stream('L1', 64, SR, seed=2, injections=[synth.Injection(40.0, 33.3, 27.1, 14.0, 0.2)])
    ds2 = gpu_prep.condition(torch.from_numpy(st2.x).cuda(), SR, st2.gps_start, 64, 8, 8, 4, 2, 31, 4, 15.0,
                             tag='L1')
    data = [ds, ds2]
    kl = consolidate.KeepLoudest(4.0, 1.072, 3, 5.0)
    base = dict(tile=32, precision='bf16', snr_threshold=4.0, exact_snr=5.5, refine=3, segment_batch=1,
                consolidate=kl, pre_consolidate=True, sgchisq_snr_threshold=5.0, self_check=False)
    IPall, sw, isw = fp_filter.svd_metric(data)
    plan = fp_filter.BankPlan(bank.m1, bank.m2, None, ds.flen, ds.df, fp_filter.Options(**base), sw, isw,
                              sg_locations=['mtotal>40:20-15,20-30,20-45'])
    copy = lambda d: fp_filter.Dataset(**{**d.__dict__, 'D': d.D.clone(), 'IP': d.IP.clone()})

    def fields(f):
        return [t.to_dict(gates=d.gates) for t, d in zip(f.run(plan), f.data)]

    _say('synchronised')
    ref = fields(fp_filter.Filter([copy(d) for d in data], fp_filter.Options(**base, timing=True)))
    assert sum(len(o['snr']) for o in ref) > 50 and sum(int(o['sg_chisq_computed'].sum()) for o in ref) > 0
    f = fp_filter.Filter([copy(d) for d in data], fp_filter.Options(**base))
    for rnd in ('pipelined', 'second run', 'after rebind'):
        _say(rnd)
        if rnd == 'after rebind':
            f.rebind(data)
        got = fields(f)
        for a, b in zip(ref, got):
            for k in a:
                np.testing.assert_array_equal(a[k], b[k], err_msg=f'{rnd}: {k}')


@pytest.mark.parametrize('precision', ['bf16', 'fp32', 'fp8', 'fp16'])
def test_second_pass_lists_equal_first_pass(dense_c1, precision):
    """Alternates from local-maximum lists made by a second GEMM pass over the windows that hold exact-band
    candidates (Options.alt_second_pass above every batch's fraction) equal those kept by the first pass
    (alt_second_pass 0): the same peak-file fields, with alternates evaluated and second-pass windows run."""
    import fp_filter
    bank, ds, _ = dense_c1
    base = dict(tile=32, precision=precision, snr_threshold=4.0, exact_snr=5.5, refine=3, alternates=3,
                alt_margin=0.05, segment_batch=1, self_check=False)
    IPall, sw, isw = fp_filter.svd_metric([ds])
    plan = fp_filter.BankPlan(bank.m1, bank.m2, None, ds.flen, ds.df, fp_filter.Options(**base), sw, isw)
    out = {}
    for name, frac in (('first', 0.0), ('second', 1.0)):
        f = fp_filter.Filter([ds], fp_filter.Options(**base, alt_second_pass=frac))
        out[name] = (f.run(plan)[0].to_dict(), dict(f.T))
        _say(name, {k: out[name][1].get(k) for k in ('n_alternates', 'second_pass_windows', 'n_exact_candidates')})
    a, b = out['first'][0], out['second'][0]
    assert len(a['snr']) > 50
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    assert out['first'][1].get('n_alternates', 0) > 0 and out['first'][1].get('second_pass_windows', 0) == 0
    assert out['second'][1].get('second_pass_windows', 0) > 0
    assert out['second'][1]['n_alternates'] == out['first'][1]['n_alternates']


def _production_tile(bank, n_tile, target_mc):
    """The 512-template chirp-time tile holding the template closest in chirp mass to target_mc."""
    import synth
    tau = synth.tau0(bank.m1, bank.m2, 15.0)
    order = np.argsort(tau, kind='stable')
    mc = (bank.m1 * bank.m2) ** 0.6 / (bank.m1 + bank.m2) ** 0.2
    j = int(np.argmin(np.abs(mc[order] - target_mc)))
    t0 = j // n_tile * n_tile
    return np.sort(order[t0:t0 + n_tile])


def test_production_shape_tile(gpu):
    """2 detectors x 512 s at 2048 Hz with 5 lines, 4 glitches and 20 injections (the production workload), one
    512-template tile of each bank (circular: the chirp-time tile around the loudest injection's chirp mass;
    eccentric: every second template):
    autogating finds the oracle's glitches, the conditioning equals the oracle's on the gated strain, the fp32 plain
    form finds the oracle's (template, window) list on H1 (float32 plain method, 16 threads) except within 1e-3 of the
    threshold, every exact SNR of the loudest templates is within 1e-4 of the float64 direct filter, and the bf16
    form with +-3-sample re-centring gives the windows of the fp32 form with the same re-centring."""
    import fp_filter, synth, workloads
    _say('production data')
    streams, banks, cond = workloads.production()
    x_dev = [torch.from_numpy(s.x).cuda() for s in streams]
    ds = workloads.condition_streams(streams, cond, x_dev)            # gates x_dev in place
    for s, d, x in zip(streams, ds, x_dev):
        gt = core.detect_loud_glitches(s.x, s.sr, s.gps_start, 50.0, cond.autogating_cluster, 15.0,
                                       cond.autogating_pad)
        assert len(d.gates) == len(gt) >= 2
        assert np.max(np.abs(np.array([g[0] for g in d.gates]) - np.array(gt))) <= 1.0 / s.sr
    _say('oracle conditioning H1')
    xg = x_dev[0].cpu().numpy()
    ref = core.condition(xg, SR, streams[0].gps_start, cond.seg_len, cond.start_pad, cond.end_pad, cond.psd_seg_len,
                         cond.psd_stride, cond.psd_nseg, cond.psd_inv_len, 15.0)
    ip = ds[0].IP.cpu().numpy()
    nz = ref['IP'] > 0
    assert np.array_equal(ip > 0, nz)
    assert np.max(np.abs(ip - ref['IP'])[nz] / ref['IP'][nz]) < 1e-4
    Dg = ds[0].D.cpu().numpy()
    assert np.linalg.norm(Dg - ref['D']) / np.linalg.norm(ref['D']) < 1e-5
    inj = streams[0].injections
    loud = max(inj, key=lambda j: j.snr)
    mc_loud = (loud.m1 * loud.m2) ** 0.6 / (loud.m1 + loud.m2) ** 0.2
    IPall, sw, isw = fp_filter.svd_metric(ds)
    sel = [_production_tile(banks[0], 512, mc_loud), np.arange(len(banks[1]))[::len(banks[1]) // 512][:512]]
    outs = {}
    opts0 = fp_filter.Options(tile=512, precision='fp32', self_check=False)
    plans = [fp_filter.BankPlan(b.m1[s_], b.m2[s_], b.e0[s_], ds[0].flen, ds[0].df, opts0, sw, isw, name=b.name)
             for b, s_ in zip(banks, sel)]
    _say('ranks', [pl.ranks for pl in plans])
    for name, p, rf in (('fp32', 'fp32', 0), ('fp32r', 'fp32', 3), ('bf16', 'bf16', 3)):
        opts = fp_filter.Options(tile=512, precision=p, refine=rf, self_check=False)
        _say('gpu', name)
        outs[name] = [[t.to_dict(gates=d.gates) for t, d in zip(tr, ds)]
                      for tr in fp_filter.Filter(ds, opts).run_banks(plans)]
    assert sum(len(o['snr']) for ob in outs['fp32'] for o in ob) > 100
    # every exact SNR of the 12 loudest templates per (bank, detector) against the float64 direct filter
    for bi, plan in enumerate(plans):
        tb = plan.tiles[0]
        row = np.empty(len(tb.ids), np.int64)
        row[tb.ids] = np.arange(len(tb.ids))            # H rows are in chirp-time order
        for di, d in enumerate(ds):
            for p in ('fp32', 'fp32r', 'bf16'):
                o = outs[p][bi][di]
                top = np.unique(o['template_id'][np.argsort(-o['snr'])][:40])[:12]
                m = np.isin(o['template_id'], top)
                snr, _ = _direct(tb.H, d.D[0], d.IP[0], d.df, d.tlen, row[o['template_id'][m]], o['sample'][m])
                assert np.max(np.abs(o['snr'][m] - snr) / snr) < 1e-4, (bi, di, p)
            a, b = outs['fp32r'][bi][di], outs['bf16'][bi][di]
            wa = set(zip(a['template_id'].tolist(), (a['sample'] // 2048).tolist()))
            wb = set(zip(b['template_id'].tolist(), (b['sample'] // 2048).tolist()))
            for t, w in wa ^ wb:
                v = [x for o_ in (a, b) for x in o_['snr'][(o_['template_id'] == t) & (o_['sample'] // 2048 == w)]]
                assert all(abs(x - 5.5) < 3e-3 * 5.5 for x in v), (bi, di, t, w, v)
    _say('oracle plain method, H1, circular tile')
    h = _host(ds[0])
    ncpu = max(1, min(16, len(os.sched_getaffinity(0))))
    res, _ = pipeline.filter_plain([h], banks[0].m1[sel[0]], banks[0].m2[sel[0]], None, thr=5.5, tile=512, eps=1e-4,
                                   dtype=np.float32, workers=ncpu, sw=sw.cpu().numpy())
    o, g = res[0], outs['fp32'][0][0]
    # the peak file's gate veto (fp_filter.py:914-916) on the oracle's list
    t_end = ds[0].gps_start + (ds[0].seg_start[0] + o['samp']) / SR
    veto = np.zeros(len(t_end), bool)
    for g_t, g_w, g_taper in ds[0].gates:
        veto |= np.abs(t_end - g_t) <= g_w + g_taper + 0.5
    o = {k: v[~veto] for k, v in o.items()}
    ko = set(zip(o['tid'].astype(int).tolist(), (o['samp'].astype(int) // 2048).tolist()))
    kg = set(zip(g['template_id'].tolist(), (g['sample'] // 2048).tolist()))
    assert len(ko) > 50
    for t, w in ko ^ kg:
        v = [x for x in o['snr'][(o['tid'] == t) & (o['samp'] // 2048 == w)]] + \
            [x for x in g['snr'][(g['template_id'] == t) & (g['sample'] // 2048 == w)]]
        assert all(abs(x - 5.5) < 1e-3 * 5.5 for x in v), (t, w, v)


def test_batched_search_vs_single_filters(gpu):
    """The batched form on six uneven chunks per detector (80 s: nothing to analyse, 192 s on the 128 s grid, 320 /
    448 s on the 256 s grid, 640 / 1024 s on the 512 s grid), both banks (96 circular + 32 eccentric templates, tiles of 64), fp32: every
    chunk's peak-file fields equal those of a Filter run on that chunk alone with the same bank products, and exact
    SNRs of the loudest templates equal the float64 direct filter to 1e-4."""
    import batched, fp_filter, synth, workloads
    rng = np.random.default_rng(3)
    banks = [synth.circular_bank(96, seed=11), synth.eccentric_bank(32, seed=12)]
    chunks, t0 = [], 1.0e9
    for L in (80, 192, 640, 320, 1024, 448):
        for d, ifo in enumerate(('H1', 'L1')):
            j = int(rng.integers(0, 96))
            inj = [synth.Injection(float(L) / 2 + 0.3 * d, float(banks[0].m1[j]), float(banks[0].m2[j]), 15.0, 0.4)]
            st = synth.make_stream(ifo, L, SR, seed=100 + 10 * L + d, lines=synth.default_lines(), gps_start=t0,
                                   injections=inj)
            chunks.append((st, workloads.chunk_conditioning(L)))
        t0 += L + 8
    x_dev = [torch.from_numpy(st.x).cuda() for st, _ in chunks]
    opts = {'fp32': fp_filter.Options(tile=64, precision='fp32', self_check=False, segment_batch=4)}
    _say('batched search')
    bs = batched.BatchedSearch(chunks, banks, x_dev, opts)
    assert bs.grids == [128, 256, 512]
    ds, outs, _ = bs.step('fp32')
    n_tr = 0
    for g in bs.grids:
        seg0 = np.cumsum([0] + [len(d.analyze) for d in ds[g]])
        for di, d in enumerate(ds[g]):
            _say('single chunk', g, di)
            single = fp_filter.Filter([d], opts['fp32']).run_banks(bs.plans[g])
            for bi in range(len(banks)):
                exp = single[bi][0].to_dict(gates=d.gates)
                got = outs[g][bi][di]
                n_tr += len(got['snr'])
                for k in exp:
                    # 'segment' counts the segments of every dataset the Filter holds
                    off = seg0[di] if k == 'segment' else 0
                    np.testing.assert_array_equal(got[k] - off, exp[k], err_msg=f'grid {g} dataset {di} bank {bi}: {k}')
                if len(got['snr']):
                    tile_of = {int(t): (tb, r) for tb in bs.plans[g][bi].tiles for r, t in enumerate(tb.ids)}
                    top = got['template_id'][np.argmax(got['snr'])]
                    m = got['template_id'] == top
                    tb, r = tile_of[int(top)]
                    seg = got['segment'][m] - seg0[di]
                    for s in np.unique(seg):
                        mm = np.flatnonzero(m)[seg == s]
                        snr, _ = _direct(tb.H, d.D[int(s)], d.IP[int(s)], d.df, d.tlen, np.full(len(mm), r),
                                         got['sample'][mm])
                        assert np.max(np.abs(got['snr'][mm] - snr) / snr) < 1e-4
    assert n_tr > 20
    # several slices in one software pipeline (each slice conditioned when the pipeline reaches it, while the
    # previous slice's last batches run) give each slice's own one-slice run
    sels = [[0, 1, 2, 3], [4, 5, 6, 7], [8, 9, 10, 11]]
    _say('pipelined slices')
    multi = bs.run('fp32', sels)
    for k, sel in enumerate(sels):
        ds1, outs1, _ = bs.step('fp32', sel)
        dsm, outsm, _ = multi[k]
        assert [g for g in bs.grids if ds1[g]] == [g for g in bs.grids if dsm[g]]
        for g in bs.grids:
            for bi in range(len(banks)):
                for a, b in zip(outs1[g][bi], outsm[g][bi]):
                    for key in a:
                        np.testing.assert_array_equal(a[key], b[key], err_msg=f'slice {k} grid {g} bank {bi}: {key}')


def test_condition_vs_reference_stage2_golden(gpu):
    """gpu_prep.condition against the reference's own gpu_stage2 on the golden seeded strains (one PSD span; four
    spans with picks 0, 0, 2, 3, 3 and the Hann-tapered truncation)."""
    import gpu_prep
    gold = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'reference_vectors.npz'))
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), 'golden'))
    import make_golden as mg
    meta = json.loads(str(gold['s2_json']))
    for (name, seed, sr, seconds, pad, opts), m in zip(mg.STAGE2, meta):
        o = {opts[i][2:].replace('-', '_'): opts[i + 1] for i in range(0, len(opts), 2)}
        h = mg.stage2_strain(seed, sr, seconds, pad)
        s = torch.from_numpy(h[int(pad * sr):len(h) - int(pad * sr)]).cuda()
        d = gpu_prep.condition(s, float(sr), float(mg.STAGE2_GPS), int(o['segment_length']),
                               int(o['segment_start_pad']), int(o['segment_end_pad']), float(o['psd_segment_length']),
                               float(o['psd_segment_stride']), int(o['psd_num_segments']),
                               float(o['psd_inverse_length']), 15.0, trunc_method=o.get('invpsd_trunc_method'))
        D, IP = d.D.cpu().numpy(), d.IP.cpu().numpy()
        gD, gIP = gold[f's2_{name}_D'], gold[f's2_{name}_IP']
        nz = gIP > 0
        st = mg.STAGE2_STRIDE
        assert np.array_equal(IP[:, ::st] > 0, nz), name
        assert np.max(np.abs(IP[:, ::st] - gIP)[nz] / gIP[nz]) < 2e-5, name
        assert np.max(np.abs(D[:, ::st] - gD)) < 1e-5 * np.abs(gD).max(), name
        assert [list(a) for a in d.analyze] == m['analyze'] and list(d.seg_start) == m['seg_start']
        assert d.flen == m['flen'] and d.tlen == m['tlen'] and d.df == m['df']


def test_condition_many_equals_condition(gpu):
    """gpu_prep.condition_many (every strain of a segment grid in a few batched launches) gives each strain's
    condition() result: the same segment layout, IP within 1e-6 and D within 1e-6 of its scale (the batched FFTs may
    round differently), for strains of different lengths and Welch segment counts (one and several PSD spans)."""
    import gpu_prep, synth, workloads
    xs, conds, sts = [], [], []
    for L, seed in ((192, 3), (320, 4), (560, 5), (448, 6)):
        st, x = synth.device_stream('H1', L, SR, seed, lines=synth.default_lines(), gps_start=1e9 + 2000 * seed)
        xs.append(x); conds.append(workloads.chunk_conditioning(L)); sts.append(st)
    c0 = conds[0]
    many = gpu_prep.condition_many(xs, SR, [s.gps_start for s in sts], 128, c0.start_pad, c0.end_pad, c0.psd_seg_len,
                                   c0.psd_stride, [c.psd_nseg for c in conds], c0.psd_inv_len, 15.0,
                                   tags=['H1'] * len(xs))
    for x, st, c, m in zip(xs, sts, conds, many):
        one = gpu_prep.condition(x.clone(), SR, st.gps_start, 128, c.start_pad, c.end_pad, c.psd_seg_len,
                                 c.psd_stride, c.psd_nseg, c.psd_inv_len, 15.0, tag='H1')
        assert m.analyze == one.analyze and m.seg_start == one.seg_start and m.gps_start == one.gps_start
        ip1, ipm = one.IP.cpu().numpy(), m.IP.cpu().numpy()
        nz = ip1 > 0
        assert np.array_equal(ipm > 0, nz)
        assert np.max(np.abs(ipm - ip1)[nz] / ip1[nz]) < 1e-6
        d1, dm = one.D.cpu().numpy(), m.D.cpu().numpy()
        assert np.max(np.abs(dm - d1)) < 1e-6 * np.abs(d1).max()
