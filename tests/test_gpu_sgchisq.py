"""The sine-Gaussian chi-squared kernel (gwgs_sg_tiles through sgchisq.sg_chisq_points) against the oracle's float64
definition (fp_filter.py:463-486) within the reference's own bound for its float32 kernel (relative 1e-4,
fp_filter.py:876), and the filter's sg_chisq fields against the oracle on the filter's own data."""
import numpy as np
import pytest
import torch

from oracle import core, sgchisq as osg

pytestmark = pytest.mark.gpu

PROD_TILES = [(20.0, float(o)) for o in (15, 30, 45, 60, 75, 90, 105, 120)]


@pytest.mark.parametrize('N', [1 << 20, 3 << 18])
def test_sg_points_match_oracle(gpu, N):
    import sgchisq
    print('sg_tiles N', N, flush=True)
    rng = np.random.default_rng(N % 1000)
    flen = N // 2 + 1
    df = 2048.0 / N
    D = ((rng.standard_normal(flen) + 1j * rng.standard_normal(flen)) * 1e-2).astype(np.complex64)
    IP = rng.uniform(0.5, 2.0, flen).astype(np.float32)
    npk = 48
    t = rng.integers(0, N, npk)
    fpeak = rng.uniform(20, 900, npk)                       # some reach above 0.9 x Nyquist: not computed
    kst = np.full(npk, int(15 / df))
    vals, ok = sgchisq.sg_chisq_points(torch.from_numpy(D).cuda(), torch.from_numpy(IP).cuda(), t, fpeak, PROD_TILES,
                                       kst, df, N)
    assert 0 < ok.sum() < npk
    for i in range(npk):
        ref = osg.sg_chisq_one(D, IP, int(t[i]), fpeak[i], PROD_TILES, int(kst[i]), df, N)
        assert (ref is not None) == bool(ok[i])
        if ref is None:
            assert vals[i] == 1.0
        else:
            assert abs(vals[i] - ref) <= 1e-4 * ref, (i, vals[i], ref)


def test_sg_points_signal_tile(gpu):
    """Closed form through the kernel: data c IP g e^{-2 pi i k t / N} on one tile gives the value |c|^2 sigma^2 / 2."""
    import sgchisq
    N = 1 << 20
    df = 2048.0 / N
    flen = N // 2 + 1
    rng = np.random.default_rng(3)
    IP = rng.uniform(0.5, 2.0, flen)
    q, off, fpeak, t, c = 20.0, 45.0, 150.0, 777777, 0.3 + 0.4j
    fcen = fpeak + off
    k = np.arange(flen)
    f = k * df
    tau = q / (2 * np.pi * fcen)
    kk, k1, kg = int((fcen - 50) / df), int((fcen + 50) / df), int(round((fcen - 50) / df))
    g = np.sqrt(np.pi) / 2 * tau * np.exp(-(np.pi * tau * (f - fcen)) ** 2) * (1 + np.exp(-q * q * f / fcen))
    g[(k < kg) | (k >= k1)] = 0
    D = (c * IP * g * np.exp(-2j * np.pi * ((k * t) % N) / N)).astype(np.complex64)
    v, ok = sgchisq.sg_chisq_points(torch.from_numpy(D).cuda(), torch.from_numpy(IP.astype(np.float32)).cuda(),
                                    np.array([t]), np.array([fpeak]), [(q, off)], np.array([0]), df, N)
    exp = abs(c) ** 2 * 4 * df * (g * g * IP).sum() / 2
    assert ok[0] and abs(v[0] - exp) <= 1e-4 * exp


def test_filter_sg_fields(gpu):
    """One detector, 64 s, 64 templates with two loud injections: every exact trigger with newsnr >= 5 of a template
    with mtotal > 12 carries the sine-Gaussian chi-squared of the oracle (on the filter's D, IP, template edges and
    sample) within 1e-4; the others carry 1 and computed = False."""
    import consolidate, fp_filter, gpu_prep, synth
    bank = synth.circular_bank(64, seed=11, mmin=5, mmax=100)
    inj = [synth.Injection(30.0, float(bank.m1[5]), float(bank.m2[5]), 12.0, 0.3),
           synth.Injection(47.0, float(bank.m1[40]), float(bank.m2[40]), 25.0, 1.1)]
    st = synth.make_stream('H1', 64, 2048.0, seed=1, injections=inj)
    ds = gpu_prep.condition(torch.from_numpy(st.x).cuda(), 2048.0, st.gps_start, 64, 8, 8, 4, 2, 31, 4, 15.0, tag='H1')
    opts = fp_filter.Options(tile=32, precision='bf16', snr_threshold=5.0, sgchisq_snr_threshold=5.0)
    IPall, sw, isw = fp_filter.svd_metric([ds])
    locs = ['mtotal>12:' + ','.join('%g-%g' % v for v in PROD_TILES)]
    plan = fp_filter.BankPlan(bank.m1, bank.m2, None, ds.flen, ds.df, opts, sw, isw, sg_locations=locs)
    f = fp_filter.Filter([ds], opts)
    tr = f.run(plan)[0].to_dict()
    ex = tr['exact']
    ns = consolidate.newsnr(tr['snr'], tr['chisq'] / np.maximum(tr['chisq_dof'], 1))
    mt = bank.m1[tr['template_id']] + bank.m2[tr['template_id']]
    want = ex & (ns >= 5.0) & (mt > 12)
    assert want.sum() > 3
    D, IP = ds.D[0].cpu().numpy(), ds.IP[0].cpu().numpy()
    where = {int(i): (tb, j) for tb in plan.tiles for j, i in enumerate(tb.ids)}
    n_done = 0
    for i in range(len(ex)):
        if not want[i]:
            assert tr['sg_chisq'][i] == 1.0 and not tr['sg_chisq_computed'][i]
            continue
        tb, j = where[int(tr['template_id'][i])]
        E = core.edges_per_template([j], [16], plan.kmin, [int(tb.kmax[j])], tb.H.cpu().numpy(), IP)
        fp = core.fpeak(E, [16], ds.df)[0]
        ref = osg.sg_chisq_one(D, IP, int(tr['sample'][i]), fp, PROD_TILES, plan.kmin, ds.df, ds.tlen)
        assert (ref is not None) == bool(tr['sg_chisq_computed'][i])
        if ref is not None:
            assert abs(tr['sg_chisq'][i] - ref) <= 1e-4 * ref, (i, tr['sg_chisq'][i], ref)
            n_done += 1
    assert n_done > 3 and f.T['n_sg_chisq_points'] == n_done
