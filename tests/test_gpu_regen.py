"""Coarse-node templates and the regenerated exact stage (the reference's --templates interp and --exact-stage regen,
imr_interp.py, exact_stage.py:472-647): the nodes against the closed-form phase, the dense interpolation against the
oracle's restatement of FineInterpolator, the regenerated sums and the edges from moments against the dense path on
the same rows, and the filter end to end against the float64 direct filter of the interpolated templates."""
import numpy as np
import pytest
import torch

from oracle import core, pipeline

pytestmark = pytest.mark.gpu

SR = 2048.0
MSUN_S = 4.925490947641267e-06


def _fisco(m1, m2):
    return 1.0 / (6.0 ** 1.5 * np.pi * (np.asarray(m1) + np.asarray(m2)) * MSUN_S)


def test_nodes_and_dense_interpolation(gpu):
    """Amplitudes at the nodes are f^(-7/6) below f_isco and 0 above; each interval's cubic passes through the
    closed-form phase -psi at both ends with its derivative (central differences of psi) at the left end; the dense
    rows equal the oracle's FineInterpolator on the same nodes to complex64 rounding, and the closed-form rows to
    1e-3 rad in phase below the ISCO interval."""
    import templates as tm
    df, flow = 1.0 / 64, 15.0
    flen = 64 * 2048 // 2 + 1
    kmin = int(flow / df)
    m1, m2, e0 = np.array([10.0, 30.0, 6.0]), np.array([8.0, 25.0, 5.0]), np.array([0.0, 0.2, 0.1])
    fi = _fisco(m1, m2)
    kend = min(int(fi.max() / df), flen - 1)
    x = tm.coarse_grid(flow, kend * df, 600, 150)
    np.testing.assert_array_equal(x, pipeline.coarse_grid(flow, kend * df, 600, 150))
    g = tm.Grid(x, df, kmin, kend)
    amp, coef = tm.nodes(m1, m2, e0, g)
    A, C = amp.cpu().numpy(), coef.cpu().numpy()
    h = np.diff(x)
    for i in range(3):
        np.testing.assert_allclose(A[i], np.where(x < fi[i], x ** (-7.0 / 6.0), 0.0), rtol=1e-13)
        ph = -pipeline.spa_phase(x, m1[i], m2[i], e0[i])
        c = C[i]
        scale = np.abs(ph).max()
        assert np.abs(c[:, 3] - ph[:-1]).max() <= 1e-13 * scale
        p_end = ((c[:, 0] * h + c[:, 1]) * h + c[:, 2]) * h + c[:, 3]
        assert np.abs(p_end - ph[1:]).max() <= 1e-12 * scale
        eps = 1e-6 * x[:-1]
        dnum = -(pipeline.spa_phase(x[:-1] + eps, m1[i], m2[i], e0[i]) -
                 pipeline.spa_phase(x[:-1] - eps, m1[i], m2[i], e0[i])) / (2 * eps)
        np.testing.assert_allclose(c[:, 2], dnum, rtol=1e-6)
    H = tm.interp_templates(amp, coef, g, flen).cpu().numpy()
    Ho = pipeline.fine_interp(A, C, x, df, kmin, kend, flen)
    assert np.abs(H - Ho).max() <= 2e-7 * np.abs(Ho).max()
    Hc, kmax = pipeline.spa_bank(m1, m2, e0, flen, df, flow)
    for i in range(3):
        j = np.searchsorted(x, fi[i]) - 1                       # the interval holding f_isco
        k1 = min(int(x[max(j, 0)] / df), kmax[i])
        d = np.angle(Ho[i, kmin:k1] * np.conj(Hc[i, kmin:k1]))
        assert np.abs(d).max() < 1e-3, (i, np.abs(d).max())
        np.testing.assert_allclose(np.abs(Ho[i, kmin:k1]), np.abs(Hc[i, kmin:k1]), rtol=1e-3)


@pytest.fixture(scope='module')
def interp_data(gpu):
    """64 s of coloured noise with two injections; 48 templates (5-60 Msun, half eccentric) as interpolated rows in
    one tile, with the dense and the regenerated exact stage."""
    import fp_filter, gpu_prep, synth
    bank = synth.circular_bank(48, seed=21, mmin=5, mmax=60)
    e0 = np.where(np.arange(48) % 2 == 1, 0.15, 0.0)
    inj = [synth.Injection(20.0, float(bank.m1[7]), float(bank.m2[7]), 14.0, 0.4),
           synth.Injection(44.0, float(bank.m1[30]), float(bank.m2[30]), 22.0, 2.0)]
    st = synth.make_stream('H1', 64, SR, seed=4, injections=inj)
    ds = gpu_prep.condition(torch.from_numpy(st.x).cuda(), SR, st.gps_start, 64, 8, 8, 4, 2, 31, 4, 15.0, tag='H1')
    IPall, sw, isw = fp_filter.svd_metric([ds])
    plans = {}
    for es in ('dense', 'regen'):
        o = fp_filter.Options(tile=48, precision='fp32', snr_threshold=5.5, templates='interp', exact_stage=es,
                              n_chirp=2000, n_log=500)
        plans[es] = (o, fp_filter.BankPlan(bank.m1, bank.m2, e0, ds.flen, ds.df, o, sw, isw))
    return bank, e0, ds, plans


def test_regenerated_sums_and_moment_edges_vs_dense(interp_data):
    """The regenerated exact stage against the dense stage on the same interpolated rows, at 300 (row, sample)
    peaks: |z| within 1e-4, phase within 1e-4 rad, chi-squared within 1e-3 (the reference's self-check
    tolerances); the edges from moments within one bin of the float64 cumsum edges, at most 1% of them differing;
    and z against the float64 direct sum over the regenerated stage's own bins of the oracle's FineInterpolator rows
    (1e-4)."""
    import exact_stage as es
    bank, e0, ds, plans = interp_data
    tb = plans['regen'][1].tiles[0]
    assert tb.stage.tile.coarse is not None
    dense = es.ExactStage(es.TemplateTile(tb.H, tb.stage.tile.nbins, tb.stage.tile.kmin, tb.stage.tile.kmax),
                          self_check=False)
    rng = np.random.default_rng(3)
    n = tb.H.shape[0]
    rows = rng.integers(0, n, 300)
    samples = rng.integers(0, ds.tlen, 300)
    seg = es.SegmentData(ds.D[0], ds.IP[0], ds.df, ds.tlen)
    req = es.PeakRequest(rows, samples)
    vr = tb.stage.evaluate(seg, req)                         # runs its own first-call self-check too
    vd = dense.evaluate(seg, req)
    rep = es.compare_values(vr.z, vr.pw, vd.z, vd.pw, tb.stage.tile.nbins[rows])
    assert rep['snr_rel_max'] < 1e-4 and rep['phase_max'] < 1e-4 and rep['chisq_rel_max'] < 1e-3, rep
    ru = np.unique(rows)
    Er = tb.stage._edges_moments([seg], ru, np.zeros(len(ru), np.int64)).cpu().numpy()
    Ed = dense.edges(seg, ru)
    dif = np.abs(Er.astype(np.int64) - Ed)
    assert dif.max() <= 1 and (dif > 0).mean() <= 1e-2, (dif.max(), (dif > 0).mean())
    # float64 direct sums over the regenerated stage's bins, on the oracle's interpolation of the same nodes
    c = tb.stage.tile.coarse
    g = c.grid
    Ho = pipeline.fine_interp(c.amp.cpu().numpy(), c.coef.cpu().numpy(), g.x, ds.df, g.kmin, g.kend, ds.flen)
    D = ds.D[0].cpu().numpy().astype(np.complex128)
    Emap = {int(r): Er[i] for i, r in enumerate(ru)}
    for i in range(0, 300, 15):
        e = Emap[int(rows[i])][:17]
        k = np.arange(e[0], e[-1])
        z = (np.conj(Ho[rows[i], e[0]:e[-1]]) * D[e[0]:e[-1]] *
             np.exp(2j * np.pi * ((k * int(samples[i])) % ds.tlen) / ds.tlen)).sum()
        assert abs(abs(complex(vr.z[i])) - abs(z)) <= 1e-4 * abs(z), i


def test_filter_interp_regen_vs_dense_and_direct(interp_data):
    """The filter with interpolated templates: the regenerated exact stage reports the dense stage's triggers (same
    (template, sample) list, SNR within 1e-4, chi-squared within 1e-3), and their SNRs equal the float64 direct
    matched filter of the oracle's interpolated templates (1e-4); the louder injection is found."""
    import fp_filter
    bank, e0, ds, plans = interp_data
    out = {}
    for es_, (o, plan) in plans.items():
        out[es_] = fp_filter.Filter([ds], o).run(plan)[0].to_dict()
    a, b = out['dense'], out['regen']
    ka = dict(zip(zip(a['template_id'].tolist(), a['sample'].tolist()), range(len(a['snr']))))
    kb = dict(zip(zip(b['template_id'].tolist(), b['sample'].tolist()), range(len(b['snr']))))
    for k in set(ka) ^ set(kb):                                # only at the threshold may one list hold a trigger
        v = a['snr'][ka[k]] if k in ka else b['snr'][kb[k]]
        assert abs(v - 5.5) < 1e-3, (k, v)
    both = sorted(set(ka) & set(kb))
    assert len(both) > 3
    ia, ib = np.array([ka[k] for k in both]), np.array([kb[k] for k in both])
    np.testing.assert_allclose(b['snr'][ib], a['snr'][ia], rtol=1e-4)
    np.testing.assert_allclose(b['chisq'][ib], a['chisq'][ia], rtol=1e-3, atol=1e-3)
    assert b['snr'].max() > 15
    plan = plans['regen'][1]
    c = plan.tiles[0].stage.tile.coarse
    g = c.grid
    Ho = pipeline.fine_interp(c.amp.cpu().numpy(), c.coef.cpu().numpy(), g.x, ds.df, g.kmin, g.kend, ds.flen)
    row = np.empty(len(plan.tiles[0].ids), np.int64)
    row[plan.tiles[0].ids] = np.arange(len(row))
    D = ds.D[0].cpu().numpy()
    IP = ds.IP.cpu().numpy()
    top = np.argsort(-b['snr'])[:12]
    H = Ho[row[b['template_id'][top]]].astype(np.complex64)
    sig, norm = core.sigmasq(H, IP, ds.df)
    for j, i in enumerate(top):
        z = core.direct_z(H[j:j + 1], D, ds.tlen)[0, b['sample'][i]]
        assert abs(b['snr'][i] - abs(z) * norm[j, 0]) <= 1e-4 * b['snr'][i], (i, b['snr'][i], abs(z) * norm[j, 0])
