"""The differentiable peak SNR (csrc/refine.hip, differentiable.py): value and gradients against a float64 reference
built from the definition, central finite differences (h = 1e-3), and the refinement's gain on an injection whose
masses lie off the bank."""
import numpy as np
import pytest
import torch

from oracle import core, pipeline

pytestmark = pytest.mark.gpu

SR = 2048.0


@pytest.fixture(scope='module')
def setup(gpu):
    """64 s of coloured noise with one SNR-25 injection whose masses sit between bank templates; 128 templates
    (12-30 Msun) in one tile; the fp32 search's loudest exact triggers near the injection."""
    import fp_filter, gpu_prep, synth
    bank = synth.circular_bank(128, seed=11, mmin=12, mmax=30)
    # off the grid: the geometric mean of two neighbouring templates' component masses
    i0 = int(np.argsort(bank.m1 + bank.m2)[64]); i1 = int(np.argsort(bank.m1 + bank.m2)[65])
    m1, m2 = np.sqrt(bank.m1[i0] * bank.m1[i1]), np.sqrt(bank.m2[i0] * bank.m2[i1])
    inj = synth.Injection(40.0, float(m1), float(m2), 25.0, 0.7)
    st = synth.make_stream('H1', 64, SR, seed=3, injections=[inj])
    ds = gpu_prep.condition(torch.from_numpy(st.x).cuda(), SR, st.gps_start, 64, 8, 8, 4, 2, 31, 4, 15.0, tag='H1')
    IPall, sw, isw = fp_filter.svd_metric([ds])
    opts = fp_filter.Options(tile=128, precision='fp32', snr_threshold=5.5, self_check=False)
    plan = fp_filter.BankPlan(bank.m1, bank.m2, None, ds.flen, ds.df, opts, sw, isw)
    trig = fp_filter.Filter([ds], opts).run(plan)[0].to_dict()
    near = np.abs(trig['end_time'] - (st.gps_start + 40.0)) < 0.1
    assert near.any()
    return dict(bank=bank, ds=ds, plan=plan, trig=trig, near=near, inj=inj, st=st, m=(m1, m2))


def _ref64(ctx, a, t, delta):
    """rho^2 in float64 from the definition: z = sum_k (sum_j a_j Vc_jk) D_k e^{2 pi i k (t + delta) / n},
    sigma^2 = 4 df sum_k |sum_j a_j Vc_jk|^2 IP_k. a complex128 (npk, r), delta float64 (npk,) (differentiable)."""
    V = ctx.Vt.T.to(torch.complex128)                                        # (r, kb)
    k = torch.arange(ctx.k0, ctx.k0 + ctx.kb, device='cuda', dtype=torch.float64)
    D = ctx.D[ctx.k0:ctx.k0 + ctx.kb].to(torch.complex128)
    ip = ctx.IP64
    g = a @ V                                                                 # (npk, kb): the template spectra
    ph = 2 * np.pi * (((k[None, :].to(torch.int64) * t[:, None]) % ctx.n).to(torch.float64)
                      + k[None, :] * delta[:, None]) / ctx.n
    z = (g * D[None, :] * torch.exp(1j * ph)).sum(1)
    sig2 = 4 * ctx.df * ((g.abs() ** 2) * ip[None, :]).sum(1)
    return (z.abs() ** 2) * (4 * ctx.df) ** 2 / sig2, z, sig2


def test_value_and_gradients_vs_float64_and_fd(setup):
    import differentiable as dfb
    ds, plan, trig, near = setup['ds'], setup['plan'], setup['trig'], setup['near']
    tb = plan.tiles[0]
    ctx = dfb.ResponseContext(tb, ds.D[0], ds.IP[0], ds.df, ds.tlen, plan.kmin)
    ctx.IP64 = ds.IP[0][ctx.k0:ctx.k0 + ctx.kb].to(torch.float64)
    order = np.argsort(-trig['snr'][near])[:4]
    tids = trig['template_id'][near][order]
    rows = torch.as_tensor([int(np.flatnonzero(tb.ids == t)[0]) for t in tids], device='cuda')
    t = torch.as_tensor(trig['sample'][near][order], device='cuda', dtype=torch.int64)
    delta = torch.tensor([0.0, 0.3, -0.45, 0.8], dtype=torch.float64, device='cuda')[:len(rows)]
    a = tb.A[rows]
    out = dfb.snr2(ctx, a, t, delta)
    # float64 reference value and autograd gradients on (Re a, Im a, delta)
    a_ri = torch.view_as_real(a.to(torch.complex128)).clone().requires_grad_(True)
    d64 = delta.clone().requires_grad_(True)
    rho2, z, _ = _ref64(ctx, torch.view_as_complex(a_ri), t, d64)
    rho2.sum().backward()
    np.testing.assert_allclose(out['rho2'].cpu().numpy(), rho2.detach().cpu().numpy(), rtol=2e-5)
    ga_ref = torch.view_as_complex(a_ri.grad).cpu().numpy()
    ga = out['grad'].cpu().numpy()
    assert np.abs(ga - ga_ref).max() <= 1e-3 * np.abs(ga_ref).max()
    np.testing.assert_allclose(out['drho2_ddelta'].cpu().numpy(), d64.grad.cpu().numpy(),
                               atol=1e-3 * float(d64.grad.abs().max()))
    # central differences, h = 1e-3, on the float64 definition: delta and 6 random coefficient coordinates
    h = 1e-3
    rp, _, _ = _ref64(ctx, a.to(torch.complex128), t, delta + h)
    rm, _, _ = _ref64(ctx, a.to(torch.complex128), t, delta - h)
    fd_t = ((rp - rm) / (2 * h)).cpu().numpy()
    np.testing.assert_allclose(out['drho2_ddelta'].cpu().numpy(), fd_t, atol=2e-3 * np.abs(fd_t).max() + 1e-6)
    rng = np.random.default_rng(0)
    for _ in range(6):
        j = int(rng.integers(0, tb.r)); part = int(rng.integers(0, 2))
        e = torch.zeros_like(a, dtype=torch.complex128)
        e[:, j] = 1.0 if part == 0 else 1j
        scale = float(a.abs().max())
        ap = a.to(torch.complex128) + h * scale * e
        am = a.to(torch.complex128) - h * scale * e
        fd = ((_ref64(ctx, ap, t, delta)[0] - _ref64(ctx, am, t, delta)[0]) / (2 * h * scale)).cpu().numpy()
        got = (ga[:, j].real if part == 0 else ga[:, j].imag)
        assert np.abs(got - fd).max() <= 2e-3 * np.abs(ga).max(), (j, part, got, fd)
    # phase: d/dphi of Re(e^{-i phi} z)^2 (4 df)^2 / sigma^2 against float64 central differences
    phi = torch.tensor([0.3, 1.1, -2.0, 2.9], device='cuda')[:len(rows)]
    o2 = dfb.snr2(ctx, a, t, delta, phi=phi, want_grad=False)
    _, z64, s64 = _ref64(ctx, a.to(torch.complex128), t, delta)

    def rp_(p):
        return ((torch.exp(-1j * p.to(torch.float64)) * z64).real ** 2 * (4 * ctx.df) ** 2 / s64)
    fdp = ((rp_(phi + h) - rp_(phi - h)) / (2 * h)).cpu().numpy()
    np.testing.assert_allclose(o2['drho2_dphi'].cpu().numpy(), fdp, atol=2e-3 * np.abs(fdp).max())
    # at delta = 0 a bank template's value is its rebuilt SNR: within the SVD bound of the exact trigger SNR
    o0 = dfb.snr2(ctx, a, t, torch.zeros_like(delta), want_grad=False)
    np.testing.assert_allclose(np.sqrt(o0['rho2'].cpu().numpy()), trig['snr'][near][order], rtol=2e-2)


def test_autograd_function_matches_kernel(setup):
    import differentiable as dfb
    ds, plan, trig, near = setup['ds'], setup['plan'], setup['trig'], setup['near']
    tb = plan.tiles[0]
    ctx = dfb.ResponseContext(tb, ds.D[0], ds.IP[0], ds.df, ds.tlen, plan.kmin)
    rows = torch.arange(3, device='cuda')
    t = torch.as_tensor(trig['sample'][near][:1].repeat(3), device='cuda', dtype=torch.int64)
    a_ri = torch.view_as_real(tb.A[rows]).clone().requires_grad_(True)
    d = torch.tensor([0.1, -0.2, 0.0], dtype=torch.float64, device='cuda', requires_grad=True)
    w = torch.tensor([1.0, 2.0, 0.5], device='cuda')
    (dfb.PeakSNR2.apply(a_ri, d, ctx, t) * w).sum().backward()
    out = dfb.snr2(ctx, tb.A[rows], t, d.detach())
    np.testing.assert_allclose(torch.view_as_complex(a_ri.grad).cpu().numpy(),
                               (out['grad'] * w[:, None]).cpu().numpy(), rtol=1e-6)
    np.testing.assert_allclose(d.grad.cpu().numpy(), (out['drho2_ddelta'] * w).cpu().numpy(), rtol=1e-6)


def test_refine_gain_off_grid_injection(setup):
    """Projected gradient ascent (20 steps) on the hull of the 8 nearest templates from the loudest trigger at the
    off-grid injection: the refined SNR is at least the nearest template's, and does not exceed the SNR of the exact
    injected waveform (float64 direct filter) by more than the noise can explain (2%)."""
    import differentiable as dfb
    import templates
    ds, plan, trig, near, st = setup['ds'], setup['plan'], setup['trig'], setup['near'], setup['st']
    tb = plan.tiles[0]
    ctx = dfb.ResponseContext(tb, ds.D[0], ds.IP[0], ds.df, ds.tlen, plan.kmin)
    i = int(np.argmax(np.where(near, trig['snr'], 0)))
    row = int(np.flatnonzero(tb.ids == trig['template_id'][i])[0])
    res = dfb.refine(ctx, tb.A, [row], [int(trig['sample'][i])], neighbours=8, steps=20)
    r0, r1 = float(res['rho2_start'][0]) ** 0.5, float(res['rho2'][0]) ** 0.5
    # the exact injected waveform, filtered in float64 over the same data: the optimum any template can reach
    m1, m2 = setup['m']
    H, _ = templates.generate(np.array([m1]), np.array([m2]), None, ds.flen, ds.df, 15.0)
    Hn = H.cpu().numpy()
    D = ds.D[0].cpu().numpy(); IP = ds.IP.cpu().numpy()
    sig, norm = core.sigmasq(Hn, IP, ds.df)
    zopt = core.direct_z(Hn.astype(np.complex64), D, ds.tlen)[0]
    w0 = int(trig['sample'][i])
    ropt = float(np.abs(zopt[w0 - 2048:w0 + 2048]).max() * norm[0, 0])
    print(f'nearest template SNR {r0:.3f}  refined {r1:.3f}  injected waveform {ropt:.3f}  '
          f'(gain {r1 / r0:.4f}, delta {float(res["delta"][0]):+.3f} samples)', flush=True)
    assert r1 >= r0 * (1 - 1e-6)
    assert r1 <= ropt * 1.02
