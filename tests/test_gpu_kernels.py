"""Parity of each HIP kernel, called through the C ABI, against the oracle on the same seeded inputs."""
import numpy as np
import pytest
import torch

from oracle import core, pipeline

pytestmark = pytest.mark.gpu


def _bf16_round(x):
    return torch.from_numpy(np.asarray(x, np.float32)).to(torch.bfloat16).to(torch.float32).numpy()


def _operands(rng, M, K, Tn, dtype):
    Ar = rng.standard_normal((M, K)).astype(np.float32)
    Ai = rng.standard_normal((M, K)).astype(np.float32)
    Z = rng.standard_normal((K, Tn)).astype(np.float32)
    if dtype in (torch.bfloat16, torch.float16):
        rd = lambda x: torch.from_numpy(x).to(dtype).to(torch.float32).numpy()
        Ar, Ai, Z = rd(Ar), rd(Ai), rd(Z)
    return Ar, Ai, Z


@pytest.mark.parametrize('dtype', [torch.float32, torch.bfloat16, torch.float16])
@pytest.mark.parametrize('M,K,Tn,variant', [(45, 64, 5 * 128 + 37, 0), (128, 40, 1024, 0), (300, 256, 2048, 0),
                                             (77, 520, 1000, 0), (130, 64, 1024, 1), (1, 16, 128, 1),
                                             (77, 520, 1000, 2), (200, 840, 1500, 3), (33, 72, 300, 3)])
def test_fused_peaks_vs_oracle(gpu, dtype, M, K, Tn, variant):
    """fp_gemm.fused_peaks (mode 0) against the float64 GEMM + block max of the same (rounded) operands. Tolerance:
    relative 1e-4 on the block maxima (fp32 accumulation); the chosen sample must reach the float64 max to 1e-4."""
    import fp_gemm
    rng = np.random.default_rng(M * 7 + K)
    Ar, Ai, Z = _operands(rng, M, K, Tn, dtype)
    V, I = fp_gemm.fused_peaks(torch.from_numpy(Ar).cuda().to(dtype), torch.from_numpy(Ai).cuda().to(dtype),
                               torch.from_numpy(Z).cuda().to(dtype), M, Tn, (16, variant))
    V, I = V.cpu().numpy(), I.cpu().numpy()
    Vr, Ir = core.gemm_absmax(Ar, Ai, Z, Tn)
    assert V.shape == Vr.shape and I.shape == Ir.shape
    np.testing.assert_allclose(V, Vr, rtol=1e-4)
    zr = Ar.astype(np.float64) @ Z; zi = Ai.astype(np.float64) @ Z
    pw = zr * zr + zi * zi
    got = np.take_along_axis(pw, I.astype(np.int64), 1)
    assert np.all(got >= Vr * (1 - 1e-4))
    assert np.all((I // 128) == np.arange(I.shape[1])[None, :])
    assert (I == Ir).mean() > 0.99


@pytest.mark.parametrize('code,variant', [(1, 0), (1, 1), (1, 2), (1, 3), (1, 4), (1, 5), (2, 0), (2, 1), (2, 3),
                                          (2, 4), (2, 5), (3, 0), (3, 1), (3, 4), (3, 5)])
def test_stacked_operands_rotated_path(gpu, code, variant):
    """Complex-stacked operands (Ai derived from Ar, as stack_a writes them) through the rotated K loops
    (mode | STACKED, Ai never read: a garbage Ai is passed) against the float64 GEMM of the same values, and equal
    window maxima and samples to the general path that reads Ai, up to fp32 summation order (relative 1e-5; bit for
    bit for the template-side rotation of variants 0, 4 and 5)."""
    import fp_gemm
    rng = np.random.default_rng(17 + variant + 10 * code)
    M, r, cw, nw = 150, (185 if code == 2 and variant in (0, 5) else 197), 2048, 3   # fp8 DMA: K % 64 == 0
    K = 2 * (-(-r // 16) * 16)
    Tn = cw * nw
    A = (rng.standard_normal((M, r)) + 1j * rng.standard_normal((M, r))) * (20 if code == 2 else 1)
    Zc = (rng.standard_normal((Tn, r)) + 1j * rng.standard_normal((Tn, r))) * (10 if code == 2 else 1)
    Ar = np.zeros((M, K), np.float32); Ai = np.zeros((M, K), np.float32); Zt = np.zeros((Tn, K), np.float32)
    Ar[:, 0:2 * r:2] = A.real; Ar[:, 1:2 * r:2] = -A.imag
    Ai[:, 0:2 * r:2] = A.imag; Ai[:, 1:2 * r:2] = A.real
    Zt[:, 0:2 * r:2] = Zc.real; Zt[:, 1:2 * r:2] = Zc.imag
    cast = (lambda x: _fp8(np.clip(x, -448, 448))) if code == 2 else \
        (lambda x: torch.from_numpy(x).to(torch.float16 if code == 3 else torch.bfloat16))
    qa, qi, qz = cast(Ar), cast(Ai), cast(Zt)
    da, di, dz = (q.to(torch.float32).numpy().astype(np.float64) for q in (qa, qi, qz))
    view = (lambda t: t.view(torch.uint8)) if code == 2 else (lambda t: t)
    qa, qi, qz = (view(q).cuda() for q in (qa, qi, qz))
    per = cw // 128
    items = fp_gemm.items_table([(0, Tn // 128, 0, 0, Tn)], 2 * per)
    out = {}
    for mode, ai in ((1 | fp_gemm.STACKED, torch.full_like(qi, 0x7b)), (1, qi)):
        V = torch.empty((M, nw), device='cuda'); I = torch.empty((M, nw), dtype=torch.int32, device='cuda')
        fp_gemm.gemm_peaks_multi(code, qz, K, qa, ai, K, M, K, items, mode, per, (2 * per, variant if mode > 1 else 2),
                                 V, I)
        out[mode] = (V.cpu().numpy(), I.cpu().numpy())
    V0, I0 = core.gemm_absmax(da, di, dz.T, Tn)
    Vr, Ir = core.window_max(V0, I0, nw, per)
    V, I = out[1 | fp_gemm.STACKED]
    np.testing.assert_allclose(V, Vr, rtol=1e-4)
    np.testing.assert_allclose(V, out[1][0], rtol=1e-5)
    if variant in (0, 4, 5):
        # the template-side rotation forms exactly the Ai the general path reads: the same products in the same order
        np.testing.assert_array_equal(V, out[1][0])
        np.testing.assert_array_equal(I, out[1][1])
    pw = (da @ dz.T) ** 2 + (di @ dz.T) ** 2
    assert np.all(np.take_along_axis(pw, I.astype(np.int64), 1) >= Vr * (1 - 1e-4))


@pytest.mark.parametrize('dtype', [torch.float32, torch.bfloat16])
@pytest.mark.parametrize('variant', [0, 1])
def test_window_mode_equals_block_max(gpu, dtype, variant):
    """The fused window epilogue (mode 1) must equal, bit for bit, the window max of the same kernel's per-block
    output (mode 0) taken by the oracle's first-index rule and by the GPU block_window_max kernel."""
    import fp_gemm, peak_kernel
    rng = np.random.default_rng(3)
    M, K, cw, nw = 200, 96, 2048, 6
    Tn = cw * nw
    Ar, Ai, Z = _operands(rng, M, K, Tn, dtype)
    args = [torch.from_numpy(a).cuda().to(dtype) for a in (Ar, Ai, Z)]
    V, I = fp_gemm.fused_peaks(*args, M, Tn, (16, variant))
    vw, iw = fp_gemm.fused_window_peaks(*args, M, Tn, cw, (32, variant))
    vo, io = core.window_max(V.cpu().numpy(), I.cpu().numpy(), nw, cw // 128)
    assert np.array_equal(vw.cpu().numpy(), vo) and np.array_equal(iw.cpu().numpy(), io)
    vg, ig = peak_kernel.block_window_max(V, I, nw, cw // 128)
    assert np.array_equal(vg.cpu().numpy(), vo) and np.array_equal(ig.cpu().numpy(), io)


def test_window_max_ties_first_index(gpu):
    """Planted ties: equal block maxima inside a window -> the first block's sample (fp_filter.py:745 with the tie
    rule of peak_kernel.py:21-33)."""
    import peak_kernel
    rng = np.random.default_rng(5)
    n, nw, per = 7, 9, 16
    V = rng.integers(0, 4, (n, nw * per)).astype(np.float32)
    I = (np.arange(nw * per)[None, :] * 128 + rng.integers(0, 128, (n, nw * per))).astype(np.int32)
    vg, ig = peak_kernel.block_window_max(torch.from_numpy(V).cuda(), torch.from_numpy(I).cuda(), nw, per)
    vo, io = core.window_max(V, I, nw, per)
    assert np.array_equal(vg.cpu().numpy(), vo) and np.array_equal(ig.cpu().numpy(), io)


def test_window_peaks_bit_exact(gpu):
    """peak_kernel.window_peaks against the oracle on identical complex64 input, with planted ties and a window
    shorter than the block (rows 5, cw 3000 and 100)."""
    import peak_kernel
    rng = np.random.default_rng(1)
    for cw, nwin, a0 in ((3000, 4, 17), (100, 30, 0)):
        z = (rng.standard_normal((5, a0 + cw * nwin + 50)) + 1j * rng.standard_normal((5, a0 + cw * nwin + 50)))
        z = z.astype(np.complex64)
        z[1, a0 + 10] = z[1, a0 + 40] = 9 + 9j           # tie inside window 0 of row 1
        z[2, a0 + cw + 5] = z[2, a0 + cw + 7] = 7 - 7j
        pm, pi = peak_kernel.window_peaks(torch.from_numpy(z).cuda(), a0, cw, nwin)
        po, io = core.window_peaks(z, a0, cw, nwin)
        assert np.array_equal(pm.cpu().numpy(), po) and np.array_equal(pi.cpu().numpy(), io)


def test_select_above_order(gpu):
    import peak_kernel
    rng = np.random.default_rng(2)
    for total in (0, 1, 4095, 4097, 100000):
        v = rng.random(total).astype(np.float32)
        v[::97] = 0.5
        g = peak_kernel.select_above(torch.from_numpy(v).cuda(), 0.5)
        assert np.array_equal(g, np.flatnonzero(v > np.float32(0.5)))
        # the one-pass candidate list: indices, their window samples and values
        iw = rng.integers(0, 2048, total).astype(np.int32)
        g2, s2, v2 = peak_kernel.select_peaks(torch.from_numpy(v).cuda(), torch.from_numpy(iw).cuda(), 0.5).result()
        exp = np.flatnonzero(v > np.float32(0.5))
        assert np.array_equal(g2, exp) and np.array_equal(s2, iw[exp]) and np.array_equal(v2, v[exp])


def test_chisq_kernel_vs_float64_definition(gpu):
    """tests/test_chisq_kernel_gpu.py of the reference, tightened: N = 2^21, 4 rows, random complex q, 40 random edges
    per row, rng(1); sum_j |Z_j|^2 against the float64 cumulative-sum definition. The reference asserts rel < 1e-3;
    this asserts rel < 1e-5."""
    import chisq_kernel
    rng = np.random.default_rng(1)
    N = 2 ** 21; F = N // 2 + 1; nrow = 4
    qh = np.zeros((nrow, N), np.complex64)
    qh[:, :F] = (rng.standard_normal((nrow, F)) + 1j * rng.standard_normal((nrow, F))).astype(np.complex64)
    rows = np.array([0, 0, 0, 1, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 3])
    times = rng.integers(0, N, len(rows))
    nb = 40
    edges_row = np.sort(rng.choice(np.arange(30 * 1024, F), size=(nrow, nb + 1), replace=False), axis=1)
    E = edges_row[rows]
    got, gz = chisq_kernel.bin_power_sums(torch.from_numpy(qh).cuda(), rows, times, E, want_z=True)
    ref, rz = core.bin_power_sums(qh, rows, times, E, want_z=True)
    assert np.max(np.abs(got - ref) / ref) < 1e-5
    assert np.max(np.abs(gz - rz) / np.abs(rz)) < 1e-4


@pytest.mark.parametrize('N', [3 * 2 ** 17, 2 ** 18])
def test_chisq_fused_any_length(gpu, N):
    """Fused mode (conj(h) d formed in the kernel) with a time length that need not be a power of two, against the
    float64 direct sums; edges include an empty bin and duplicate peaks."""
    import chisq_kernel
    rng = np.random.default_rng(N % 1000)
    F = N // 2 + 1
    H = (rng.standard_normal((3, F)) + 1j * rng.standard_normal((3, F))).astype(np.complex64)
    d = (rng.standard_normal(F) + 1j * rng.standard_normal(F)).astype(np.complex64)
    rows = np.array([0, 1, 1, 2, 2, 2, 1])
    times = rng.integers(0, N, len(rows)); times[-1] = times[1]
    e = np.sort(rng.choice(np.arange(100, F), size=(3, 17), replace=False), axis=1)
    e[1, 5] = e[1, 4]                                   # an empty bin
    E = e[rows]
    pw, z = chisq_kernel.bin_power_sums(torch.from_numpy(H).cuda(), rows, times, E, want_z=True,
                                        d=torch.from_numpy(d).cuda(), n_time=N)
    rp, rz = core.bin_power_sums(H, rows, times, E, want_z=True, d=d, n_time=N)
    assert np.max(np.abs(pw - rp) / rp) < 1e-5
    assert np.max(np.abs(z - rz) / np.abs(rz)) < 1e-4


def test_edges_vs_float64_cumsum(gpu):
    """Edges kernel against multi_filter.edges_per_template's float64 definition: within 1 bin, at most 1e-3 of edges
    differing (tests/test_exact_stage.py:224-225)."""
    import chisq_edges
    rng = np.random.default_rng(4)
    F = 65537; kmin = 960
    f = np.maximum(np.arange(F), 1) / 64.0
    H = np.zeros((6, F), np.complex64)
    kmax = np.array([30000, 65536, 12000, 50000, 2000, 40000])
    for i in range(6):
        H[i, kmin:kmax[i]] = (f[kmin:kmax[i]] ** (-7 / 6) * np.exp(1j * rng.uniform(0, 6, kmax[i] - kmin)))
    ips = (1.0 / (1 + (f / 200.0) ** 2) * (np.arange(F) >= kmin)).astype(np.float32)
    rows = np.array([0, 1, 2, 3, 4, 5, 3])
    nb = np.array([16, 16, 16, 8, 4, 16, 8])
    E = chisq_edges.edges_blocked(None, rows, nb, kmin, kmax[rows], torch.from_numpy(H).cuda(),
                                  torch.from_numpy(ips).cuda())
    Er = core.edges_per_template(rows, nb, kmin, kmax[rows], H, ips)
    assert E.shape == Er.shape
    d = np.abs(E - Er)
    assert d.max() <= 1 and (d > 0).mean() <= 1e-3


def test_templates_vs_oracle(gpu):
    """GPU stationary-phase chirps against the float64 oracle: float32 rounding of the same formula (<= 2e-6 of the
    row's peak amplitude); identical end bins."""
    import templates
    import synth
    cb = synth.circular_bank(12, seed=3)
    eb = synth.eccentric_bank(8, seed=4)
    m1 = np.r_[cb.m1, eb.m1]; m2 = np.r_[cb.m2, eb.m2]; e0 = np.r_[cb.e0, eb.e0]
    flen, df = 65537, 1 / 64
    H, kmax = templates.generate(m1, m2, e0, flen, df, 15.0)
    Hr, kr = pipeline.spa_bank(m1, m2, e0, flen, df, 15.0)
    assert np.array_equal(kmax, kr)
    Hg = H.cpu().numpy()
    err = np.abs(Hg - Hr).max(1) / np.abs(Hr).max(1)
    assert err.max() < 2e-6


def test_sigmasq_vs_oracle(gpu):
    import templates
    rng = np.random.default_rng(6)
    H = (rng.standard_normal((9, 4097)) + 1j * rng.standard_normal((9, 4097))).astype(np.complex64)
    IP = rng.random((3, 4097)).astype(np.float32)
    sig, norm = templates.sigmasq(torch.from_numpy(H).cuda(), torch.from_numpy(IP).cuda(), 0.25)
    sr, nr = core.sigmasq(H, IP, 0.25)
    np.testing.assert_allclose(sig.cpu().numpy(), sr, rtol=1e-12)
    np.testing.assert_allclose(norm.cpu().numpy(), nr, rtol=1e-12)
    # rows zero outside [kmin, kend): the range-limited sum over several chunks equals the full-row oracle
    H2 = (rng.standard_normal((600, 40000)) + 1j * rng.standard_normal((600, 40000))).astype(np.complex64)
    H2[:, :123] = 0; H2[:, 31000:] = 0
    IP2 = rng.random((17, 40000)).astype(np.float32)
    sig2, _ = templates.sigmasq(torch.from_numpy(H2).cuda(), torch.from_numpy(IP2).cuda(), 0.5, 123, 31000)
    np.testing.assert_allclose(sig2.cpu().numpy(), core.sigmasq(H2, IP2, 0.5)[0], rtol=1e-12)


def _fp8(x):
    """Host e4m3 (OCP) rounding of float32 values already within +-448, as torch rounds (nearest even)."""
    return torch.from_numpy(np.asarray(x, np.float32)).to(torch.float8_e4m3fn)


@pytest.mark.parametrize('M,K,Tn,mode,variant', [(77, 128, 1024, 0, 1), (130, 784, 2048, 1, 1), (5, 48, 300, 0, 1),
                                                 (130, 784, 2048, 1, 2), (77, 208, 1024, 0, 3)])
def test_fp8_gemm_vs_oracle(gpu, M, K, Tn, mode, variant):
    """fp8 e4m3 operands (bytes rounded on the host) against the float64 GEMM of the same dequantized values, with a
    per-template row scale applied to |z|^2. Products of e4m3 values are exact in fp32: relative 1e-4."""
    import fp_gemm
    rng = np.random.default_rng(K)
    Ar, Ai = (rng.standard_normal((M, K)) * 30).astype(np.float32), (rng.standard_normal((M, K)) * 30).astype(np.float32)
    Z = (rng.standard_normal((K, Tn)) * 2).astype(np.float32)
    qa, qi, qz = _fp8(Ar), _fp8(Ai), _fp8(Z)
    da, di, dz = (q.to(torch.float32).numpy().astype(np.float64) for q in (qa, qi, qz))
    rs = (rng.random(M) + 0.5).astype(np.float32)
    if mode == 0:
        V, I = fp_gemm.fused_peaks(qa.cuda(), qi.cuda(), qz.cuda(), M, Tn, (16, variant),
                                   rowscale=torch.from_numpy(rs).cuda())
        Vr, Ir = core.gemm_absmax(da, di, dz, Tn)
    else:
        V, I = fp_gemm.fused_window_peaks(qa.cuda(), qi.cuda(), qz.cuda(), M, Tn, 2048, (16, variant),
                                          rowscale=torch.from_numpy(rs).cuda())
        V0, I0 = core.gemm_absmax(da, di, dz, Tn)
        Vr, Ir = core.window_max(V0, I0, Tn // 2048, 16)
    np.testing.assert_allclose(V.cpu().numpy(), Vr * rs[:, None], rtol=1e-4)
    pw = (da @ dz) ** 2 + (di @ dz) ** 2
    got = np.take_along_axis(pw, I.cpu().numpy().astype(np.int64), 1)
    assert np.all(got >= Vr * (1 - 1e-4))


def test_fp8_operand_conversion(gpu):
    """Device e4m3 rounding (pack_zt, stack_a_fp8) equals torch's host rounding, with saturation at +-448 and the
    row scale 224 / max |row|."""
    import _dev, _lib
    rng = np.random.default_rng(11)
    rp, tlen, a0, Tn = 16, 4096, 100, 1000
    Zc = (rng.standard_normal((rp, tlen)) + 1j * rng.standard_normal((rp, tlen))).astype(np.complex64) * 50
    Zc[3, a0 + 7] = 900 + 1j                           # saturates
    cs = (rng.random(rp) + 0.5).astype(np.float32)
    Zt = torch.empty((Tn, 2 * rp), dtype=torch.uint8, device='cuda')
    csd = torch.from_numpy(cs).cuda()
    Zcd = torch.from_numpy(Zc).cuda()                  # kept alive until the kernel has run
    _lib.call('gwgs_pack_zt', 2, Zcd.data_ptr(), tlen, rp, a0, Tn, Tn, csd.data_ptr(), Zt.data_ptr(), _dev.stream())
    ref = Zc[:, a0:a0 + Tn] * cs[:, None]
    ref2 = np.empty((Tn, 2 * rp), np.float32); ref2[:, 0::2] = ref.real.T; ref2[:, 1::2] = ref.imag.T
    exp = _fp8(np.clip(ref2, -448, 448)).view(torch.uint8).numpy()
    assert np.array_equal(Zt.cpu().numpy(), exp)
    n, r, K = 9, 13, 32
    A = (rng.standard_normal((n, r)) + 1j * rng.standard_normal((n, r))).astype(np.complex64)
    rs = (rng.random(n) + 0.1).astype(np.float32)
    Ar = torch.empty((n, K), dtype=torch.uint8, device='cuda'); Ai = torch.empty_like(Ar)
    pws = torch.empty(n, dtype=torch.float32, device='cuda')
    Ad, rsd = torch.from_numpy(A).cuda(), torch.from_numpy(rs).cuda()
    _lib.call('gwgs_stack_a_fp8', Ad.data_ptr(), n, r, rsd.data_ptr(), None, K, Ar.data_ptr(), Ai.data_ptr(),
              pws.data_ptr(), _dev.stream())
    As = A * rs[:, None]
    ra = 224.0 / np.maximum(np.abs(As.real), np.abs(As.imag)).max(1)
    np.testing.assert_allclose(pws.cpu().numpy(), 1 / ra ** 2, rtol=1e-6)
    Xr = np.zeros((n, K), np.float32); Xr[:, 0:2 * r:2] = As.real * ra[:, None]; Xr[:, 1:2 * r:2] = -As.imag * ra[:, None]
    got = Ar.cpu().view(torch.float8_e4m3fn).to(torch.float32).numpy()
    exp = _fp8(Xr).to(torch.float32).numpy()
    assert np.abs(got - exp).max() <= 1e-6 * 448


@pytest.mark.parametrize('dtype,variant', [(torch.bfloat16, 0), (torch.bfloat16, 1), (torch.float32, 0),
                                           (torch.bfloat16, 2), (torch.bfloat16, 3)])
def test_multi_segment_launch_equals_single(gpu, dtype, variant):
    """One launch over three segments of different lengths (item table, per-segment row scales) equals the
    single-segment launches, bit for bit."""
    import fp_gemm
    rng = np.random.default_rng(9)
    M, K, cw = 150, 64, 2048
    nws = [3, 1, 5]
    Tns = [nw * cw for nw in nws]
    Ar, Ai, _ = _operands(rng, M, K, 8, dtype)
    Zs = [_operands(rng, 1, K, T, dtype)[2].T.copy() for T in Tns]
    scales = (rng.random((3, M)) + 0.5).astype(np.float32)
    A_r, A_i = torch.from_numpy(Ar).cuda().to(dtype), torch.from_numpy(Ai).cuda().to(dtype)
    Zt = torch.from_numpy(np.concatenate(Zs)).cuda().to(dtype)
    row0 = np.r_[0, np.cumsum(Tns)]; col0 = np.r_[0, np.cumsum(nws)]
    code = fp_gemm.DTYPES[dtype]
    items = fp_gemm.items_table([(int(row0[i]), Tns[i] // 128, int(col0[i]), i, Tns[i]) for i in range(3)], 32)
    V = torch.empty((M, sum(nws)), device='cuda'); I = torch.empty((M, sum(nws)), dtype=torch.int32, device='cuda')
    fp_gemm.gemm_peaks_multi(code, Zt, K, A_r, A_i, K, M, K, items, 1, 16, (32, variant), V, I,
                             torch.from_numpy(scales).cuda())
    for i in range(3):
        v, ii = fp_gemm.fused_window_peaks(A_r, A_i, Zt[row0[i]:row0[i + 1]], M, Tns[i], cw, (32, variant), zt=True,
                                           rowscale=torch.from_numpy(scales[i].copy()).cuda())
        assert np.array_equal(V[:, col0[i]:col0[i + 1]].cpu().numpy(), v.cpu().numpy())
        assert np.array_equal(I[:, col0[i]:col0[i + 1]].cpu().numpy(), ii.cpu().numpy())


def test_time_slides_vs_brute_force(gpu):
    """GPU slide coincidences against the brute-force restatement of coinc_rank.find_pairs: same pairs, order,
    slides and dt, including wrap-around at the span ends."""
    import slides
    rng = np.random.default_rng(12)
    P, step, K, window = 1000.0, 0.1, 100, 0.015
    H = dict(x=rng.uniform(0, P, 400), tid=rng.integers(0, 5, 400))
    L = dict(x=rng.uniform(0, P, 300), tid=rng.integers(0, 5, 300))
    L['x'][:5] = H['x'][:5] + 0.3                      # planted coincidences at slide 3
    L['x'][5] = P - 0.01; H['x'][5] = 0.04; L['tid'][5] = H['tid'][5]     # across the wrap
    h, l, k, dt = slides.find_pairs(H, L, P, step, K, window)
    hr, lr, kr, dr = core.find_pairs(H, L, P, step, K, window)
    assert len(h) == len(hr) > 5
    assert np.array_equal(h, hr) and np.array_equal(l, lr) and np.array_equal(k, kr)
    np.testing.assert_allclose(dt, dr, atol=1e-12)


@pytest.mark.parametrize('n,r,kmin,kend,a0,Tn', [
    (1 << 13, 11, 120, 2049, 512, 4096),              # 64 x 128: 16 s chunks band-limited to 256 Hz
    (1 << 14, 23, 240, 2000, 1024, 12288),            # 128 x 128, 2 of 16 column inputs per thread
    (1 << 15, 40, 480, (1 << 14) + 1, 2048, 12288),   # 128 x 256: 16 s chunks at 2048 Hz, the whole band
    (1 << 16, 37, 5, 20000, 300, 1 << 15),            # 256 x 256
    (1 << 17, 16, 0, (1 << 16) + 1, 0, 1 << 17),      # 256 x 512, whole series
    (1 << 19, 9, 3840, 120000, 4096, 65536),          # 512 x 1024
    (1 << 20, 21, 7680, 230000, 98304, 917504),       # 1024 x 1024, a 512 s segment at 2048 Hz
    (1 << 20, 35, 10240, 128240, 98304, 917504),      # the same below 250 Hz: 4 of 32 column inputs per thread
    (1 << 21, 3, 100, 500000, 5, 1000),               # 1024 x 2048
    (1 << 22, 2, 0, 1 << 21, (1 << 22) - 4096, 4096),  # 2048 x 2048, the last rows
])
def test_basis_series_vs_oracle(gpu, n, r, kmin, kend, a0, Tn):
    """The fused four-step transform (gwgs_basis_series) against the oracle's float64 basis series of the same Vc and
    D (fp_filter.py:597-613): fp32 output within 2e-6 rms and 2e-5 max of the series' rms; zero columns for rows
    r..rp; the bf16 and fp8 (with per-basis scales) outputs are the device's own fp32 values rounded to nearest even,
    bit for bit."""
    import _dev, _lib
    rng = np.random.default_rng(n + r)
    flen = n // 2 + 1
    rp = -(-r // 16) * 16
    K = 2 * rp
    Vc = np.zeros((rp, flen), np.complex64)
    Vc[:r, kmin:kend] = (rng.standard_normal((r, kend - kmin)) + 1j * rng.standard_normal((r, kend - kmin)))
    D = (rng.standard_normal(flen) + 1j * rng.standard_normal(flen)).astype(np.complex64)
    ref = core.basis_series(np.conj(Vc[:r].astype(np.complex128)), D.astype(np.complex128), a0, Tn, n).T  # (Tn, r)
    lib = _lib.load()
    ws = lib.gwgs_basis_series_workspace(r, n)
    assert ws == r * n * 8
    work = torch.empty(ws, dtype=torch.uint8, device='cuda')
    Vd, Dd = torch.from_numpy(Vc).cuda(), torch.from_numpy(D).cuda()
    cs = (rng.random(rp) * 0.02 + 0.01).astype(np.float32)
    csd = torch.from_numpy(cs).cuda()
    outs = {}
    for code, dt in ((0, torch.float32), (1, torch.bfloat16), (2, torch.uint8)):
        Zt = torch.full((Tn, K), 7, dtype=dt, device='cuda')
        _lib.call('gwgs_basis_series', code, Vd.data_ptr(), flen, r, rp, Dd.data_ptr(), kmin, kend, n, a0, Tn,
                  csd.data_ptr() if code == 2 else None, 0, None, work.data_ptr(), Zt.data_ptr(), K, _dev.stream())
        outs[code] = Zt.cpu()
    z32 = outs[0].numpy()
    got = z32[:, 0:2 * r:2].astype(np.float64) + 1j * z32[:, 1:2 * r:2]
    rms = np.sqrt(np.mean(np.abs(ref) ** 2))
    err = np.abs(got - ref)
    assert np.sqrt(np.mean(err ** 2)) < 2e-6 * rms and err.max() < 2e-5 * rms, (err.max() / rms)
    assert np.all(z32[:, 2 * r:] == 0)
    assert torch.equal(outs[1], outs[0].to(torch.bfloat16))
    scaled = np.clip(z32 * np.repeat(cs, 2)[None, :], -448, 448)
    assert np.array_equal(outs[2].numpy(), _fp8(scaled).view(torch.uint8).numpy())
    # fp16 intermediate (with a per-row scale): float32 output within 5e-4 rms / 5e-3 max of the series' rms
    ysc = torch.from_numpy((rng.random(r) * 4 + 0.25).astype(np.float32)).cuda()
    Zh = torch.full((Tn, K), 7, dtype=torch.float32, device='cuda')
    _lib.call('gwgs_basis_series', 0, Vd.data_ptr(), flen, r, rp, Dd.data_ptr(), kmin, kend, n, a0, Tn, None, 1,
              ysc.data_ptr(), work.data_ptr(), Zh.data_ptr(), K, _dev.stream())
    zh = Zh.cpu().numpy()
    eh = np.abs(zh[:, 0:2 * r:2].astype(np.float64) + 1j * zh[:, 1:2 * r:2] - ref)
    assert np.sqrt(np.mean(eh ** 2)) < 5e-4 * rms and eh.max() < 5e-3 * rms, (eh.max() / rms)
    assert np.all(zh[:, 2 * r:] == 0)
    # the filter's bf16 form: fp16 intermediate and bf16 output, within the bf16 rounding (2^-9 of each value) plus
    # the intermediate's; padding columns exactly zero
    Zb = torch.full((Tn, K), 7, dtype=torch.bfloat16, device='cuda')
    _lib.call('gwgs_basis_series', 1, Vd.data_ptr(), flen, r, rp, Dd.data_ptr(), kmin, kend, n, a0, Tn, None, 1,
              ysc.data_ptr(), work.data_ptr(), Zb.data_ptr(), K, _dev.stream())
    zb = Zb.float().cpu().numpy()
    eb = np.abs(zb[:, 0:2 * r:2].astype(np.float64) + 1j * zb[:, 1:2 * r:2] - ref)
    assert np.sqrt(np.mean(eb ** 2)) < 3e-3 * rms and eb.max() < 3e-2 * rms, (eb.max() / rms)
    assert np.all(zb[:, 2 * r:] == 0)


def test_basis_series_empty_band(gpu):
    """kmin == kend: every input is outside the band, so the series (and the padding rows) are exactly zero."""
    import _dev, _lib
    n, r, rp, flen = 1 << 17, 5, 16, (1 << 16) + 1
    Vd = torch.ones((rp, flen), dtype=torch.complex64, device='cuda')
    Dd = torch.ones(flen, dtype=torch.complex64, device='cuda')
    work = torch.empty(_lib.load().gwgs_basis_series_workspace(r, n), dtype=torch.uint8, device='cuda')
    for kmin in (0, 3000):
        Zt = torch.full((4096, 2 * rp), 7, dtype=torch.float32, device='cuda')
        _lib.call('gwgs_basis_series', 0, Vd.data_ptr(), flen, r, rp, Dd.data_ptr(), kmin, kmin, n, 100, 4096, None,
                  0, None, work.data_ptr(), Zt.data_ptr(), 2 * rp, _dev.stream())
        assert torch.count_nonzero(Zt).item() == 0


def test_basis_series_padding_rows(gpu):
    """rp well above r (a whole row block past the last basis row, read through an empty buffer): the padding
    columns are zero and the live ones equal those of a call with the smallest padding, bit for bit, for the
    complex64 and the fp16 intermediate."""
    import _dev, _lib
    rng = np.random.default_rng(11)
    n, r, flen, kmin, kend = 1 << 17, 3, (1 << 16) + 1, 40, 30000
    Vc = np.zeros((64, flen), np.complex64)
    Vc[:r, kmin:kend] = rng.standard_normal((r, kend - kmin)) + 1j * rng.standard_normal((r, kend - kmin))
    D = (rng.standard_normal(flen) + 1j * rng.standard_normal(flen)).astype(np.complex64)
    Vd, Dd = torch.from_numpy(Vc).cuda(), torch.from_numpy(D).cuda()
    work = torch.empty(_lib.load().gwgs_basis_series_workspace(r, n), dtype=torch.uint8, device='cuda')
    ysc = torch.full((64,), 0.5, device='cuda')
    Tn, a0 = 8192, 1000
    for ytype in (0, 1):
        outs = []
        for rp in (16, 48):
            Zt = torch.full((Tn, 2 * rp), 7, dtype=torch.bfloat16, device='cuda')
            _lib.call('gwgs_basis_series', 1, Vd.data_ptr(), flen, r, rp, Dd.data_ptr(), kmin, kend, n, a0, Tn, None,
                      ytype, ysc.data_ptr() if ytype else None, work.data_ptr(), Zt.data_ptr(), 2 * rp, _dev.stream())
            outs.append(Zt.cpu())
        assert torch.count_nonzero(outs[1][:, 2 * r:]).item() == 0
        assert torch.equal(outs[0][:, :2 * r], outs[1][:, :2 * r])
        assert torch.count_nonzero(outs[0][:, :2 * r]).item() > 0


@pytest.mark.parametrize('n,ytype,dtype', [(1 << 18, 1, 1), (1 << 20, 1, 1), (1 << 18, 0, 0), (1 << 16, 1, 2)])
def test_basis_series_batch_equals_per_segment(gpu, n, ytype, dtype):
    """gwgs_basis_series_batch over segments of different data, analysed spans, output rows and fp16 scales (one
    launch of each pass) equals one gwgs_basis_series call per segment, bit for bit, and leaves the Zt rows between
    the segments untouched."""
    import ctypes
    import _dev, _lib
    rng = np.random.default_rng(n + ytype)
    flen, r, kmin = n // 2 + 1, 37, 60
    kend = n // 20                                        # a band below n / 16: the column pass's pruned inputs
    rp, K = 48, 96
    Vc = np.zeros((rp, flen), np.complex64)
    Vc[:r, kmin:kend] = rng.standard_normal((r, kend - kmin)) + 1j * rng.standard_normal((r, kend - kmin))
    Vd = torch.from_numpy(Vc).cuda()
    nseg = 5
    Ds = [torch.from_numpy((rng.standard_normal(flen) + 1j * rng.standard_normal(flen)).astype(np.complex64)).cuda()
          for _ in range(nseg)]
    a0 = [n // 8, 0, n // 4 + 128, n // 2, 1000]
    Tn = [n // 2, n // 8, n // 4, n // 2 - 256, 4096]
    gap = 7
    zrow = list(np.cumsum([0] + [t + gap for t in Tn])[:nseg])
    rows = int(zrow[-1] + Tn[-1] + gap)
    ysc = [torch.full((rp,), float(v), device='cuda') for v in rng.random(nseg) * 4 + 0.25]
    cs = torch.from_numpy((rng.random(rp) * 0.02 + 0.01).astype(np.float32)).cuda() if dtype == 2 else None
    dt = {0: torch.float32, 1: torch.bfloat16, 2: torch.uint8}[dtype]
    lib = _lib.load()
    one = torch.empty(lib.gwgs_basis_series_workspace(r, n), dtype=torch.uint8, device='cuda')
    work = torch.empty(nseg * r * n * (4 if ytype else 8), dtype=torch.uint8, device='cuda')
    ref = torch.full((rows, K), 7, dtype=dt, device='cuda')
    for i in range(nseg):
        _lib.call('gwgs_basis_series', dtype, Vd.data_ptr(), flen, r, rp, Ds[i].data_ptr(), kmin, kend, n, a0[i],
                  Tn[i], _dev.ptr(cs), ytype, ysc[i].data_ptr() if ytype else None, one.data_ptr(),
                  ref[zrow[i]].data_ptr(), K, _dev.stream())
    got = torch.full((rows, K), 7, dtype=dt, device='cuda')
    arr = lambda ct, v: (ct * nseg)(*v)
    _lib.call('gwgs_basis_series_batch', dtype, Vd.data_ptr(), flen, r, rp, nseg,
              arr(ctypes.c_void_p, [d.data_ptr() for d in Ds]), kmin, kend, n, arr(ctypes.c_long, a0),
              arr(ctypes.c_long, Tn), arr(ctypes.c_long, [int(z) for z in zrow]), _dev.ptr(cs), ytype,
              arr(ctypes.c_void_p, [y.data_ptr() if ytype else None for y in ysc]), work.data_ptr(), got.data_ptr(), K,
              _dev.stream())
    assert torch.equal(got.view(torch.uint8) if dtype == 2 else got, ref.view(torch.uint8) if dtype == 2 else ref)
    assert torch.count_nonzero(got[:Tn[0]].float() if dtype != 2 else got[:Tn[0]]).item() > 0


def test_vd_power_vs_float64(gpu):
    """gwgs_vd_power: sum over [kmin, kend) of |Vc D|^2 per row, against float64 (relative 1e-12)."""
    import _dev, _lib
    rng = np.random.default_rng(5)
    r, flen, kmin, kend = 7, 5000, 30, 4100
    Vc = (rng.standard_normal((r, flen)) + 1j * rng.standard_normal((r, flen))).astype(np.complex64)
    D = (rng.standard_normal(flen) + 1j * rng.standard_normal(flen)).astype(np.complex64)
    out = torch.empty(r, dtype=torch.float64, device='cuda')
    Vd, Dd = torch.from_numpy(Vc).cuda(), torch.from_numpy(D).cuda()
    _lib.call('gwgs_vd_power', Vd.data_ptr(), flen, r, Dd.data_ptr(), kmin, kend, out.data_ptr(), _dev.stream())
    exp = (np.abs(Vc[:, kmin:kend].astype(np.complex128) * D[kmin:kend]) ** 2).sum(1)
    np.testing.assert_allclose(out.cpu().numpy(), exp, rtol=1e-12)


def test_autotuner_picks_a_running_config(gpu):
    """fp_gemm.pick_cfg (the reference's sweep, fp_gemm.py:47-64) on a small stacked shape: every candidate runs,
    the chosen one is among the fastest, and every variant gives the same window maxima and samples to fp32
    summation order (variants differ only in how operands reach the MFMAs)."""
    import fp_gemm
    cfg, res = fp_gemm.pick_cfg(M=96, K=160, Tn=8 * 2048, reps=1)
    times = {k: v for k, v in res.items() if isinstance(v, float)}
    assert len(times) == len(res), res
    assert res[str(cfg)] == min(times.values())
    # in list mode: tiles per item only
    cfg2, res2 = fp_gemm.pick_cfg(M=96, K=160, Tn=8 * 2048, reps=1, lists=True)
    assert all(isinstance(v, float) for v in res2.values()), res2
    assert {eval(k)[1] for k in res2} == {fp_gemm.VARIANTS[-1]} and res2[str(cfg2)] == min(res2.values())
    rng = np.random.default_rng(3)
    M, r, Tn, per = 96, 80, 4 * 2048, 16
    K = 2 * r
    A = rng.standard_normal((M, r)) + 1j * rng.standard_normal((M, r))
    Ar = np.zeros((M, K), np.float32); Ai = np.zeros((M, K), np.float32)
    Ar[:, 0::2], Ar[:, 1::2] = A.real, -A.imag
    Ai[:, 0::2], Ai[:, 1::2] = A.imag, A.real
    Z = rng.standard_normal((Tn, K)).astype(np.float32)
    qa, qi, qz = (torch.from_numpy(x).cuda().to(torch.bfloat16) for x in (Ar, Ai, Z))
    outs = []
    for variant in fp_gemm.VARIANTS:
        V = torch.empty((M, Tn // 2048), device='cuda'); I = torch.empty_like(V, dtype=torch.int32)
        items = fp_gemm.items_table([(0, Tn // 128, 0, 0, Tn)], 32)
        fp_gemm.gemm_peaks_multi(1, qz, K, qa, qi, K, M, K, items, 1 | fp_gemm.STACKED, per, (32, variant), V, I)
        outs.append((V.cpu().numpy(), I.cpu().numpy()))
    for V, I in outs[1:]:
        np.testing.assert_allclose(V, outs[0][0], rtol=1e-5)
        assert (I == outs[0][1]).mean() > 0.99


def _window_local_maxima(p, Tn, cw, nw, lm):
    """Per (row, window) the lm largest local maxima of p over the window's valid samples (t < Tn): p[t] >= p[t - 1]
    and p[t] > p[t + 1], the window's first sample and its last valid sample open on that side; larger value first,
    then smaller sample; (-1, -1) padding."""
    M = p.shape[0]
    VL = np.full((M, nw * lm), -1.0, np.float32)
    IL = np.full((M, nw * lm), -1, np.int32)
    for w in range(nw):
        lo, hi = w * cw, min((w + 1) * cw, Tn)
        if hi <= lo:
            continue
        q = p[:, lo:hi]
        left = np.ones_like(q, bool); left[:, 1:] = q[:, 1:] >= q[:, :-1]
        right = np.ones_like(q, bool); right[:, :-1] = q[:, :-1] > q[:, 1:]
        for m in range(M):
            t = np.flatnonzero(left[m] & right[m])
            o = np.lexsort((t, -q[m, t]))[:lm]
            VL[m, w * lm:w * lm + len(o)] = q[m, t[o]]
            IL[m, w * lm:w * lm + len(o)] = lo + t[o]
    return VL, IL


@pytest.mark.parametrize('code,stacked,r,sep,gate', [(1, True, 32, 0, 0), (1, True, 24, 0, 0), (1, False, 32, 0, 0),
                                                     (0, False, 32, 0, 0), (2, True, 32, 0, 0), (3, True, 32, 0, 0),
                                                     (3, True, 24, 0, 0), (1, True, 32, 16, 0), (0, False, 32, 5, 0),
                                                     (3, True, 24, 40, 0), (1, True, 32, 0, 0.6),
                                                     (0, False, 32, 0, 0.8), (1, True, 32, 0, -0.7)])
def test_local_maxima_lists(gpu, code, stacked, r, sep, gate):
    """gemm_peaks_multi_lm over three segments (one with a partial last window, rows past its end holding large
    values that must be masked): integer operands make every |z|^2 exact in fp32 whatever the summation order, so
    the window maxima equal the mode-1 launch's bit for bit, and with sep = 0 the local-maximum lists equal their
    definition bit for bit, ties (frequent with integers) included; with a gate, so do the entries at or above gate x
    the window maximum (gate < 0: gate |gate| with a floor at the windows' median maximum, entries at or above
    |gate| x max(window maximum, floor)). With sep > 0: every entry is a local maximum, entries are more than sep apart and sorted, and
    the first is the largest local maximum (earliest on ties). The alternates kernel then equals its rule on the same
    lists."""
    import fp_gemm
    import peak_kernel
    import _lib, _dev
    rng = np.random.default_rng(40 + code + 7 * r + stacked + sep + int(10 * abs(gate)))
    M, cw, LM = 70, 1024, fp_gemm.LM
    K = 2 * r
    nws, Tns = [3, 1, 2], [3 * cw, cw, 2 * cw - 300]
    per = cw // 128
    A = rng.integers(-3, 4, (M, r)) + 1j * rng.integers(-3, 4, (M, r))
    Ar = np.zeros((M, K), np.float32); Ai = np.zeros((M, K), np.float32)
    Ar[:, 0::2] = A.real; Ar[:, 1::2] = -A.imag
    Ai[:, 0::2] = A.imag; Ai[:, 1::2] = A.real
    Zs = []
    for nw, Tn in zip(nws, Tns):
        Zc = rng.integers(-3, 4, (nw * cw, r)) + 1j * rng.integers(-3, 4, (nw * cw, r))
        Zc[Tn:] = 100 + 100j                         # past the segment's end: never reported
        z = np.zeros((nw * cw, K), np.float32); z[:, 0::2] = Zc.real; z[:, 1::2] = Zc.imag
        Zs.append(z)
    Zt = np.concatenate(Zs)
    scales = (2.0 ** rng.integers(-2, 3, (3, M))).astype(np.float32)
    cast = {0: lambda x: torch.from_numpy(x), 1: lambda x: torch.from_numpy(x).to(torch.bfloat16),
            2: lambda x: _fp8(x).view(torch.uint8), 3: lambda x: torch.from_numpy(x).to(torch.float16)}[code]
    qa, qi, qz = (cast(x).cuda() for x in (Ar, Ai, Zt))
    row0 = np.r_[0, np.cumsum([nw * cw for nw in nws])]; col0 = np.r_[0, np.cumsum(nws)]
    items = fp_gemm.items_table([(int(row0[i]), nws[i] * per, int(col0[i]), i, Tns[i]) for i in range(3)], 2 * per)
    nwsum = int(col0[-1])
    V = torch.empty((M, nwsum), device='cuda'); I = torch.empty((M, nwsum), dtype=torch.int32, device='cuda')
    VL = torch.empty((M, nwsum * LM), device='cuda'); IL = torch.empty((M, nwsum * LM), dtype=torch.int32,
                                                                         device='cuda')
    sc = torch.from_numpy(scales).cuda()
    ai = torch.full_like(qi, 0x7b) if stacked and code != 0 else qi
    V1 = torch.empty_like(V); I1 = torch.empty_like(I)
    fp_gemm.gemm_peaks_multi(code, qz, K, qa, qi, K, M, K, items, 1, per, (2 * per, 2), V1, I1, sc)
    floor = float(np.median(V1.cpu().numpy())) if gate < 0 else 0.0
    gate = abs(gate)
    fp_gemm.gemm_peaks_multi_lm(code, qz, K, qa, ai, K, M, K, items, stacked, per, sep, (2 * per, 0), V, I, VL, IL,
                                sc, gate=gate, floor=floor)
    torch.cuda.synchronize()
    assert np.array_equal(V.cpu().numpy(), V1.cpu().numpy())
    assert np.array_equal(I.cpu().numpy(), I1.cpu().numpy())
    vl, il = VL.cpu().numpy(), IL.cpu().numpy()
    for i in range(3):
        z = Zt[row0[i]:row0[i + 1]].astype(np.float64)
        p = ((Ar.astype(np.float64) @ z.T) ** 2 + (Ai.astype(np.float64) @ z.T) ** 2) * scales[i][:, None]
        p = p.astype(np.float32)
        sl = slice(col0[i] * LM, col0[i + 1] * LM)
        if sep == 0:
            ev, ei = _window_local_maxima(p, Tns[i], cw, nws[i], LM)
            if gate > 0:
                # entries at or above gate x the window maximum (its first entry) must match; others may be absent
                top = np.repeat(np.maximum(ev.reshape(M, nws[i], LM)[:, :, 0], np.float32(floor)), LM,
                                1).reshape(ev.shape)
                keep = ev >= np.float32(gate) * top
                g_keep = vl[:, sl] >= np.float32(gate) * top
                assert np.array_equal(keep, g_keep), f'segment {i} gated entries'
                assert np.array_equal(np.where(keep, vl[:, sl], -1), np.where(keep, ev, -1))
                assert np.array_equal(np.where(keep, il[:, sl], -1), np.where(keep, ei, -1))
                continue
            assert np.array_equal(vl[:, sl], ev), f'segment {i} values'
            assert np.array_equal(il[:, sl], ei), f'segment {i} samples'
            continue
        ev, ei = _window_local_maxima(p, Tns[i], cw, nws[i], cw)     # every local maximum, sorted
        gv, gi = vl[:, sl].reshape(M, nws[i], LM), il[:, sl].reshape(M, nws[i], LM)
        ev, ei = ev.reshape(M, nws[i], cw), ei.reshape(M, nws[i], cw)
        for m in range(M):
            for w in range(nws[i]):
                ok = gi[m, w] >= 0
                lmset = dict(zip(ei[m, w][ei[m, w] >= 0].tolist(), ev[m, w][ei[m, w] >= 0].tolist()))
                assert ok[0] and (gv[m, w, 0], gi[m, w, 0]) == (ev[m, w, 0], ei[m, w, 0])
                assert all(lmset.get(int(t)) == v for t, v in zip(gi[m, w][ok], gv[m, w][ok]))
                t_ok = np.sort(gi[m, w][ok])
                assert np.all(np.diff(t_ok) > sep)
                vv = gv[m, w][ok]
                assert np.all((vv[:-1] > vv[1:]) | ((vv[:-1] == vv[1:]) & (gi[m, w][ok][:-1] < gi[m, w][ok][1:])))
    # the alternates of a candidate list on these lists
    cut = float(np.quantile(V.cpu().numpy(), 0.7))
    idx, samp, val, cnt = peak_kernel.select_launch(V, I, cut)
    w, na, R, ratio, xcut = idx.numel(), LM - 1, 3, 0.9, float(np.quantile(V.cpu().numpy(), 0.8))
    alt = torch.empty((w, na), dtype=torch.int32, device='cuda')
    _lib.call('gwgs_window_alternates', VL.data_ptr(), IL.data_ptr(), VL.stride(0), LM, idx.data_ptr(),
              samp.data_ptr(), val.data_ptr(), cnt.data_ptr(), w, nwsum, cw, float(np.float32(xcut)),
              float(np.float32(ratio)), R,
              na, alt.data_ptr(), _dev.stream())
    torch.cuda.synchronize()
    n = int(cnt.cpu()[0])
    assert n > 10
    g, s0, v0, a = (x.cpu().numpy()[:n] for x in (idx, samp, val, alt))
    for c in range(n):
        exp = []
        if v0[c] >= np.float32(xcut):
            m, col = divmod(int(g[c]), nwsum)
            for k in range(LM):
                t, v = il[m, col * LM + k], vl[m, col * LM + k]
                if t >= 0 and v >= np.float32(ratio) * v0[c] and abs(t - s0[c]) > R and len(exp) < na:
                    exp.append(t % cw)
        assert list(a[c]) == exp + [-1] * (na - len(exp)), c
    # the second pass: one-window items for the windows holding candidates at or above xcut, built on the device,
    # and the list-mode launch over them with the item count in device memory list those windows as above
    hseg = fp_gemm.hot_segments(row0, col0, Tns)
    flags = torch.empty(nwsum, dtype=torch.int32, device='cuda')
    it2 = torch.empty(nwsum * fp_gemm.ITEM.itemsize, dtype=torch.uint8, device='cuda')
    nit = torch.full((1,), 99, dtype=torch.int32, device='cuda')
    hcut = np.float32(np.sort(V.cpu().numpy().max(0))[nwsum - 2])      # the two loudest windows hot
    fp_gemm.hot_items(idx, val, cnt, nwsum, float(hcut), hseg, per, flags, it2, nit)
    VL2 = torch.full_like(VL, -7.0); IL2 = torch.full_like(IL, -7)
    fp_gemm.gemm_peaks_multi_lm_dyn(code, qz, K, qa, ai, K, M, K, it2, nit, nwsum, stacked, per, sep, (2 * per, 0),
                                    torch.empty_like(V), torch.empty_like(I), VL2, IL2, sc, gate=gate, floor=floor)
    torch.cuda.synchronize()
    hot = np.unique(g[v0 >= hcut] % nwsum)
    k = int(nit.cpu()[0])
    assert k == len(hot) and 0 < k < nwsum
    tab = it2.cpu().numpy()[:k * fp_gemm.ITEM.itemsize].view(fp_gemm.ITEM)
    tab = tab[np.argsort(tab['col0'] * 1_000_000 + tab['t_first'])]
    si = np.searchsorted(col0, hot, side='right') - 1
    assert np.array_equal(tab['zrow0'], row0[si]) and np.array_equal(tab['seg'], si)
    assert np.array_equal(tab['t_first'], (hot - col0[si]) * per)
    assert np.array_equal(tab['t_end'] - tab['t_first'], np.full(k, per))
    assert np.array_equal(tab['Tn'], np.asarray(Tns)[si])
    vl2, il2 = VL2.cpu().numpy().reshape(M, nwsum, LM), IL2.cpu().numpy().reshape(M, nwsum, LM)
    cold = np.setdiff1d(np.arange(nwsum), hot)
    assert np.array_equal(vl2[:, hot], vl.reshape(M, nwsum, LM)[:, hot])
    assert np.array_equal(il2[:, hot], il.reshape(M, nwsum, LM)[:, hot])
    assert np.all(vl2[:, cold] == -7.0) and np.all(il2[:, cold] == -7)
    # a device count above max_items is clamped to it: the first k items only
    VL3 = torch.full_like(VL, -7.0); IL3 = torch.full_like(IL, -7)
    big = torch.full((1,), k + 1000, dtype=torch.int32, device='cuda')
    fp_gemm.gemm_peaks_multi_lm_dyn(code, qz, K, qa, ai, K, M, K, it2, big, k, stacked, per, sep, (2 * per, 0),
                                    torch.empty_like(V), torch.empty_like(I), VL3, IL3, sc, gate=gate, floor=floor)
    torch.cuda.synchronize()
    assert np.array_equal(VL3.cpu().numpy(), VL2.cpu().numpy()) and np.array_equal(IL3.cpu().numpy(),
                                                                                    IL2.cpu().numpy())


def test_climb_points_vs_float64(gpu):
    """gwgs_climb_points against the same climb on float64 exact sums: data holding a template at a known sample
    plus noise; starts on either slope of a broad peak (reached from far: several re-centred ranges), at the peak,
    with bounds that cut the climb off before the peak (the climb stops on the bound), and on a wide band (two
    phasor runs per lane) with a sharp peak."""
    import _lib, _dev
    rng = np.random.default_rng(5)
    N, nseg, nedge, W = 32768, 3, 17, 3
    flen = N // 2 + 1
    k = np.arange(flen)
    bands = [(40, 400), (40, 400), (40, 16000)]
    amp = np.stack([rng.random(flen) * ((k >= a) & (k < b)) for a, b in bands])
    H = (amp * np.exp(1j * rng.uniform(0, 2 * np.pi, (nseg, flen)))).astype(np.complex64)
    t0 = np.array([1000, 20500, 30000])
    D = np.stack([H[s] * np.exp(-2j * np.pi * k * t0[s] / N) +
                  0.05 * (rng.standard_normal(flen) + 1j * rng.standard_normal(flen)) for s in range(nseg)])
    D = D.astype(np.complex64)
    E = np.stack([np.linspace(a, b, nedge) for a, b in bands]).astype(np.int32)
    rows, segs, start, lo, hi = [], [], [], [], []
    for s in range(nseg):
        for off, (a, b) in ((-25, (-100, 100)), (-7, (-100, 100)), (0, (-100, 100)), (4, (-100, 100)),
                            (30, (-100, 100)), (-40, (-100, -10)), (40, (12, 100))):
            rows.append(s); segs.append(s); start.append(t0[s] + off); lo.append(t0[s] + a); hi.append(t0[s] + b)
    rows, segs, start, lo, hi = (np.array(x) for x in (rows, segs, start, lo, hi))
    Hd, Dd = torch.from_numpy(H).cuda(), torch.from_numpy(D).cuda()
    tab = np.array([Dd[s].data_ptr() for s in range(nseg)], np.int64)
    r_d, s_d, e_d, st_d, lo_d, hi_d, tab_d, E_d = _dev.to_dev_many(rows.astype(np.int32), segs.astype(np.int32),
                                                                   segs.astype(np.int32), start, lo, hi, tab, E)
    out = torch.empty(len(rows), dtype=torch.int64, device='cuda')
    _lib.call('gwgs_climb_points', Hd.data_ptr(), flen, tab_d.data_ptr(), N, r_d.data_ptr(), s_d.data_ptr(),
              E_d.data_ptr(), nedge, e_d.data_ptr(), st_d.data_ptr(), lo_d.data_ptr(), hi_d.data_ptr(), W, 1024,
              len(rows), out.data_ptr(), _dev.stream())
    got = out.cpu().numpy()
    prod = np.conj(H.astype(np.complex128)) * D.astype(np.complex128)

    def climb(s, c, l, h):
        a, b = bands[s]
        kk = k[a:b]
        while True:
            ts = np.clip(c + np.arange(-W, W + 1), l, h - 1)
            z = np.abs(prod[s, a:b] @ np.exp(2j * np.pi * np.outer(kk, ts) / N)) ** 2
            j = int(np.argmax(z))
            if (j == 0 and ts[0] > l) or (j == 2 * W and ts[-1] < h - 1):
                c = ts[j]
                continue
            return ts[j]

    exp = np.array([climb(s, c, l, h) for s, c, l, h in zip(segs, start, lo, hi)])
    assert np.array_equal(got, exp), (got, exp)
    for s in (0, 1):                      # the broad peaks: every start on the hill reaches the top
        g = got[segs == s]
        assert np.array_equal(g[:5], [t0[s]] * 5) and g[5] == t0[s] - 11 and g[6] == t0[s] + 12, (s, g)
