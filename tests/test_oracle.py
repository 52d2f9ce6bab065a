"""Pins the oracle (CPU): golden vectors written from the reference's own numpy functions, textbook routines (scipy),
closed forms and brute force on tiny inputs. Each check is chosen so that a plausible slip (an index, a sign, a
dropped term, a transposed operand) fails it."""
import json
import os

import numpy as np
import pytest
import scipy.signal

from oracle import core, pipeline

GOLD = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'reference_vectors.npz'))


def test_segments_and_spans_golden():
    for c, exp in zip(GOLD['seg_cases'], json.loads(str(GOLD['seg_json']))):
        n, sr, st, L, sp, ep, ts, te = c
        got = core.segment_slices(int(n), int(sr), float(st), int(L), int(sp), int(ep), int(ts), int(te))
        assert [[list(a), list(b)] for a, b in got] == exp
    for c, exp in zip(GOLD['span_cases'], json.loads(str(GOLD['span_json']))):
        assert [list(s) for s in core.overlapping_psd_spans(*[int(v) for v in c])] == exp


def test_median_bias_golden_and_limit():
    got = [core.median_bias(n) for n in range(1, 80)] + [core.median_bias(1001)]
    np.testing.assert_array_equal(got, GOLD['median_bias'])
    assert abs(core.median_bias(999) - np.log(2)) < 1e-3 and core.median_bias(3) == 1 + 1 / 3 - 1 / 2


def test_truncation_golden():
    P = GOLD['ist_psd']
    np.testing.assert_allclose(core.inverse_spectrum_truncation(P, 1 / 8.0, 512, 15.0), GOLD['ist_plain'], rtol=1e-10)
    np.testing.assert_allclose(core.inverse_spectrum_truncation(P, 1 / 8.0, 512, 15.0, 'hann'), GOLD['ist_hann'],
                               rtol=1e-10)
    np.testing.assert_allclose(core.inverse_spectrum_truncation(P, 1 / 8.0, 256, None), GOLD['ist_noflow'], rtol=1e-10)


def test_truncation_compact_support():
    """The truncated filter's impulse response (irfft of 1/sqrt(out) with the input's zero phase) vanishes outside
    +-L/2, and away from the cut the output tracks the input PSD."""
    P = GOLD['ist_psd']
    out = core.inverse_spectrum_truncation(P, 1 / 8.0, 512, 15.0)
    assert np.all(out > 0)
    mid = slice(400, 1500)
    assert np.median(np.abs(out[mid] / P[mid] - 1)) < 0.2


def test_welch_golden_float32_path():
    """The reference computes float32 PSDs from float32 input; the oracle is float64: agreement to float32 rounding."""
    x = GOLD['welch_x']
    np.testing.assert_allclose(core.welch(x, 4096, 2048, 1 / 2048.0)[0], GOLD['welch_med'], rtol=2e-5)
    np.testing.assert_allclose(core.welch(x, 4096, 4096, 1 / 2048.0, num_segments=6)[0], GOLD['welch_even'], rtol=2e-5)
    np.testing.assert_allclose(core.welch(x, 4096, 2048, 1 / 2048.0, avg='mean')[0], GOLD['welch_mean'], rtol=2e-5)


def test_welch_vs_scipy():
    """scipy.signal.welch with the same symmetric Hann window, density scaling and median average (scipy divides the
    median by the same bias series)."""
    rng = np.random.default_rng(0)
    x = rng.standard_normal(4096 * 9)
    for avg in ('median', 'mean'):
        f, p = scipy.signal.welch(x, fs=2048.0, window=np.hanning(4096), nperseg=4096, noverlap=2048, detrend=False,
                                  scaling='density', average=avg)
        q, df = core.welch(x, 4096, 2048, 1 / 2048.0, avg=avg)
        np.testing.assert_allclose(q, p, rtol=1e-10)
        assert df == f[1]


def test_welch_white_noise_level():
    """White noise of variance s^2 has one-sided PSD 2 s^2 dt."""
    rng = np.random.default_rng(1)
    s, dt = 3.0, 1 / 1024.0
    q, _ = core.welch(rng.standard_normal(1 << 20) * s, 1024, 512, dt, avg='mean')
    assert abs(np.mean(q[10:-10]) / (2 * s * s * dt) - 1) < 0.01


def test_interpolate_and_cluster_golden():
    np.testing.assert_allclose(core.interpolate(GOLD['interp_in'], 0.25, 1 / 64.0, 4097), GOLD['interp_out'],
                               rtol=1e-15)
    np.testing.assert_array_equal(core.cluster_over_window(GOLD['cl_idx'], GOLD['cl_vals'], 2048), GOLD['cl_keep'])


def test_gate_window_shape():
    w = core.gate_window(2048.0, 0.25, 0.25)
    assert len(w) == 2048 and w[0] == 1.0 and np.all(w[512:1536] == 0) and np.allclose(w[::-1], w)


def _tiny_bank():
    flen, df = 4097, 1 / 8.0
    m1 = np.array([30.0, 31.0, 33.0, 40.0, 41.0, 60.0]); m2 = np.array([25.0, 25.5, 24.0, 30.0, 30.0, 40.0])
    H, kmax = pipeline.spa_bank(m1, m2, None, flen, df, 15.0)
    f = np.arange(flen) * df
    ip = np.where(f >= 15, 1.0 / (1 + (f / 100.0) ** 2 + (30.0 / np.maximum(f, 1)) ** 4), 0.0)
    return H.astype(np.complex64), kmax, ip, df


def test_svd_invariants():
    """Vh rows orthonormal; conj(A) @ Vh reconstructs every sw-weighted template with relative residual^2 <= eps;
    with eps -> 0 the rank is the tile size and the reconstruction is exact."""
    H, kmax, ip, df = _tiny_bank()
    sw = np.sqrt(ip)
    kmin = int(15 / df)
    for eps in (1e-2, 1e-4, 0.0):
        r, A, Vh, worst = core.svd_tile(H, sw, kmin, eps)
        np.testing.assert_allclose(Vh @ Vh.conj().T, np.eye(r), atol=1e-9)
        Hw = H.astype(np.complex128) * sw
        res = np.linalg.norm(np.conj(A) @ Vh - Hw, axis=1) ** 2 / np.linalg.norm(Hw, axis=1) ** 2
        assert res.max() <= max(eps, 1e-12) * 1.0001 and worst <= max(eps, 1e-12)
    assert r == H.shape[0]


def test_compressed_snr_equals_direct_filter():
    """At full rank the basis series and stacked A rebuild the direct filter z exactly (fp_filter.py:592 identity),
    and the Re/Im stacking of stacked_A / stacked_Z gives the same complex product."""
    H, kmax, ip, df = _tiny_bank()
    N = 2 * (H.shape[1] - 1)
    rng = np.random.default_rng(3)
    D = ((rng.standard_normal(H.shape[1]) + 1j * rng.standard_normal(H.shape[1])) * ip).astype(np.complex64)
    sw = np.sqrt(ip); isw = np.where(sw > 0, 1 / np.maximum(sw, 1e-30), 0)
    r, A, Vh, _ = core.svd_tile(H, sw, int(15 / df), 0.0)
    zb = core.basis_series(Vh, D * isw, 100, 1000, N)
    zd = core.direct_z(H, D, N)[:, 100:1100]
    np.testing.assert_allclose(A @ zb, zd, rtol=1e-7, atol=1e-9 * np.abs(zd).max())
    Are, Aim = core.stacked_A(A, np.ones(len(A)), 2.0, 2 * r + 4)
    Z = core.stacked_Z(zb, 2.0, 2 * r + 4)
    np.testing.assert_allclose(Are @ Z + 1j * (Aim @ Z), A @ zb, rtol=1e-9, atol=1e-9 * np.abs(zd).max())


def test_direct_z_vs_dft_loop():
    rng = np.random.default_rng(4)
    N = 32
    h = rng.standard_normal((2, 17)) + 1j * rng.standard_normal((2, 17))
    d = rng.standard_normal(17) + 1j * rng.standard_normal(17)
    z = core.direct_z(h, d, N)
    for i in range(2):
        for t in range(N):
            s = sum(np.conj(h[i, k]) * d[k] * np.exp(2j * np.pi * k * t / N) for k in range(17))
            assert abs(z[i, t] - s) < 1e-10


def test_gemm_absmax_brute_force():
    rng = np.random.default_rng(5)
    M, K, Tn = 3, 5, 300
    Ar, Ai, Z = rng.standard_normal((M, K)), rng.standard_normal((M, K)), rng.standard_normal((K, Tn))
    Z[:, 140] = Z[:, 129]                      # a tie inside block 1
    V, I = core.gemm_absmax(Ar, Ai, Z, Tn)
    assert V.shape == (M, 3)
    for m in range(M):
        for b in range(3):
            best, bi = -1.0, b * 128
            for t in range(b * 128, min(Tn, (b + 1) * 128)):
                zr = sum(Ar[m, k] * Z[k, t] for k in range(K)); zi = sum(Ai[m, k] * Z[k, t] for k in range(K))
                p = zr * zr + zi * zi
                if p > best:
                    best, bi = p, t
            assert abs(V[m, b] - best) < 1e-9 * best and I[m, b] == bi


def test_window_max_and_peaks_brute_force():
    rng = np.random.default_rng(6)
    V = rng.integers(0, 3, (4, 12)).astype(float)
    I = rng.integers(0, 1000, (4, 12))
    vw, iw = core.window_max(V, I, 3, 4)
    for r in range(4):
        for w in range(3):
            j = max(range(4), key=lambda j: (V[r, w * 4 + j], -j))
            assert vw[r, w] == V[r, w * 4 + j] and iw[r, w] == I[r, w * 4 + j]
    z = (rng.integers(-2, 3, (2, 50)) + 1j * rng.integers(-2, 3, (2, 50))).astype(np.complex64)
    pm, pi = core.window_peaks(z, 3, 10, 4)
    for r in range(2):
        for w in range(4):
            seg = z[r, 3 + w * 10:3 + (w + 1) * 10]
            p = [float(v.real) ** 2 + float(v.imag) ** 2 for v in seg]      # small integers: exact
            j = p.index(max(p))
            assert pi[r, w] == j and pm[r, w] == np.float32(p[j])


def test_candidates_strict_threshold():
    vw = np.array([[25.0, 25.0000001, 30.0]]); iw = np.array([[5, 6, 7]])
    tb, wb, samp, apx = core.candidates(vw, iw, 5.0, 0.0, 100)
    assert list(wb) == [1, 2] and list(samp) == [106, 107]


def test_edges_brute_force():
    rng = np.random.default_rng(7)
    H = rng.standard_normal((2, 300)) + 1j * rng.standard_normal((2, 300))
    ips = rng.random(300)
    E = core.edges_per_template(np.array([0, 1]), np.array([4, 7]), 10, np.array([250, 290]), H, ips)
    for i, (nb, k1) in enumerate(((4, 250), (7, 290))):
        w = np.abs(H[i]) ** 2 * ips
        cs = np.cumsum(w); tot = cs[k1 - 1]
        for j in range(7):
            if j < nb:
                k = next(k for k in range(300) if cs[k] / tot > j / nb)
                assert E[i, j] == min(max(k, 10), k1)
            else:
                assert E[i, j] == k1
        assert E[i, 7] == k1


def test_bin_sums_sum_to_direct_z():
    """sum_j Z_j over bins covering the template's band equals the direct filter z(t); pw >= |z|^2 / nb
    (Cauchy-Schwarz), with equality only for equal bins."""
    H, kmax, ip, df = _tiny_bank()
    N = 2 * (H.shape[1] - 1)
    rng = np.random.default_rng(8)
    D = ((rng.standard_normal(H.shape[1]) + 1j * rng.standard_normal(H.shape[1])) * ip).astype(np.complex64)
    kmin = int(15 / df)
    rows = np.array([0, 2, 5]); t = np.array([7, 1000, N - 1])
    E = core.edges_per_template(rows, np.full(3, 16), kmin, kmax[rows], H, ip)
    pw, z = core.bin_power_sums(H, rows, t, E, want_z=True, d=D, n_time=N)
    zd = core.direct_z(H[rows], D, N)
    for i in range(3):
        assert abs(z[i] - zd[i, t[i]]) < 1e-9 * abs(zd[i]).max()
        assert pw[i] >= abs(z[i]) ** 2 / 16 * (1 - 1e-12)
    zz, pp = core.dense_float64(H, D, rows, t, E, np.full(6, 16), N)
    np.testing.assert_allclose(zz, z, rtol=1e-12); np.testing.assert_allclose(pp, pw, rtol=1e-12)


def test_spa_phase_time_frequency():
    """The stationary-phase time t(f) = psi'(f) / 2 pi of the 1PN chirp equals -(tau0 + tau2)(f), the textbook
    Newtonian and first post-Newtonian chirp times (Sathyaprakash & Dhurandhar 1991): pins the sign, the 0PN and
    1PN coefficients and the mass combinations."""
    m1, m2 = 12.0, 7.0
    M = (m1 + m2) * pipeline.MSUN_S; eta = m1 * m2 / (m1 + m2) ** 2
    for f in (20.0, 45.0, 90.0):
        h = 1e-4 * f
        t = (pipeline.spa_phase(f + h, m1, m2) - pipeline.spa_phase(f - h, m1, m2)) / (2 * h) / (2 * np.pi)
        tau0 = 5.0 / (256 * np.pi * f * eta) * (np.pi * M * f) ** (-5.0 / 3)
        tau2 = 5.0 / (192 * np.pi * f * eta) * (743.0 / 336 + 11.0 * eta / 4) / (np.pi * M * f)
        assert abs(t + tau0 + tau2) < 1e-6 * (tau0 + tau2)
    # eccentric term: at f = f0 it is -(2355/1462) e0^2 times the Newtonian term
    f = 15.0; e0 = 0.2
    newt = 3.0 / (128 * eta * (np.pi * M * f) ** (5.0 / 3))
    d = pipeline.spa_phase(f, m1, m2, e0) - pipeline.spa_phase(f, m1, m2)
    assert abs(d / newt + 2355.0 / 1462 * e0 * e0) < 1e-12


def test_newsnr_closed_form():
    s = np.array([8.0, 8.0, 8.0]); rc = np.array([0.5, 1.0, 2.0])
    ns = core.newsnr(s, rc)
    assert ns[0] == 8 and ns[1] == 8 and abs(ns[2] - 8 / (4.5 ** (1 / 6))) < 1e-12


def test_planted_mistake_is_caught():
    """A checker that would miss a conjugate slip pins nothing: flipping conj in the direct filter must break the
    full-rank identity above."""
    H, kmax, ip, df = _tiny_bank()
    N = 2 * (H.shape[1] - 1)
    rng = np.random.default_rng(9)
    D = ((rng.standard_normal(H.shape[1]) + 1j * rng.standard_normal(H.shape[1])) * ip).astype(np.complex64)
    good = core.direct_z(H, D, N)
    bad = np.fft.ifft(np.pad(H.astype(np.complex128) * D, ((0, 0), (0, N - H.shape[1]))), axis=1) * N
    assert np.abs(bad - good).max() > 1e-3 * np.abs(good).max()


def test_find_pairs_brute_force_small():
    """The brute-force slide search on a hand-built case: a pair 0.3 s apart is found at slide -3 with dt 0 and a
    pair across the span end is found through the wrap."""
    H = dict(x=np.array([10.0, 0.005]), tid=np.array([1, 2]))
    L = dict(x=np.array([10.3, 99.995, 50.0]), tid=np.array([1, 2, 1]))
    h, l, k, dt = core.find_pairs(H, L, 100.0, 0.1, 10, 0.015)
    assert list(zip(h, l, k)) == [(0, 0, -3), (1, 1, 0)]
    assert abs(dt[0]) < 1e-12 and abs(dt[1] + 0.01) < 1e-9


def _stage2_cases():
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), 'golden'))
    import make_golden as mg
    meta = json.loads(str(GOLD['s2_json']))
    for (name, seed, sr, seconds, pad, opts), m in zip(mg.STAGE2, meta):
        o = {opts[i][2:].replace('-', '_'): opts[i + 1] for i in range(0, len(opts), 2)}
        h = mg.stage2_strain(seed, sr, seconds, pad)
        s = h[int(pad * sr):len(h) - int(pad * sr)]
        args = (int(o['segment_length']), int(o['segment_start_pad']), int(o['segment_end_pad']),
                float(o['psd_segment_length']), float(o['psd_segment_stride']), int(o['psd_num_segments']),
                float(o['psd_inverse_length']), 15.0)
        yield name, s, sr, mg, m, args, o.get('invpsd_trunc_method')


def _check_stage2(name, D, IP, analyze, seg_start, meta, stride):
    """D and IP on the golden bin subsample (the reference's float32 PSD path against a float64 one: 2e-5 per IP
    bin, 1e-5 of max |D|), the exact zero pattern around kmin, and the segment layout exactly."""
    gD, gIP = GOLD[f's2_{name}_D'], GOLD[f's2_{name}_IP']
    nz = gIP > 0
    assert np.array_equal(IP[:, ::stride] > 0, nz)
    assert np.max(np.abs(IP[:, ::stride] - gIP)[nz] / gIP[nz]) < 2e-5
    assert np.max(np.abs(D[:, ::stride] - gD)) < 1e-5 * np.abs(gD).max()
    kmin = int(15.0 / meta['df'])
    assert np.array_equal(IP[:, kmin - 32:kmin + 32] > 0, GOLD[f's2_{name}_IPk'] > 0)
    assert not np.any(D[:, kmin - 32:kmin]) and np.all(IP[:, kmin - 32:kmin] == 0)
    assert [list(a) for a in analyze] == meta['analyze'] and list(seg_start) == meta['seg_start']


def test_condition_vs_reference_stage2_golden():
    """core.condition against the reference's own gpu_stage2 (numpy backend) on seeded strains: one PSD span, and a
    1,100 s strain with four PSD spans and five segments whose largest-overlap picks are 0, 0, 2, 3, 3 (a pick moved
    to the next span changes IP by about 2x, far beyond the tolerance), with the Hann-tapered truncation."""
    for name, s, sr, mg, meta, args, trunc in _stage2_cases():
        r = core.condition(s, sr, mg.STAGE2_GPS, *args, trunc_method=trunc)
        _check_stage2(name, r['D'], r['IP'], r['analyze'], r['seg_start'], meta, mg.STAGE2_STRIDE)
        if name == 'multi_span':
            seg = core.segment_slices(len(s), sr, mg.STAGE2_GPS, *args[:3])
            spans = core.overlapping_psd_spans(len(s), int(args[3] * sr), int(args[4] * sr), args[5])
            assert core.pick_psd(seg, spans) == [0, 0, 2, 3, 3]
            # a planted pick (the next span) is caught by the golden IP
            alt = [(p + 1) % len(spans) for p in core.pick_psd(seg, spans)]
            P = r['psd'][alt]
            ip2 = np.where(P > 0, 1.0 / np.where(P > 0, P, 1.0), 0.0)
            ip2[:, :int(15.0 / r['df'])] = 0
            gIP = GOLD['s2_multi_span_IP']
            nz = gIP > 0
            assert np.max(np.abs(ip2[:, ::mg.STAGE2_STRIDE] - gIP)[nz] / gIP[nz]) > 0.1


def test_pick_psd_rules():
    """Largest overlap wins, the first span wins ties, a segment touching no span gets None."""
    segs = [((0, 100), None), ((50, 150), None), ((90, 200), None), ((1000, 1100), None)]
    spans = [(0, 120), (40, 160), (80, 200)]
    # overlaps: seg0 120/60/20 -> 0; seg1 70/100/70 -> 1; seg2 30/70/110 -> 2; seg3 none
    assert core.pick_psd(segs, spans) == [0, 1, 2, None]
    assert core.pick_psd([((10, 60), None)], [(0, 40), (30, 70)]) == [0]        # 30 vs 30: first wins


def test_gate_taper_hand_case():
    """pycbc gate_data at non-integer sample times near both edges: zeros start exactly at int(t_start sr) + pad,
    the ramps are 0.5 (1 + cos(pi i / pad)) and its mirror, samples outside the window keep their value, and a gate
    starting before the data clips at sample 0."""
    sr, n, gps = 256.0, 4096, 1000000000.0
    gw, pw = 0.25, 0.125
    win, pad = int(2 * sr * (gw + pw)), int(sr * pw)                   # 192, 32
    for gt_rel in (3.3017, 15.89):                                     # offsets 748 (near mid) and 3972 (right edge)
        x = np.full(n, 2.0, np.float32)
        core.gate_taper(x, sr, gps, [(gps + gt_rel, gw, pw)])
        off = int((gt_rel - gw - pw) * sr)
        ramp = 0.5 * (1.0 + np.cos(np.pi * np.arange(pad) / pad))
        exp = np.full(n, 2.0)
        w = np.concatenate([ramp, np.zeros(win - 2 * pad), ramp[::-1]])
        i2 = min(win, n - off)
        exp[off:off + i2] = 2.0 * w[:i2]
        np.testing.assert_array_equal(x, exp.astype(np.float32))
        assert x[off + pad] == 0 and x[off + pad - 1] > 0 and x[off - 1] == 2.0
    x = np.full(n, 2.0, np.float32)
    core.gate_taper(x, sr, gps, [(gps + 0.1, gw, pw)])                 # t_start = -0.275 s: offset -70
    off = int((0.1 - gw - pw) * sr)
    assert off == -70
    w = np.concatenate([0.5 * (1 + np.cos(np.pi * np.arange(pad) / pad)), np.zeros(win - 2 * pad),
                        0.5 * (1 + np.cos(np.pi * np.arange(pad) / pad))[::-1]])
    exp = np.full(n, 2.0)
    exp[0:win + off] = 2.0 * w[-off:]
    np.testing.assert_array_equal(x, exp.astype(np.float32))
    y = np.full(n, 2.0, np.float32)
    core.gate_taper(y, sr, gps, [(gps + 40.0, gw, pw), (gps - 5.0, gw, pw)])       # both outside the data
    assert np.all(y == 2.0)


def test_white_noise_snr_normalisation():
    """sigma^2 = 4 df sum |h|^2 IP and norm = 4 df / sigma make rho = |z| norm have <rho^2> = 2 on Gaussian noise of
    the PSD 1/IP (D = stilde IP, E|stilde_k|^2 = S_k / (2 df)); a factor 2 slip in sigma^2, df or the one-sided sum
    gives 1 or 4."""
    rng = np.random.default_rng(5)
    N, sr = 1 << 15, 2048.0
    flen, df = N // 2 + 1, sr / N
    f = np.arange(flen) * df
    import synth
    S = synth.analytic_psd(f)
    kmin = int(15.0 / df)
    IP = np.where(np.arange(flen) >= kmin, 1.0 / S, 0.0)
    H, _ = pipeline.spa_bank(np.array([10.0, 30.0, 7.0]), np.array([8.0, 25.0, 6.0]), None, flen, df, 15.0)
    sig, norm = core.sigmasq(H, IP[None, :], df)
    acc = []
    for _ in range(24):
        st = np.sqrt(S / (4 * df)) * (rng.standard_normal(flen) + 1j * rng.standard_normal(flen))
        z = core.direct_z(H, st * IP, N)
        acc.append((np.abs(z) ** 2 * norm[:, 0][:, None] ** 2).mean(1))
    m = np.mean(acc, 0)
    assert np.all(np.abs(m - 2.0) < 0.06), m


def test_resample_ldas_golden():
    """The 'ldas' FIR decimation by 2 and 4 equals the reference's resample_ldas bit for bit (float64, same tap
    order), including the zeroed start-up samples and the tail past the data."""
    for f in (2, 4):
        got = core.resample_ldas(GOLD['resample_x'], f)
        np.testing.assert_array_equal(got, GOLD[f'resample_{f}'])
        ntaps = 20 * f + 1
        assert np.all(got[:(2 * (ntaps // 2) - (ntaps + 1) // 2 + f - 1) // f] == 0) and got[-1] == 0


def test_coarse_grid_layout():
    """pipeline.coarse_grid: ascending unique nodes with both ends; its chirp part uniform in f^(-5/3)."""
    x = pipeline.coarse_grid(15.0, 1000.0, n_chirp=300, n_log=100)
    assert x[0] == 15.0 and x[-1] == 1000.0 and np.all(np.diff(x) > 0)
    assert len(x) <= 402
    u = np.linspace(15.0 ** (-5 / 3), 1000.0 ** (-5 / 3), 300) ** (-3 / 5)
    assert np.isin(u[1:-1], x).mean() > 0.99


def test_fine_interp_reproduces_cubic_phase_and_linear_amplitude():
    """pipeline.fine_interp on nodes of a global cubic phase (expanded at each node: c0 = phi, c1 = phi', c2 =
    phi''/2, c3 = phi'''/6) and a linear amplitude gives a(f) exp(i phi(f)) at every bin, clipped to the node span:
    a wrong interval (search side, clip) or a wrong t or d fails it."""
    x = np.array([20.0, 21.5, 23.0, 30.0, 31.0, 40.0])
    al, be, ga, de = 3e-3, -0.2, 7.0, 0.4
    ph = lambda f: ((al * f + be) * f + ga) * f + de
    dph = lambda f: (3 * al * f + 2 * be) * f + ga
    coef = np.stack([np.full(5, al), (3 * al * x[:-1] + be), dph(x[:-1]), ph(x[:-1])], -1)[None]
    amp = (2.0 + 0.1 * x)[None]
    df, kmin, kend, flen = 0.25, 76, 170, 200          # f from 19 Hz (below x_0: clipped) to 42.25 Hz (above x_J)
    out = pipeline.fine_interp(amp, coef, x, df, kmin, kend, flen)
    f = np.clip(np.arange(kmin, kend) * df, x[0], x[-1])
    np.testing.assert_allclose(out[0, kmin:kend], (2.0 + 0.1 * f) * np.exp(1j * ph(f)), rtol=1e-12, atol=1e-12)
    assert not out[0, :kmin].any() and not out[0, kend:].any()
