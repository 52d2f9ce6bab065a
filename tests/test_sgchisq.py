"""Sine-Gaussian chi-squared (fp_filter.py:463-486, sg_kernel.py:66-81) on the CPU: the oracle against a per-bin loop
and a closed form, the host tile arithmetic against the oracle's, the location parsing."""
import math

import numpy as np

from oracle import sgchisq as osg

PROD_TILES = [(20.0, float(o)) for o in (15, 30, 45, 60, 75, 90, 105, 120)]


def _seg(rng, N, sr=2048.0):
    flen = N // 2 + 1
    D = (rng.standard_normal(flen) + 1j * rng.standard_normal(flen)) * 1e-3
    IP = rng.uniform(0.5, 2.0, flen)
    return D, IP, sr / N


def test_oracle_matches_loop():
    rng = np.random.default_rng(4)
    N = 4096
    D, IP, df = _seg(rng, N)
    for trial in range(12):
        t = int(rng.integers(0, N))
        fpeak = float(rng.uniform(30, 600))
        vals = PROD_TILES[:int(rng.integers(1, 9))]
        kst = int(rng.integers(0, 60))
        a = osg.sg_chisq_one(D, IP, t, fpeak, vals, kst, 0.5, N)
        b = osg.sg_chisq_loop(D, IP, t, fpeak, vals, kst, 0.5, N)
        assert (a is None) == (b is None)
        if a is not None:
            assert abs(a - b) <= 1e-12 * b, (trial, a, b)
    # above 0.9 x Nyquist (921.6 Hz at 2048 Hz): not computed
    assert osg.sg_chisq_one(D, IP, 5, 900.0, PROD_TILES, 0, df, N) is None


def test_oracle_closed_form_single_tile():
    """Data equal to c IP g e^{-2 pi i k t / N} on the tile's bins gives |z|^2 = |c|^2 sigma^2, so the value (one
    tile: |z|^2 / 2) is |c|^2 4 df sum g^2 IP / 2."""
    rng = np.random.default_rng(6)
    N, sr = 1 << 14, 2048.0
    df = sr / N
    flen = N // 2 + 1
    IP = rng.uniform(0.5, 2.0, flen)
    q, off, fpeak, t, c = 12.0, 40.0, 180.0, 3333, 0.7 - 0.2j
    fcen = fpeak + off
    kk, k1, kg = int((fcen - 50) / df), int((fcen + 50) / df), int(round((fcen - 50) / df))
    k = np.arange(flen)
    f = k * df
    tau = q / (2 * math.pi * fcen)
    g = math.sqrt(math.pi) / 2 * tau * np.exp(-(math.pi * tau * (f - fcen)) ** 2) * (1 + np.exp(-q * q * f / fcen))
    g[(k < kg) | (k >= k1)] = 0
    D = c * IP * g * np.exp(-2j * np.pi * ((k * t) % N) / N)
    D[:kk] = 0
    v = osg.sg_chisq_one(D, IP, t, fpeak, [(q, off)], 0, df, N)
    exp = abs(c) ** 2 * 4 * df * (g * g * IP).sum() / 2
    assert abs(v - exp) <= 1e-10 * exp


def test_host_tile_ranges_match_oracle():
    import sgchisq
    rng = np.random.default_rng(8)
    for df in (1 / 512, 1 / 64, 0.5):
        fpeak = rng.uniform(20, 700, 200)
        kst = rng.integers(0, int(30 / df), 200)
        a = sgchisq.tile_ranges(fpeak, PROD_TILES, kst, df)
        b = osg.tile_ranges(fpeak, PROD_TILES, kst, df)
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
        # and the scalar Python arithmetic of the per-peak definition
        for i in range(0, 200, 37):
            kk = int(kst[i])
            for j, (q, off) in enumerate(PROD_TILES):
                fcen = fpeak[i] + off
                flo = max(kk * df, fcen - 50)
                kk = int(flo / df)
                assert (a[0][i, j], a[1][i, j], a[2][i, j]) == (kk, int((fcen + 50) / df), int(round(flo / df)))


def test_parse_locations():
    import sgchisq
    m1 = np.array([10.0, 25.0, 40.0, 3.0, 60.0])
    m2 = np.array([5.0, 10.0, 30.0, 2.0, 1.5])
    locs = sgchisq.parse_locations(['mtotal>30:20-15,20-30', 'mtotal>60&q<2:10-45'], m1, m2)
    assert locs[0] is None and locs[3] is None
    assert locs[1] == [(20.0, 15.0), (20.0, 30.0)] and locs[4] == [(20.0, 15.0), (20.0, 30.0)]
    assert locs[2] == [(10.0, 45.0)]                         # mtotal 70, q 1.33: the later entry wins
    assert sgchisq.parse_locations(['else:5-10'], m1, m2) == [[(5.0, 10.0)]] * 5
