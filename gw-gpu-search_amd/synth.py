"""Seeded synthetic inputs for the filtering path: coloured Gaussian detector noise with an analytic PSD, narrow
lines, Gaussian glitches, injected chirps, and the template banks (component masses, eccentricity).

This module only makes *inputs* (host numpy arrays and parameter tables). Templates themselves are generated on the
GPU by templates.py; the frequency-domain chirp here is used only to inject signals into the synthetic strain.

Conventions follow the reference's test data (tests/test_exact_stage.py:56-83): a one-sided PSD S(f), frequency
domain Gaussian noise with sd = sqrt(S / (4 df)) on the real and imaginary parts, stilde = dt rfft(x), and strain
scaled by DYN_RANGE_FAC before conditioning (gpu_prep.py:21, 380).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

DYN_RANGE_FAC = 5.9029581035870565e20      # pycbc.DYN_RANGE_FAC (gpu_prep.py:21)
MSUN_S = 4.925490947641267e-06              # G Msun / c^3 in seconds
F_FLAT = 15.0                               # the analytic PSD is held flat below this frequency
LINE_FREQS = (60.0, 120.0, 180.0, 500.0, 1000.0)
LINE_HEIGHT = 1e3
LINE_WIDTH = 0.05                           # Hz, Lorentzian half width (tests/test_exact_stage.py:62-63 form)


def analytic_psd(f, lines=()):
    """S(f) = 1e-47 [(f/20)^-8 + 3 (f/100)^-2 + 1 + (f/300)^2] (strain^2/Hz), held at its 15 Hz value below 15 Hz,
    times 1 + h / (1 + ((f - f0)/w)^2) for each (f0, h, w) in lines."""
    f = np.asarray(f, np.float64)
    fc = np.maximum(f, F_FLAT)
    S = 1e-47 * ((fc / 20.0) ** -8 + 3.0 * (fc / 100.0) ** -2 + 1.0 + (fc / 300.0) ** 2)
    for f0, h, w in lines:
        S = S * (1.0 + h / (1.0 + ((f - f0) / w) ** 2))
    return S


def default_lines():
    return [(f0, LINE_HEIGHT, LINE_WIDTH) for f0 in LINE_FREQS]


def colored_noise(n, sr, seed, lines=()):
    """n samples of stationary Gaussian noise with one-sided PSD analytic_psd (unscaled strain, float64).
    E|stilde_k|^2 = S_k / (2 df) with stilde = dt rfft(x)."""
    rng = np.random.default_rng(seed)
    df = sr / n
    flen = n // 2 + 1
    S = analytic_psd(np.arange(flen) * df, lines)
    sd = np.sqrt(S / (4 * df))
    X = sd * (rng.standard_normal(flen) + 1j * rng.standard_normal(flen))
    X[0] = 0.0
    if n % 2 == 0:
        X[-1] = X[-1].real * np.sqrt(2.0)
    return np.fft.irfft(X * sr, n)


def colored_noise_device(n, sr, seed, lines=()):
    """colored_noise generated on the GPU (torch's seeded Philox generator instead of numpy's; the same spectrum and
    conventions), float64 tensor of n samples. For workloads too large to draw on the host in reasonable time."""
    import torch
    g = torch.Generator(device='cuda')
    g.manual_seed(int(seed))
    df = sr / n
    flen = n // 2 + 1
    S = torch.from_numpy(analytic_psd(np.arange(flen) * df, lines)).cuda()
    sd = torch.sqrt(S / (4 * df))
    re = torch.randn(flen, generator=g, device='cuda', dtype=torch.float64)
    im = torch.randn(flen, generator=g, device='cuda', dtype=torch.float64)
    X = torch.complex(sd * re, sd * im)
    X[0] = 0.0
    if n % 2 == 0:
        X[-1] = X[-1].real * np.sqrt(2.0)
    return torch.fft.irfft(X * sr, n)


def device_stream(ifo, duration, sr, seed, lines=(), gps_start=1.0e9):
    """A noise-only Stream whose strain lives on the GPU: (Stream with an empty host array, float32 tensor of
    strain x DYN_RANGE_FAC)."""
    n = int(round(duration * sr))
    x = (colored_noise_device(n, sr, seed, lines) * DYN_RANGE_FAC).to(dtype=__import__('torch').float32)
    return Stream(ifo, np.zeros(0, np.float32), float(sr), gps_start), x


def mass_params(m1, m2):
    M = (m1 + m2) * MSUN_S
    eta = m1 * m2 / (m1 + m2) ** 2
    return M, eta


def f_isco(m1, m2):
    """Schwarzschild ISCO gravitational-wave frequency 1 / (6^1.5 pi M)."""
    M, _ = mass_params(np.asarray(m1, float), np.asarray(m2, float))
    return 1.0 / (6.0 ** 1.5 * np.pi * M)


def tau0(m1, m2, flow):
    """Newtonian chirp time from flow, 5 / (256 pi flow eta) (pi M flow)^(-5/3) (s): the bank sort key
    (fp_filter.py:419-420)."""
    M, eta = mass_params(np.asarray(m1, float), np.asarray(m2, float))
    return 5.0 / (256.0 * np.pi * flow * eta) * (np.pi * M * flow) ** (-5.0 / 3.0)


def spa_phase(f, m1, m2, e0=0.0, f0=15.0):
    """TaylorF2 phase to 1PN with t_c = phi_c = 0, plus the leading eccentric term (Yunes, Arun, Berti & Will 2009,
    PRD 80 084001): 3 / (128 eta v^5) [1 + (3715/756 + 55 eta / 9) v^2 - (2355/1462) e0^2 (f/f0)^(-19/9)] - pi/4."""
    M, eta = mass_params(m1, m2)
    v = (np.pi * M * f) ** (1.0 / 3.0)
    pn = 1.0 + (3715.0 / 756.0 + 55.0 * eta / 9.0) * v * v
    if e0:
        pn = pn - (2355.0 / 1462.0) * e0 * e0 * (f / f0) ** (-19.0 / 9.0)
    return 3.0 / (128.0 * eta * v ** 5) * pn - np.pi / 4.0


def chirp_fd(flen, df, m1, m2, flow, e0=0.0):
    """h(f) = f^(-7/6) exp(-i psi(f)) on k df for k in [kmin, kmax), zero elsewhere; kmin = int(flow/df),
    kmax = min(int(f_isco/df), flen - 1). complex128."""
    kmin = int(flow / df)
    kmax = min(int(f_isco(m1, m2) / df), flen - 1)
    h = np.zeros(flen, np.complex128)
    if kmax > kmin:
        f = np.arange(kmin, kmax) * df
        h[kmin:kmax] = f ** (-7.0 / 6.0) * np.exp(-1j * spa_phase(f, m1, m2, e0))
    return h


@dataclass
class Bank:
    """Template parameters, in bank order. e0 = 0 for circular templates."""
    m1: np.ndarray
    m2: np.ndarray
    e0: np.ndarray
    name: str = 'bank'

    def __len__(self):
        return len(self.m1)

    def tau0(self, flow):
        return tau0(self.m1, self.m2, flow)


def _loguniform_pairs(rng, n, lo, hi, mtot_max):
    m1, m2 = [], []
    while len(m1) < n:
        a, b = np.exp(rng.uniform(np.log(lo), np.log(hi), (2, 4 * n)))
        ok = a + b <= mtot_max
        m1 += list(np.maximum(a, b)[ok]); m2 += list(np.minimum(a, b)[ok])
    return np.array(m1[:n]), np.array(m2[:n])


def circular_bank(n, seed=11, mmin=5.0, mmax=150.0, mtot_max=None):
    """Component masses log-uniform in [mmin, mmax]; pairs with m1 + m2 above mtot_max (default mmax) are redrawn,
    so every template keeps f_isco >= 29 Hz and a usable band above a 15 Hz cutoff."""
    rng = np.random.default_rng(seed)
    m1, m2 = _loguniform_pairs(rng, n, mmin, mmax, mtot_max or mmax)
    return Bank(m1, m2, np.zeros(n), 'circular')


def eccentric_bank(n, seed=12, mtot=(45.0, 100.0), q=(1.0, 3.0), emax=0.3):
    """Total mass log-uniform in mtot, mass ratio m1/m2 uniform in q, e0 uniform in [0, emax] at 15 Hz."""
    rng = np.random.default_rng(seed)
    M = np.exp(rng.uniform(np.log(mtot[0]), np.log(mtot[1]), n))
    qq = rng.uniform(q[0], q[1], n)
    m2 = M / (1 + qq)
    return Bank(M - m2, m2, rng.uniform(0.0, emax, n), 'eccentric')


def sigma_analytic(h, df, S):
    ok = S > 0
    return np.sqrt(4 * df * np.sum(np.abs(h[ok]) ** 2 / S[ok]))


@dataclass
class Injection:
    t0: float            # merger time, s from the stream start
    m1: float
    m2: float
    snr: float           # optimal SNR in this detector against the analytic PSD
    phase: float


@dataclass
class Stream:
    """One detector's synthetic strain: x (float32, already x DYN_RANGE_FAC), its sample rate and GPS start, and the
    truth tables."""
    ifo: str
    x: np.ndarray
    sr: float
    gps_start: float
    injections: list = field(default_factory=list)
    glitches: list = field(default_factory=list)   # (t, sigma_s, optimal snr)


def make_stream(ifo, duration, sr, seed, lines=(), n_glitch=0, glitch_snr=50.0, glitch_width=5e-3,
                injections=(), gps_start=1.0e9, flow=15.0):
    """Noise + glitches + injections. injections: list of Injection (SNR scaled against the analytic PSD with the
    same lines). Returns a Stream with x = float32(strain x DYN_RANGE_FAC)."""
    n = int(round(duration * sr))
    x = colored_noise(n, sr, seed, lines)
    df = sr / n
    flen = n // 2 + 1
    f = np.arange(flen) * df
    S = analytic_psd(f, lines)
    S[: int(flow / df)] = 0
    rng = np.random.default_rng(seed + 1000)
    glitches = []
    for _ in range(n_glitch):
        tg = rng.uniform(0.1, 0.9) * duration
        G = np.sqrt(2 * np.pi) * glitch_width * np.exp(-2 * (np.pi * glitch_width * f) ** 2) * np.exp(-2j * np.pi * f * tg)
        a = glitch_snr / sigma_analytic(G, df, S)
        x += np.fft.irfft(a * G * sr, n)
        glitches.append((tg, glitch_width, glitch_snr))
    for inj in injections:
        h = chirp_fd(flen, df, inj.m1, inj.m2, flow)
        amp = inj.snr / sigma_analytic(h, df, S)
        s = amp * h * np.exp(-2j * np.pi * f * inj.t0 + 1j * inj.phase)
        x += np.fft.irfft(s * sr, n)
    return Stream(ifo, (x * DYN_RANGE_FAC).astype(np.float32), float(sr), gps_start, list(injections), glitches)


def network_injections(n, duration, seed=21, snr=(8.0, 30.0), masses=(5.0, 150.0), mtot_max=150.0, start=60.0,
                       end_pad=20.0, ndet=2):
    """n chirps with network SNR uniform in snr, split equally over ndet detectors (rho / sqrt(ndet) each), merger
    times spread over [start, duration - end_pad] at least 8 s apart, per-detector delays up to 10 ms."""
    rng = np.random.default_rng(seed)
    m1, m2 = _loguniform_pairs(rng, n, masses[0], masses[1], mtot_max)
    ts = np.sort(rng.uniform(start, duration - end_pad, n))
    for i in range(1, n):
        ts[i] = max(ts[i], ts[i - 1] + 8.0)
    rho = rng.uniform(snr[0], snr[1], n)
    per_det = []
    for d in range(ndet):
        dt = rng.uniform(-0.01, 0.01, n) if d else np.zeros(n)
        ph = rng.uniform(0, 2 * np.pi, n)
        per_det.append([Injection(float(ts[i] + dt[i]), float(m1[i]), float(m2[i]), float(rho[i] / np.sqrt(ndet)),
                                  float(ph[i])) for i in range(n) if ts[i] + dt[i] < duration - end_pad])
    return per_det
