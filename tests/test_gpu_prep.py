"""GPU conditioning (gpu_prep) against the oracle's float64 restatement on the same seeded strain."""
import numpy as np
import pytest
import torch

from oracle import core

pytestmark = pytest.mark.gpu


def test_welch_median_and_mean(gpu):
    import gpu_prep
    rng = np.random.default_rng(7)
    x = (rng.standard_normal((2, 40000)) * 3).astype(np.float32)
    for avg in ('median', 'mean'):
        p, df = gpu_prep.welch(torch.from_numpy(x).cuda(), 4096, 2048, 1 / 2048., avg=avg)
        pr, dfr = core.welch(x, 4096, 2048, 1 / 2048., avg=avg)
        assert df == dfr
        np.testing.assert_allclose(p.cpu().numpy(), pr, rtol=1e-9)
    # an even segment count (numpy's median averages the two middle values)
    p, _ = gpu_prep.welch(torch.from_numpy(x[0]).cuda(), 4096, 4096, 1 / 2048., num_segments=8)
    pr, _ = core.welch(x[0], 4096, 4096, 1 / 2048., num_segments=8)
    np.testing.assert_allclose(p.cpu().numpy(), pr, rtol=1e-9)


@pytest.mark.parametrize('nseg', [63, 64, 65, 200, 301, 512, 513, 1500, 6000])
def test_welch_median_many_segments(gpu, nseg):
    """The median over up to 512 segments (odd and even counts, every staging size) and, through the one-bin-per-
    workgroup path, over hours of data (the autogating PSD of a 12,272 s dataset has about 6,000) against the
    oracle; the mean too for the largest."""
    import gpu_prep
    rng = np.random.default_rng(nseg)
    x = (rng.standard_normal((2, 128 * (nseg + 1))) * 3).astype(np.float32)
    p, _ = gpu_prep.welch(torch.from_numpy(x).cuda(), 256, 128, 1 / 2048., num_segments=nseg)
    pr, _ = core.welch(x, 256, 128, 1 / 2048., num_segments=nseg)
    np.testing.assert_allclose(p.cpu().numpy(), pr, rtol=1e-9)
    if nseg == 6000:
        p, _ = gpu_prep.welch(torch.from_numpy(x).cuda(), 256, 128, 1 / 2048., num_segments=nseg, avg='mean')
        pr, _ = core.welch(x, 256, 128, 1 / 2048., num_segments=nseg, avg='mean')
        np.testing.assert_allclose(p.cpu().numpy(), pr, rtol=1e-9)


def test_interpolate_and_truncation(gpu):
    import gpu_prep
    rng = np.random.default_rng(8)
    P = (1 + rng.random((2, 513))) * np.geomspace(1e3, 1, 513)[None]
    Pi = gpu_prep.interpolate(torch.from_numpy(P).cuda(), 0.25, 1 / 32.)
    Pr = core.interpolate(P, 0.25, 1 / 32.)
    np.testing.assert_allclose(Pi.cpu().numpy(), Pr, rtol=1e-12)
    kmin = int(15.0 * 32)
    for method in (None, 'hann'):
        T = gpu_prep.inverse_spectrum_truncation(Pi, 1 / 32., 1024, 15.0, method).cpu().numpy()
        Tr = core.inverse_spectrum_truncation(Pr, 1 / 32., 1024, 15.0, method)
        # float64 FFT rounding (rocFFT vs pocketfft); below kmin the filtered response is tiny and 1 / |.|^2 of it
        # amplifies that rounding, so the band used downstream is held tighter than the stop band
        np.testing.assert_allclose(T[:, kmin:], Tr[:, kmin:], rtol=1e-9)
        np.testing.assert_allclose(T, Tr, rtol=1e-5)


def test_condition_vs_oracle(gpu):
    """D and IP from the GPU conditioning against the float64 oracle on a 64 s stream with lines: IP within float32
    rounding (rel 1e-6 of the band max), D within rel-L2 1e-5 (float32 segment FFT)."""
    import gpu_prep, synth
    st = synth.make_stream('H1', 64, 2048.0, seed=1, lines=synth.default_lines())
    ds = gpu_prep.condition(torch.from_numpy(st.x).cuda(), 2048.0, st.gps_start, 64, 8, 8, 4, 2, 31, 4, 15.0)
    ref = core.condition(st.x, 2048.0, st.gps_start, 64, 8, 8, 4, 2, 31, 4, 15.0)
    assert ds.analyze == ref['analyze'] and ds.seg_start == ref['seg_start']
    IP = ds.IP.cpu().numpy().astype(np.float64)
    assert np.abs(IP - ref['IP']).max() <= 1e-6 * np.abs(ref['IP']).max()
    D = ds.D.cpu().numpy().astype(np.complex128)
    assert np.linalg.norm(D - ref['D']) / np.linalg.norm(ref['D']) < 1e-5


def test_autogating_finds_glitches(gpu):
    """Four Gaussian glitches (5 ms, optimal SNR 400) in 128 s of noise: the GPU whitened magnitude equals the
    oracle's, and the gate times equal the oracle's detections, each within 10 ms of a planted glitch."""
    import gpu_prep, synth
    st = synth.make_stream('H1', 128, 2048.0, seed=3, n_glitch=4, glitch_snr=400.0)
    x = torch.from_numpy(st.x.copy()).cuda()
    mag = gpu_prep.whitened_magnitude(x, 2048.0, 15.0, 16.0).cpu().numpy()
    mr = core.whiten_for_gating(st.x, 2048.0, 15.0, 16.0)
    assert np.abs(mag - mr).max() < 1e-4 * mr.max()
    gates = gpu_prep.autogate(x, 2048.0, st.gps_start, 50.0, cluster=0.5, pad=16.0, flow=15.0)
    ref = core.detect_loud_glitches(st.x, 2048.0, st.gps_start, 50.0, 0.5, 15.0, 16.0)
    assert [g[0] for g in gates] == ref
    truth = [st.gps_start + g[0] for g in st.glitches if 16.5 < g[0] < 128 - 16.5]
    assert len(ref) == len(truth) and all(min(abs(t - r) for r in ref) < 0.01 for t in truth)
    xg = st.x.copy()
    core.gate_taper(xg, 2048.0, st.gps_start, [(g, 0.25, 0.25) for g in ref])
    assert np.array_equal(x.cpu().numpy(), xg)
