"""The drop-in boundary on the CPU: the peak file (fields, dtypes, order, gates, attributes) and its atomic write,
the conditioning stage API's option handling, the template-hash and bank-file helpers, and the restated IIR
high-pass's defining property."""
import os

import numpy as np
import pytest


def _triggers():
    import fp_filter
    tr = fp_filter.Triggers(sg=True)
    rng = np.random.default_rng(3)
    for part in range(3):
        n = 7
        tr.add(tid=rng.integers(0, 5, n), t=1e9 + rng.uniform(0, 100, n), snr=rng.uniform(5, 9, n).astype(np.float32),
               ph=rng.uniform(-3, 3, n).astype(np.float32), chi=rng.uniform(0, 40, n).astype(np.float32),
               dof=np.full(n, 30, np.int32), sg=rng.uniform(1, 2, n), ex=np.ones(n, bool),
               samp=rng.integers(0, 1000, n), seg=np.zeros(n, np.int64))
    return tr


def test_peak_file_fields_and_order(tmp_path):
    """fp_filter.py:910-937: every dataset of the reference's peak file, its dtype, the (template_id, end_time) order,
    the gate veto |t - g_t| > width + taper + 0.5, template_hash by template_id, and the attrs / gates groups."""
    import fp_filter
    tr = _triggers()
    gates = [(float(tr.parts['t'][1][2]) + 0.3, 0.25, 0.25)]       # vetoes that trigger (and any within 1 s)
    f = tr.to_dict(gates=gates)
    th = np.arange(5, dtype=np.int64) * 1000 + 7
    sig2 = np.arange(10, dtype=np.float64).reshape(5, 2)
    arr = fp_filter.peak_file_arrays(f, th, sig2, gates, '{"a": 1}', 'abc')
    path = str(tmp_path / 'H1-1-2.npz')
    fp_filter.write_peaks(path, arr)
    got = fp_filter.read_peaks(path)
    assert set(fp_filter.PEAK_FIELDS) <= set(got)
    assert {'attrs/n_gates', 'attrs/provenance', 'attrs/config_md5', 'gates/time', 'gates/width',
            'gates/taper'} <= set(got)
    dt = dict(template_id=np.int64, template_hash=np.int64, snr=np.float32, end_time=np.float64, sigmasq=np.float64,
              coa_phase=np.float32, peak_sigmasq=np.float64, exact=np.bool_, chisq=np.float32, chisq_dof=np.int32,
              newsnr=np.float64, sg_chisq=np.float32, sg_chisq_computed=np.bool_)
    for k, t in dt.items():
        assert got[k].dtype == t, k
    n = len(got['snr'])
    assert all(len(got[k]) == n for k in dt if k != 'sigmasq') and got['sigmasq'].shape == (5, 2)
    t_all = np.concatenate(tr.parts['t'])
    assert n == int((np.abs(t_all - gates[0][0]) > 1.0).sum()) and n < len(t_all)
    assert np.all(np.abs(got['end_time'] - gates[0][0]) > 1.0)
    order = np.lexsort((got['end_time'], got['template_id']))
    np.testing.assert_array_equal(order, np.arange(n))
    np.testing.assert_array_equal(got['template_hash'], th[got['template_id']])
    assert int(got['attrs/n_gates']) == 1 and str(got['attrs/config_md5']) == 'abc'
    np.testing.assert_array_equal(got['gates/time'], [gates[0][0]])


def test_interrupted_write_leaves_no_file(tmp_path, monkeypatch):
    """A write that fails part way leaves neither the final file nor the temporary one; an earlier complete file at
    the same path is kept as it was (os.replace is the only step that makes a file visible)."""
    import fp_filter
    path = str(tmp_path / 'L1-1-2.npz')
    good = {'snr': np.ones(3, np.float32)}
    fp_filter.write_peaks(path, good)
    before = open(path, 'rb').read()

    def broken(f, **arrays):
        f.write(b'PK\x03\x04 partial')
        raise OSError('disk full')

    monkeypatch.setattr(np, 'savez', broken)
    with pytest.raises(OSError):
        fp_filter.write_peaks(path, {'snr': np.zeros(5, np.float32)})
    assert open(path, 'rb').read() == before and not os.path.exists(path + '.part')
    path2 = str(tmp_path / 'V1-1-2.npz')
    with pytest.raises(OSError):
        fp_filter.write_peaks(path2, good)
    assert not os.path.exists(path2) and not os.path.exists(path2 + '.part')


def test_stage_options():
    """gpu_prep.options (gpu_prep.py:306-311): the reference's argv values parse; options this path does not
    implement and missing stores are refused."""
    import gpu_prep
    av = ['--hdf-store', 's.npz', '--channel-name', 'H1:GWOSC_STRAIN', '--gps-start-time', '1380077576',
          '--gps-end-time', '1380089848', '--trig-start-time', '1380077720', '--trig-end-time', '1380089832',
          '--low-frequency-cutoff', '20', '--sample-rate', '2048', '--strain-high-pass', '15', '--pad-data', '8',
          '--segment-length', '1024', '--segment-start-pad', '144', '--segment-end-pad', '16', '--psd-estimation',
          'median', '--psd-segment-length', '16', '--psd-segment-stride', '8', '--psd-num-segments', '63',
          '--psd-inverse-length', '16', '--snr-threshold', '5.0', '--autogating-threshold', '50',
          '--autogating-cluster', '0.5', '--autogating-width', '0.25', '--autogating-taper', '0.25',
          '--autogating-pad', '16', '--processing-scheme', 'cpu:1']
    o = gpu_prep.options(av)
    assert (o.gps_start_time, o.segment_length, o.psd_num_segments, o.autogating_cluster) == (1380077576, 1024, 63, 0.5)
    assert gpu_prep.n_resampled(av) == (1380089848 - 1380077576 + 16) * 2048
    with pytest.raises(AssertionError):
        gpu_prep.options(av + ['--fake-strain', 'aLIGO'])
    with pytest.raises(AssertionError):
        gpu_prep.options([x for x in av if x not in ('--hdf-store', 's.npz')])
    with pytest.raises(NotImplementedError):
        gpu_prep.injection_runs(av + ['--injection-file', 'inj.xml'])


def test_highpass_property():
    """The restated high-pass: a sinusoid at the cutoff keeps 1 - attenuation of its amplitude through both passes,
    one at 4x the cutoff passes unchanged to 1e-3, one at a third of it is suppressed below 1e-3."""
    import gpu_prep
    sr, f0 = 2048.0, 15.0
    t = np.arange(int(64 * sr)) / sr
    mid = slice(len(t) // 4, 3 * len(t) // 4)
    for f, lo, hi in ((f0, 0.899, 0.901), (4 * f0, 0.999, 1.001), (f0 / 3, 0.0, 1e-3)):
        y = gpu_prep.highpass_cpu(np.sin(2 * np.pi * f * t), sr, 0.0, f0)
        a = np.sqrt(2 * np.mean(y[mid] ** 2))
        assert lo <= a <= hi, (f, a)


def test_bank_file_and_hashes(tmp_path):
    import fp_filter
    p = str(tmp_path / 'bank.npz')
    np.savez(p, mass1=np.array([10.0, 20.0, 30.0]), mass2=np.array([5.0, 15.0, 25.0]))
    m1, m2, e0, th = fp_filter.load_bank(p, 'SPAtmplt')
    assert e0 is None and len(set(th.tolist())) == 3
    np.testing.assert_array_equal(th, fp_filter.template_hashes(m1, m2))
    assert fp_filter.load_bank(p, 'SPAtmplt', max_templates=2)[0].tolist() == [10.0, 20.0]
    with pytest.raises(NotImplementedError):
        fp_filter.load_bank(p, 'IMRPhenomD')
    with pytest.raises(NotImplementedError):
        fp_filter._chisq_bins("0.72*get_freq('fSEOBNRv4Peak',params.mass1,params.mass2,params.spin1z,params.spin2z)**0.7")
    assert fp_filter._chisq_bins("'16'") == 16
