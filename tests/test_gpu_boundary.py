"""The drop-in boundary on the GPU: the reference's conditioning stage API (gpu_stage2 against the reference's own
outputs, the 'ldas' resampling bit for bit), and fp_filter.main driven end to end with the reference's pycbc_inspiral
values (data/inspiral_argv_H1_000.txt, with a fixed chi-squared bin count) on a synthetic store, its peak file
against the oracle's plain method on the same conditioned data."""
import glob
import json
import os
import sys

import numpy as np
import pytest
import torch

from oracle import core, pipeline

pytestmark = pytest.mark.gpu

GOLD = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'reference_vectors.npz'))


def test_gpu_stage2_golden(gpu):
    """gpu_prep.gpu_stage2(xp, h, argv) on the golden strains with their padding and argv: the reference's own
    gpu_stage2 outputs (D and IP on the stored bins, layout exactly)."""
    import gpu_prep
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), 'golden'))
    import make_golden as mg
    for (name, seed, sr, seconds, pad, opts), m in zip(mg.STAGE2, json.loads(str(GOLD['s2_json']))):
        h = mg.stage2_strain(seed, sr, seconds, pad)
        r = gpu_prep.gpu_stage2(None, h, mg.stage2_argv(sr, seconds, pad, opts))
        D, IP = r['D'].cpu().numpy(), r['IP'].cpu().numpy()
        st = mg.STAGE2_STRIDE
        gD, gIP = GOLD[f's2_{name}_D'], GOLD[f's2_{name}_IP']
        nz = gIP > 0
        assert np.array_equal(IP[:, ::st] > 0, nz)
        assert np.max(np.abs(IP[:, ::st] - gIP)[nz] / gIP[nz]) < 2e-5
        assert np.max(np.abs(D[:, ::st] - gD)) < 1e-5 * np.abs(gD).max()
        assert [list(a) for a in r['analyze']] == m['analyze'] and list(r['seg_start']) == m['seg_start']
        assert (r['flen'], r['tlen'], r['df'], r['gps_start']) == (m['flen'], m['tlen'], m['df'], m['gps_start'])


def test_resample_ldas_golden(gpu):
    """gwgs_fir_decimate equals the reference's resample_ldas bit for bit (float64 taps and sums, no fused
    multiply-add)."""
    import gpu_prep
    x = torch.from_numpy(GOLD['resample_x']).cuda()
    for f in (2, 4):
        np.testing.assert_array_equal(gpu_prep.resample_ldas(x, f).cpu().numpy(), GOLD[f'resample_{f}'])


REF_ARGV = ("--bank-file bank.hdf --hdf-store store.hdf --channel-name H1:GWOSC_STRAIN --gps-start-time 1380077576 "
            "--gps-end-time 1380089848 --trig-start-time 1380077720 --trig-end-time 1380089832 --output trig.hdf "
            "--approximant SPAtmplt --order -1 --low-frequency-cutoff 20 --sample-rate 2048 --strain-high-pass 15 "
            "--pad-data 8 --segment-length 1024 --segment-start-pad 144 --segment-end-pad 16 --psd-estimation median "
            "--psd-segment-length 16 --psd-segment-stride 8 --psd-num-segments 63 --psd-inverse-length 16 "
            "--snr-threshold 5.0 --cluster-window 1 --cluster-function symmetric --chisq-bins 16 "
            "--newsnr-threshold 4.0 --keep-loudest-interval 1.072 --keep-loudest-num 100 --keep-loudest-stat newsnr "
            "--autogating-threshold 50 --autogating-cluster 0.5 --autogating-width 0.25 --autogating-taper 0.25 "
            "--autogating-pad 16 --processing-scheme cpu:1 --fft-backends fftw")


def test_main_end_to_end(gpu, tmp_path):
    """fp_filter.main on one H1 dataset of 1,216 s stored at 4096 Hz (resampled to 2048 Hz, high-passed, autogated,
    two 1024 s segments), 24 SPA templates (closed form on every bin, --templates pycbc), FP32 GEMM: the peak file's
    (template, window) list equals the oracle's
    plain method on the dataset's conditioned data except within 1e-3 of the threshold, exact SNR within 1e-4,
    chi-squared within 1e-3 and phase within 1e-3 rad of it at the same sample, end times from the segment layout,
    the 2-D sigmasq equal to the oracle's per first segment, and template hashes from the bank."""
    import fp_filter, gpu_prep, synth
    sr0, T, gps0 = 4096.0, 1216, 1000000000
    bank = synth.circular_bank(24, seed=5, mmin=8, mmax=60)
    inj = [synth.Injection(600.3, float(bank.m1[3]), float(bank.m2[3]), 14.0, 0.7)]
    st = synth.make_stream('H1', T, sr0, seed=9, injections=inj, gps_start=float(gps0))
    store = str(tmp_path / 'H1-store.npz')
    gpu_prep.write_store(store, 'H1:GWOSC_STRAIN', [(gps0, gps0 + T, st.x.astype(np.float64) / synth.DYN_RANGE_FAC)])
    bankf = str(tmp_path / 'bank.npz')
    np.savez(bankf, mass1=bank.m1, mass2=bank.m2)
    argvf = str(tmp_path / 'argv.txt')
    open(argvf, 'w').write(REF_ARGV)
    dsf = str(tmp_path / 'datasets.txt')
    open(dsf, 'w').write(f'H1 {gps0} {gps0 + T} {store}\n')
    out, cache = str(tmp_path / 'out'), str(tmp_path / 'cache')
    print('main', flush=True)
    fp_filter.main(['--datasets', dsf, '--bank', bankf, '--approximant', 'SPAtmplt', '--argv', argvf, '--out', out,
                    '--precision', 'fp32', '--prep-cache', cache, '--templates', 'pycbc'])
    pk = fp_filter.read_peaks(os.path.join(out, 'peaks', f'H1-{gps0}-{gps0 + T}.npz'))
    assert os.path.exists(os.path.join(out, 'timing.json'))
    meta = json.load(open(glob.glob(os.path.join(cache, '*.json'))[0]))
    with np.load(meta['file']) as z:
        D, IP = z['D'], z['IP']
    assert len(meta['analyze']) == 2 and meta['sr'] == 2048.0 and meta['tlen'] == 1 << 21
    data = [dict(D=D, IP=IP, analyze=[tuple(a) for a in meta['analyze']], seg_start=meta['seg_start'],
                 tlen=meta['tlen'], df=meta['df'], sr=meta['sr'], gps_start=meta['gps_start'])]
    print('oracle', flush=True)
    res, _ = pipeline.filter_plain(data, bank.m1, bank.m2, None, thr=5.0, eps=1e-3, tile=1024, flow=20.0)
    o = res[0]
    # the oracle's end times and the peak file's gate veto
    t_o = meta['gps_start'] + (np.asarray(meta['seg_start'])[o['seg'].astype(int)] + o['samp']) / meta['sr']
    keep = np.ones(len(t_o), bool)
    for g_t, g_w, g_tp in meta['gates']:
        keep &= np.abs(t_o - g_t) > g_w + g_tp + 0.5
    key = lambda t, e: (int(t), int(round((e - gps0) * 2048)) // 2048)
    ko = {key(t, e): i for i, (t, e) in enumerate(zip(o['tid'], t_o)) if keep[i]}
    kg = {key(t, e): i for i, (t, e) in enumerate(zip(pk['template_id'], pk['end_time']))}
    assert len(kg) > 20
    for k in set(ko) ^ set(kg):
        v = o['snr'][ko[k]] if k in ko else pk['snr'][kg[k]]
        assert abs(v - 5.0) < 5e-3, (k, v)
    both = [k for k in set(ko) & set(kg) if abs(t_o[ko[k]] - pk['end_time'][kg[k]]) < 1e-9]
    assert len(both) >= 0.98 * len(set(ko) & set(kg))
    i_o, i_g = np.array([ko[k] for k in both]), np.array([kg[k] for k in both])
    assert np.max(np.abs(pk['snr'][i_g] - o['snr'][i_o]) / o['snr'][i_o]) < 1e-4
    assert np.max(np.abs(pk['chisq'][i_g] - o['chi'][i_o]) / np.maximum(o['chi'][i_o], 1e-3)) < 1e-3
    assert np.max(np.abs(np.angle(np.exp(1j * (pk['coa_phase'][i_g] - o['ph'][i_o]))))) < 1e-3
    assert pk['snr'].max() > 12                                      # the injection
    H, _ = pipeline.spa_bank(bank.m1, bank.m2, None, meta['tlen'] // 2 + 1, meta['df'], 20.0)
    sig, _ = core.sigmasq(H.astype(np.complex64), IP, meta['df'])
    np.testing.assert_allclose(pk['sigmasq'], sig, rtol=1e-6)
    np.testing.assert_array_equal(pk['template_hash'], fp_filter.template_hashes(bank.m1, bank.m2)[pk['template_id']])
    assert int(pk['attrs/n_gates']) == len(meta['gates']) and len(str(pk['attrs/config_md5'])) == 32
