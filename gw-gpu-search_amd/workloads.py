"""The synthetic workloads the filtering path is measured and checked on (inputs only; seeds fixed).

  small      1 detector x 64 s, 64 circular templates (5-100 Msun), segments of 64 s with 8 s pads.
  production 2 detectors x 512 s with 5 lines, 4 glitches (whitened peak about 60), 20 injections at network SNR
             8-30; the 4,612-template circular bank (5-150 Msun) and the 1,268-template eccentric bank
             (45-100 Msun total, e0 in [0, 0.3]); one 512 s segment per detector with 48 s / 16 s pads; PSD by
             16 s median Welch (stride 8 s, 63 segments) in float64, 16 s inverse-spectrum truncation.
Seeds: H1 = 1, L1 = 2, circular bank = 11, eccentric bank = 12, injections = 21.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

import synth

SR = 2048.0
FLOW = 15.0
SEEDS = dict(H1=1, L1=2, bank=11, bank_ecc=12, injections=21)


@dataclass
class Conditioning:
    seg_len: int
    start_pad: int
    end_pad: int
    psd_seg_len: float
    psd_stride: float
    psd_nseg: int
    psd_inv_len: float
    autogating_threshold: float | None = None
    autogating_cluster: float = 0.5
    autogating_pad: float = 16.0


def small():
    st = synth.make_stream('H1', 64, SR, seed=SEEDS['H1'])
    bank = synth.circular_bank(64, seed=SEEDS['bank'], mmin=5.0, mmax=100.0)
    return [st], [bank], Conditioning(64, 8, 8, 4, 2, 31, 4)


def production(duration=512, n_circ=4612, n_ecc=1268, n_inj=20, glitch_snr=300.0):
    """glitch_snr is the optimal SNR of each 5 ms Gaussian glitch; 300 gives a whitened peak near 65, above the
    autogating threshold of 50 (the reference's production setting)."""
    inj = synth.network_injections(n_inj, duration, seed=SEEDS['injections'], ndet=2)
    lines = synth.default_lines()
    streams = [synth.make_stream(ifo, duration, SR, seed=SEEDS[ifo], lines=lines, n_glitch=4, glitch_snr=glitch_snr,
                                 injections=inj[i]) for i, ifo in enumerate(('H1', 'L1'))]
    banks = [synth.circular_bank(n_circ, seed=SEEDS['bank'])]
    if n_ecc:
        banks.append(synth.eccentric_bank(n_ecc, seed=SEEDS['bank_ecc']))
    return streams, banks, Conditioning(duration, 48, 16, 16, 8, 63, 16, autogating_threshold=50.0)


# segment lengths of the batched form's grids: a chunk takes the longest that leaves it a full segment beyond the pads
GRIDS = (512, 256, 128)


def chunk_conditioning(L, grids=GRIDS):
    """Segmenting of one uneven chunk of L seconds (the batched form), with the production pads (48 s / 16 s): the
    longest segment length g of grids with L >= g + 64 s, else the shortest from L >= that length up (512 s segments
    from 576 s, 256 s from 320 s, 128 s from 128 s with the default grids; a longer segment analyses a larger share
    of what it transforms: 448 / 512, 192 / 256, 64 / 128). The 16 s / 8 s median-Welch PSD uses as many segments as
    the chunk holds, up to 63; chunks shorter than the shortest segment leave nothing to analyse (None)."""
    gs = sorted(grids, reverse=True)
    if L < gs[-1]:
        return None
    seg = next((g for g in gs[:-1] if L >= g + 64), gs[-1])
    nseg = int(min(63, (L - 16) // 8 + 1))
    return Conditioning(seg, 48, 16, 16, 8, nseg, 16)


def batched(n_chunks=200, ndet=2, seed=31, lmin=64, lmax=1024, step=16, n_circ=4612, n_ecc=1268, grids=GRIDS):
    """The batched form: per detector n_chunks chunks with lengths lmin + step U{0 .. (lmax - lmin) / step} s (segment
    gaps), independent seeded noise with the production lines; both banks. Returns (chunks, banks): chunks is a list
    of (Stream, Conditioning or None)."""
    rng = np.random.default_rng(seed)
    lens = lmin + step * rng.integers(0, (lmax - lmin) // step + 1, (ndet, n_chunks))
    lines = synth.default_lines()
    chunks = []
    t0 = 1.0e9
    for c in range(n_chunks):
        for d, ifo in enumerate(('H1', 'L1')[:ndet]):
            L = int(lens[d, c])
            st = synth.make_stream(ifo, L, SR, seed=1000 * (d + 1) + c, lines=lines, gps_start=t0)
            chunks.append((st, chunk_conditioning(L, grids)))
        t0 += int(lens[:, c].max()) + 8
    banks = [synth.circular_bank(n_circ, seed=SEEDS['bank'])]
    if n_ecc:
        banks.append(synth.eccentric_bank(n_ecc, seed=SEEDS['bank_ecc']))
    return chunks, banks


def batched_device(n_chunks=200, ndet=2, seed=31, lmin=64, lmax=1024, step=16, n_circ=4612, n_ecc=1268,
                   noise_seed=0, grids=GRIDS):
    """batched() with the chunks' noise drawn on the GPU (synth.device_stream): the same chunk lengths, GPS layout and
    banks; per chunk the noise seed is noise_seed + 1000 (d + 1) + c. Returns (chunks, x_dev, banks)."""
    rng = np.random.default_rng(seed)
    lens = lmin + step * rng.integers(0, (lmax - lmin) // step + 1, (ndet, n_chunks))
    lines = synth.default_lines()
    chunks, x_dev = [], []
    t0 = 1.0e9
    for c in range(n_chunks):
        for d, ifo in enumerate(('H1', 'L1')[:ndet]):
            L = int(lens[d, c])
            st, x = synth.device_stream(ifo, L, SR, noise_seed + 1000 * (d + 1) + c, lines=lines, gps_start=t0)
            chunks.append((st, chunk_conditioning(L, grids)))
            x_dev.append(x)
        t0 += int(lens[:, c].max()) + 8
    banks = [synth.circular_bank(n_circ, seed=SEEDS['bank'])]
    if n_ecc:
        banks.append(synth.eccentric_bank(n_ecc, seed=SEEDS['bank_ecc']))
    return chunks, x_dev, banks


def condition_streams(streams, cond: Conditioning, x_dev=None):
    """GPU conditioning of each stream (autogating first when configured). x_dev: optional list of float32 strain
    tensors already on the GPU (they are gated in place)."""
    import torch
    import gpu_prep
    out = []
    for i, st in enumerate(streams):
        x = x_dev[i] if x_dev is not None else torch.from_numpy(st.x).cuda()
        gates = []
        if cond.autogating_threshold is not None:
            gates = gpu_prep.autogate(x, st.sr, st.gps_start, cond.autogating_threshold, cond.autogating_cluster,
                                      pad=cond.autogating_pad, flow=FLOW)
        out.append(gpu_prep.condition(x, st.sr, st.gps_start, cond.seg_len, cond.start_pad, cond.end_pad,
                                      cond.psd_seg_len, cond.psd_stride, cond.psd_nseg, cond.psd_inv_len, FLOW,
                                      tag=st.ifo, gates=gates))
    return out


def analysed_seconds(datasets, cw):
    return sum((a1 - a0) // cw * cw / d.sr for d in datasets for a0, a1 in d.analyze)
