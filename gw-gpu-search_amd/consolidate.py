"""Trigger consolidation: pycbc_inspiral's keep-loudest per interval applied to one dataset's peaks after every tile
(the reference's consolidate, fp_filter.py:629-663), on the host where the trigger lists live, and its
pre-consolidation of the GEMM candidates on the device (the reference's pre_keep, fp_filter.py:665-695), which
drops before the candidate list is compacted every candidate that consolidation would drop anyway
(csrc/consolidate.hip).

Why the pre-consolidation is exact: a peak that is not among the `num` loudest of its window in one tile and
segment's candidates is not among the loudest of the union either, and a peak at or below the num-th loudest
newsnr already kept in its window (the floor consolidate leaves) can never enter: its newsnr is its SNR and it comes
later in the list, so it loses ties. Applying consolidate after every tile gives the result of applying it once at
the end, for the same reason (fp_filter.py:630-632)."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

import _dev
import _lib


@dataclass
class KeepLoudest:
    """pycbc_inspiral's --newsnr-threshold, --keep-loudest-interval (s), --keep-loudest-num (fp_filter.py:267-268;
    the ranking statistic is newsnr) and the reference's --consolidate-exempt-snr (fp_filter.py:133): peaks with
    SNR at or above it are kept whatever their rank."""
    newsnr_threshold: float | None = None
    interval: float | None = None
    num: int | None = None
    exempt_snr: float | None = None

    def __post_init__(self):
        assert not self.interval or (self.num is not None and self.num >= 1), 'keep-loudest needs num >= 1'

    @classmethod
    def from_pycbc(cls, opt, exempt_snr=None):
        """From multi_filter.pycbc_parser()'s parsed options."""
        assert opt.keep_loudest_stat in (None, 'newsnr'), opt.keep_loudest_stat
        return cls(opt.newsnr_threshold, opt.keep_loudest_interval, opt.keep_loudest_num, exempt_snr)


def newsnr(snr, rchisq):
    """pycbc.events.ranking.newsnr (PyCBC 2.8.3, q = 6, n = 2), which the reference's consolidation and peak files
    use (fp_filter.py:250, 639, 930): float64 snr times (0.5 (1 + rchisq^3))^(-1/6) where rchisq > 1."""
    ns = np.array(snr, ndmin=1, dtype=np.float64)
    r = np.array(rchisq, ndmin=1, dtype=np.float64)
    i = np.nonzero(r > 1.0)[0]
    ns[i] *= (0.5 * (1.0 + r[i] ** 3.0)) ** (-1.0 / 6.0)
    return ns


def _window(t, gps_start, sr, interval):
    """Keep-loudest window of end times t (fp_filter.py:648-649): int64(round((t - gps_start) sr) / (interval sr))."""
    return (np.round((np.asarray(t, np.float64) - gps_start) * sr) / (interval * sr)).astype(np.int64)


def consolidate_parts(parts, gates, gps_start, sr, kl: KeepLoudest, want_floor=False):
    """fp_filter.py:629-663 on a list of parts (dicts of equal-length arrays: snr float32, chi float32, dof int32, t
    float64, and any other fields): the gate veto as at write, newsnr >= the threshold (or exempt), and with an interval
    the `num` loudest by newsnr per window among the non-exempt peaks, ties by position. Returns (the kept peaks as
    one dict, floor, number dropped); the floor (want_floor and an interval): per window up to the last ranked one,
    the num-th loudest newsnr of every window holding at least num ranked peaks, -inf elsewhere, else None."""
    c = {k: np.concatenate([p_[k] for p_ in parts]) for k in parts[0]}
    snr = c['snr'].astype(np.float32, copy=False)
    ns = newsnr(snr, c['chi'].astype(np.float32, copy=False) / np.maximum(c['dof'].astype(np.int32, copy=False), 1))
    exempt = snr >= np.float32(kl.exempt_snr) if kl.exempt_snr is not None else np.zeros(len(ns), bool)
    keep = (ns >= (kl.newsnr_threshold or 0)) | exempt
    t = c['t'].astype(np.float64, copy=False)
    for g_t, g_w, g_taper in gates:
        keep &= np.abs(t - g_t) > g_w + g_taper + 0.5
    floor = None
    if kl.interval:
        idx = np.nonzero(keep & ~exempt)[0]
        wt = _window(t[idx], gps_start, sr, kl.interval)
        od = np.lexsort((-ns[idx], wt))
        wt_s = wt[od]
        first = np.r_[0, np.nonzero(np.diff(wt_s))[0] + 1] if len(od) else np.zeros(0, np.int64)
        size = np.diff(np.r_[first, len(od)])
        rank = np.arange(len(od)) - np.repeat(first, size)
        ranked = np.zeros(len(keep), bool)
        ranked[idx[od[rank < kl.num]]] = True
        keep = ranked | (keep & exempt)
        if want_floor:
            full = size >= kl.num
            floor = np.full(int(wt_s[-1]) + 1 if len(wt_s) else 1, -np.inf)
            floor[wt_s[first[full]]] = ns[idx[od[first[full] + kl.num - 1]]]
    return {k: v[keep] for k, v in c.items()}, floor, int((~keep).sum())


def snapshot(tr):
    """The leading parts of a Triggers whose fields are final, as a list of dicts, and their count."""
    o = tr.parts
    n = tr.ready_parts() if hasattr(tr, 'ready_parts') else len(o['snr'])
    return [{k: o[k][i] for k in o} for i in range(n)], n


def replace_prefix(tr, n, kept):
    """Put the consolidated peaks of the first n parts in their place (later parts unchanged)."""
    for k in tr.parts:
        tr.parts[k] = [kept[k]] + tr.parts[k][n:]
    tr.settled = 1


def consolidate(tr, gates, gps_start, sr, kl: KeepLoudest, want_floor=False):
    """fp_filter.py:629-663 on one dataset's Triggers (in place; on its leading parts whose fields are final, the
    later parts kept after them unchanged). Returns (floor or None, number dropped); (None, 0) when there is no
    finished part (the caller keeps its previous floor)."""
    parts, n = snapshot(tr)
    if n == 0:
        return None, 0
    kept, floor, dropped = consolidate_parts(parts, gates, gps_start, sr, kl, want_floor)
    replace_prefix(tr, n, kept)
    return floor, dropped


class PreKeep:
    """The pre-consolidation of one batch's window maxima on the device (gwgs_pre_keep): candidates that
    consolidate() would drop are overwritten in vw before the candidate compaction. thr: the SNR threshold;
    exact_snr and margin: the exact band's lower edge exact_snr (1 - margin) (its candidates are all kept)."""

    def __init__(self, kl: KeepLoudest, thr, exact_snr, margin, sr, cw):
        self.kl, self.thr, self.sr, self.cw = kl, float(thr), float(sr), int(cw)
        self.exact_cut = np.float32(exact_snr * (1 - margin))
        self.wdiv = float(kl.interval) * self.sr if kl.interval else 1.0
        self.flags = (1 if kl.exempt_snr is not None else 0) | (2 if kl.interval else 0)
        self._work = None

    def _win(self, gps, samp):
        """The kernel's window of absolute sample samp (float64 as in the kernel; Python floats are IEEE doubles)."""
        t = gps + samp / self.sr
        return int(round((t - gps) * self.sr) / self.wdiv)

    def tables(self, segs):
        """segs: per batch segment (sbase, gps_start, col0, ncol, gates, floor device tensor or None). Returns the
        host tables (segi int64 (nseg, 10), segd, gates (ngate, 2), floor list) and the batch's window count."""
        segi = np.zeros((len(segs), 10), np.int64)
        segd = np.zeros(len(segs), np.float64)
        gt, fls, at = [], [], {}
        nwin = fl0 = 0
        for i, (sbase, gps, col0, ncol, gates, floor) in enumerate(segs):
            if self.kl.interval and ncol > 0:
                wt0 = self._win(gps, int(sbase))
                nw = self._win(gps, int(sbase) + ncol * self.cw - 1) - wt0 + 1
            else:
                wt0 = nw = 0
            nfl = 0 if floor is None else int(floor.numel())
            if nfl and id(floor) not in at:          # segments of one dataset share its floor
                at[id(floor)] = fl0
                fls.append(floor)
                fl0 += nfl
            segi[i] = (sbase, wt0, col0, ncol, nwin, nw, len(gt), len(gates), at[id(floor)] if nfl else 0, nfl)
            gt += [(float(g_t), g_w + g_taper + 0.5) for g_t, g_w, g_taper in gates]
            segd[i] = gps
            nwin += nw
        return segi, segd, np.asarray(gt, np.float64).reshape(-1, 2), fls, nwin

    def launch(self, vw, iw, cut, segs):
        """Enqueue gwgs_pre_keep on the batch's (n, nwsum) window maxima vw and samples iw (modified in place).
        cut: the candidate cut (float32) the compaction uses."""
        n, nwsum = vw.shape
        assert vw.dtype == torch.float32 and iw.dtype == torch.int32 and vw.is_contiguous() and iw.is_contiguous()
        assert tuple(iw.shape) == (n, nwsum)
        if n * nwsum == 0:
            return
        segi, segd, gates, fls, nwin = self.tables(segs)
        segi_d, segd_d, gates_d = _dev.to_dev_many(segi, segd, gates)
        floors_d = torch.cat(fls) if fls else None
        need = int(_lib.load().gwgs_pre_keep_workspace(n * nwsum, nwin)) if self.flags & 2 else 0
        if need and (self._work is None or self._work.numel() < need):
            self._work = None
            self._work = torch.empty(need, dtype=torch.uint8, device=vw.device)
        kl = self.kl
        _lib.call('gwgs_pre_keep', vw.data_ptr(), iw.data_ptr(), n, nwsum, len(segi), segi_d.data_ptr(),
                  segd_d.data_ptr(), gates_d.data_ptr(),
                  None if floors_d is None else floors_d.data_ptr(), self.sr, self.wdiv, int(kl.num or 0),
                  float(cut), float(self.exact_cut), float(np.float32(self.thr)),
                  float(np.float32(kl.newsnr_threshold or 0)),
                  float(np.float32(kl.exempt_snr)) if kl.exempt_snr is not None else 0.0, self.flags, nwin,
                  None if self._work is None else self._work.data_ptr(), need, _dev.stream())
