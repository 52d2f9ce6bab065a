"""Exact complex SNR numerator z and chi-squared bin power sum_j |Z_j|^2 at candidate peaks of one template tile
(the reference's exact_stage.py, dense form: the path fp_filter takes with --exact-stage dense and
--templates pycbc, exact_stage.py:196-203, 704-731).

Usage (fp_filter):
    stage = ExactStage(TemplateTile(H, nbins, kmin, kmax))                           # per tile
    pv = stage.evaluate(SegmentData(D, IP, df, tlen), PeakRequest(rows, samples))    # per (tile, segment)
    snr = |pv.z| norm;  chisq = (nb pv.pw - |pv.z|^2) norm^2;  sg reference frequency = pv.fpeak

Edges come from the float64 GPU edges kernel (chisq_edges), the sums from the fused chi-squared kernel
(chisq_kernel, conj(H) D formed in-kernel, phasors reseeded from the exact 64-bit phase), z and pw accumulated in
float64 over bins. The first evaluate checks up to CHECK_PEAKS peaks against a float64 direct sum of the same
bins and raises if CHECK_TOL fails (exact_stage.py:269-289). tlen need not be a power of two.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

import _dev
import _lib
import chisq_kernel
from chisq_edges import edges_blocked, edges_rows_dev, edges_rows_dev_multi

MAX_PEAKS = 8
CLIMB_MAX_RANGES = 1024     # bound on a climb's re-centred ranges (a cluster window of 2048 samples needs < 700)
CHECK_ROWS = 16
CHECK_PEAKS = 64
CHECK_TOL = dict(edge_bins=1, edge_frac=1e-2, snr_rel=1e-4, phase_rad=1e-4, chisq_rel=1e-3)


@dataclass(frozen=True, eq=False)
class TemplateTile:
    """One tile. H (n, flen) complex64 on the GPU holds the dense rows. nbins per row (0 = no chi-squared, one bin).
    Bins cover [kmin, kmax[row]). coarse (a Coarse): the rows are the interpolated templates of those nodes, and the
    exact stage regenerates them instead of reading H (exact_stage.py:53-65, 472-647); H then serves the climbs and
    the first call's self-check."""
    H: object
    nbins: np.ndarray
    kmin: int
    kmax: np.ndarray
    coarse: object = None

    def __post_init__(self):
        n = self.H.shape[0]
        object.__setattr__(self, 'nbins', np.asarray(self.nbins, np.int64))
        object.__setattr__(self, 'kmax', np.asarray(self.kmax, np.int64))
        assert self.H.ndim == 2 and self.H.dtype == torch.complex64 and self.H.is_contiguous()
        assert self.nbins.shape == (n,) and self.kmax.shape == (n,)
        if self.coarse is not None:
            g = self.coarse.grid
            assert tuple(self.coarse.amp.shape) == (n, g.nx) and tuple(self.coarse.coef.shape) == (n, g.J, 4)
            assert g.kmin == self.kmin and np.all(self.kmax <= g.kend), 'interpolated rows lie in [kmin, kend)'


@dataclass(frozen=True, eq=False)
class Coarse:
    """The nodes a tile's rows were interpolated from (templates.nodes): grid (templates.Grid, shared by the bank),
    amp (n, nx) and coef (n, J, 4) float64 on the GPU, the tile's rows in tile order; coef32 / amp2 the regenerated
    exact stage's float32 tables (templates.regen_tables)."""
    grid: object
    amp: object
    coef: object

    def __post_init__(self):
        import templates
        c32, a2 = templates.regen_tables(self.amp, self.coef)
        object.__setattr__(self, 'coef32', c32)
        object.__setattr__(self, 'amp2', a2)


@dataclass(frozen=True, eq=False)
class SegmentData:
    """One segment: D = stilde / psd (complex64, flen) and IP = 1 / psd (float32, flen), both zero below kmin."""
    D: object
    IP: object
    df: float
    tlen: int

    def __post_init__(self):
        flen = self.tlen // 2 + 1
        assert self.tlen > 0
        assert tuple(self.D.shape) == (flen,) and self.D.dtype == torch.complex64
        assert tuple(self.IP.shape) == (flen,) and self.IP.dtype == torch.float32


@dataclass(frozen=True, eq=False)
class PeakRequest:
    """Peaks of one (tile, segment): tile row and sample index. Duplicates are allowed; results keep this order."""
    rows: np.ndarray
    samples: np.ndarray

    def __post_init__(self):
        object.__setattr__(self, 'rows', np.asarray(self.rows, np.int64))
        object.__setattr__(self, 'samples', np.asarray(self.samples, np.int64))
        assert self.rows.ndim == 1 and self.rows.shape == self.samples.shape


@dataclass(frozen=True, eq=False)
class PeakValues:
    """Per requested peak, in request order: z = sum_j Z_j (complex64), pw = sum_j |Z_j|^2 (float32), fpeak =
    (2 e[nb-1] - e[nb-2]) df from this stage's edges (Hz); samples: where they were evaluated (the requested ones,
    or the refined ones with evaluate(refine > 0))."""
    z: np.ndarray
    pw: np.ndarray
    fpeak: np.ndarray
    samples: np.ndarray = None


def _fpeak(E, nb, df):
    r = np.arange(len(E))
    return (2 * E[r, np.maximum(nb - 1, 0)] - E[r, np.maximum(nb - 2, 0)]) * df


class ExactStage:
    """Exact z, pw and fpeak at requested peaks of one tile, segment after segment. Built once per tile."""

    def __init__(self, tile: TemplateTile, self_check=True):
        self.tile = tile
        self.self_check = self_check
        self.check: dict | None = None

    def _nbe(self, rows_u):
        return np.maximum(self.tile.nbins[rows_u], 1)

    def _edges_moments(self, segs, rows, rseg):
        """Edges of (segment rseg[i], row rows[i]) pairs from the segments' moments of 1/psd and the rows' node
        amplitudes (gwgs_edges_moments, exact_stage.py:230-250), one launch; (n, B + 1) int32."""
        c, g = self.tile.coarse, self.tile.coarse.grid
        nbe = self._nbe(rows)
        B = int(nbe.max())
        used = set(np.asarray(rseg).tolist())
        mom = [g.moments(segs[s].IP) if s in used else None for s in range(len(segs))]
        mi = np.array([m[1].data_ptr() if m is not None else 0 for m in mom], np.int64)
        pk = np.array([m[0].data_ptr() if m is not None else 0 for m in mom], np.int64)
        r_d, nb_d, k1_d, sg_d, mi_d, pk_d = _dev.to_dev_many(
            rows.astype(np.int32), nbe.astype(np.int32), self.tile.kmax[rows].astype(np.int32),
            np.asarray(rseg, np.int32), mi, pk)
        out = torch.empty((len(rows), B + 1), dtype=torch.int32, device='cuda')
        _lib.call('gwgs_edges_moments', c.amp.data_ptr(), g.nx, mi_d.data_ptr(), pk_d.data_ptr(), sg_d.data_ptr(),
                  g.kb_d.data_ptr(), r_d.data_ptr(), nb_d.data_ptr(), k1_d.data_ptr(), B + 1, g.kmin, len(rows),
                  out.data_ptr(), _dev.stream())
        return out

    def edges(self, seg: SegmentData, rows) -> np.ndarray:
        """(len(rows), B + 1) chi-squared bin edges for the given tile rows in this segment."""
        rows = np.asarray(rows, np.int64)
        E = edges_blocked(None, rows, self._nbe(rows), self.tile.kmin, self.tile.kmax[rows], self.tile.H, seg.IP)
        E[self.tile.nbins[rows] == 0, 0] = self.tile.kmin
        return E

    def evaluate(self, seg: SegmentData, req: PeakRequest, refine=0, lo=None, hi=None) -> PeakValues:
        """Exact values at the requested peaks. refine = w > 0: each peak is re-centred on the largest exact |z| among
        the samples t - w .. t + w that lie in [lo, hi) (per-peak bounds, e.g. its cluster window; first sample on
        ties), all candidates evaluated in the same pass; a peak whose best lies on the edge of that range (inside
        its bounds) is evaluated again around that sample, until every best is interior or on a bound, so each peak
        climbs to a local maximum of the exact |z| whichever sample of its slope it started from. PeakValues.samples
        holds the chosen samples. refine = 0 is the reference's rule: values at the requested sample."""
        return self.evaluate_async(seg, req, refine, lo, hi).result()

    def evaluate_async(self, seg: SegmentData, req: PeakRequest, refine=0, lo=None, hi=None) -> 'PendingValues':
        """evaluate() without waiting: edges and sums are enqueued on the current stream and their copies to the
        host on the side stream; PendingValues.result() waits for them and returns the PeakValues."""
        rows, samples = req.rows, req.samples
        npk = len(rows)
        if npk == 0:
            return PendingValues(self, seg, rows, samples[:, None], None, None, None)
        if self.tile.coarse is not None:
            return self.evaluate_async_multi([seg], np.zeros(npk, np.int64), req, refine, lo, hi)
        n = self.tile.H.shape[0]
        assert 0 <= rows.min() and rows.max() < n and 0 <= samples.min() and samples.max() < seg.tlen
        t = self.tile
        # peaks sorted by row (stable): distinct rows, each peak's edge row, and row-sorted groups for the sums
        ob = np.argsort(rows, kind='stable')
        rs = rows[ob]
        first = np.r_[True, rs[1:] != rs[:-1]]
        ur = rs[first]
        inv = np.empty(npk, np.int64)
        inv[ob] = np.cumsum(first) - 1
        E_dev = edges_rows_dev(ur, self._nbe(ur), t.kmin, t.kmax[ur], t.H, seg.IP)
        zero = np.flatnonzero(t.nbins[ur] == 0)
        if len(zero):
            E_dev[_dev.to_dev(zero), 0] = t.kmin
        if refine > 0:
            lo = np.zeros(npk, np.int64) if lo is None else np.asarray(lo, np.int64)
            hi = np.full(npk, seg.tlen, np.int64) if hi is None else np.asarray(hi, np.int64)
            off = np.arange(-refine, refine + 1)
            cand = np.clip(samples[:, None] + off[None, :], lo[:, None], hi[:, None] - 1)   # repeats at the bounds,
        else:                                                                              # removed by argmax order
            cand = samples[:, None]
        noff = cand.shape[1]
        out, zo = chisq_kernel.launch_sorted(t.H, np.repeat(rs, noff), cand[ob].ravel(), E_dev,
                                             np.repeat(inv[ob], noff), want_z=True, d=seg.D, n_time=seg.tlen,
                                             row_span=np.maximum(t.kmax - t.kmin, 0))
        pv = PendingValues(self, seg, rows, cand, ob, _dev.Fetch(out, zo, E_dev), inv)
        if refine > 0:
            pv.climb = (lo, hi, self._follow([seg], np.zeros(npk, np.int64), rows, lo, hi, refine, E_dev, ur, inv))
        return pv

    def _follow(self, segs, sidx, rows, lo, hi, refine, E_dev, uk, inv):
        """The climb of peaks e from samples c (gwgs_climb_points on the device, on the current stream), then their
        exact values re-centred once more within +-refine of the samples reached, so that the sample reported is
        chosen by the same sums as for peaks that did not climb; returns that PendingValues (copies on the current
        stream)."""
        def follow(e, c):
            n = len(e)
            rows_d, seg_d, ei_d, st_d, lo_d, hi_d, tab = _dev.to_dev_many(
                rows[e].astype(np.int32), sidx[e].astype(np.int32), inv[e].astype(np.int32),
                np.asarray(c, np.int64), lo[e].astype(np.int64), hi[e].astype(np.int64),
                np.array([sg.D.data_ptr() for sg in segs], np.int64))
            fin = torch.empty(n, dtype=torch.int64, device=rows_d.device)
            E = E_dev.contiguous()
            _lib.call('gwgs_climb_points', self.tile.H.data_ptr(), self.tile.H.shape[1], tab.data_ptr(),
                      int(segs[0].tlen), rows_d.data_ptr(), seg_d.data_ptr(), E.data_ptr(), E.shape[1],
                      ei_d.data_ptr(), st_d.data_ptr(), lo_d.data_ptr(), hi_d.data_ptr(), min(int(refine), 3),
                      CLIMB_MAX_RANGES, n, fin.data_ptr(), _dev.stream())
            fin_h = _dev.Fetch(fin, here=True).get()[0]
            return self.evaluate_async_multi(segs, sidx[e], PeakRequest(rows[e], fin_h), refine, lo[e], hi[e],
                                             edges=(E, uk), here=True, climb=False)
        return follow

    def evaluate_async_multi(self, segs, sidx, req: PeakRequest, refine=0, lo=None, hi=None, edges=None,
                             here=False, climb=True) -> 'PendingValues':
        """evaluate_async over the peaks of several segments of equal tlen and df in one edges launch and one sums
        launch: segs is a list of SegmentData, sidx the segment of each peak. Results in request order; the first-call
        self-check runs on the peaks of the first requested segment. edges: (device edges, their sorted keys
        segment n + row) of an earlier call on the same segments holding every (segment, row) requested here, used
        instead of recomputing them; here: the copies to the host run on the current stream; climb: peaks whose best
        lies on the edge of their range climb on (with refine > 0)."""
        rows, samples = req.rows, req.samples
        sidx = np.asarray(sidx, np.int64)
        npk = len(rows)
        seg0 = segs[int(sidx[0])] if npk else segs[0]
        if npk == 0:
            return PendingValues(self, seg0, rows, samples[:, None], None, None, None)
        assert all(sg.tlen == seg0.tlen and sg.df == seg0.df for sg in segs)
        n = self.tile.H.shape[0]
        assert 0 <= rows.min() and rows.max() < n and 0 <= samples.min() and samples.max() < seg0.tlen
        assert sidx.shape == rows.shape and 0 <= sidx.min() and sidx.max() < len(segs)
        t = self.tile
        key = sidx * n + rows
        ob = np.argsort(key, kind='stable')
        ks = key[ob]
        first = np.r_[True, ks[1:] != ks[:-1]]
        uk = ks[first]
        inv = np.empty(npk, np.int64)
        inv[ob] = np.cumsum(first) - 1
        ur, us = uk % n, uk // n
        if edges is not None:
            E_prev, k_prev = edges
            E_dev = E_prev[_dev.to_dev(np.searchsorted(k_prev, uk))]
        else:
            if t.coarse is not None:
                E_dev = self._edges_moments(segs, ur, us)
            else:
                E_dev = edges_rows_dev_multi(ur, us, self._nbe(ur), t.kmin, t.kmax[ur], t.H,
                                             np.array([sg.IP.data_ptr() for sg in segs], np.int64))
            zero = np.flatnonzero(t.nbins[ur] == 0)
            if len(zero):
                E_dev[_dev.to_dev(zero), 0] = t.kmin
        if refine > 0:
            lo = np.zeros(npk, np.int64) if lo is None else np.asarray(lo, np.int64)
            hi = np.full(npk, seg0.tlen, np.int64) if hi is None else np.asarray(hi, np.int64)
            off = np.arange(-refine, refine + 1)
            cand = np.clip(samples[:, None] + off[None, :], lo[:, None], hi[:, None] - 1)
        else:
            cand = samples[:, None]
        noff = cand.shape[1]
        out, zo = chisq_kernel.launch_sorted_multi(t.H, np.repeat(ks, noff), n, cand[ob].ravel(), E_dev,
                                                   np.repeat(inv[ob], noff),
                                                   np.array([sg.D.data_ptr() for sg in segs], np.int64), seg0.tlen,
                                                   row_span=np.maximum(t.kmax - t.kmin, 0), regen=t.coarse)
        pv = PendingValues(self, seg0, rows, cand, ob, _dev.Fetch(out, zo, E_dev, here=here), inv)
        pv.check_mask = sidx == sidx[0]
        pv.keep_alive = segs
        if refine > 0 and climb:
            pv.climb = (lo, hi, self._follow(segs, sidx, rows, lo, hi, refine, E_dev, uk, inv))
        return pv


class PendingValues:
    """Exact values of one evaluate_async call; result() waits for the device work and builds the PeakValues (the
    refine choice, fpeak from the edges, and the first-call self-check)."""

    climb = None   # (lo, hi, evaluate_async of (peak indices, new centres)) with refine > 0

    def __init__(self, stage, seg, rows, cand, order, fetch, inv):
        self.stage, self.seg, self.rows, self.cand = stage, seg, rows, cand
        self.order, self.fetch, self.inv = order, fetch, inv

    def result(self) -> PeakValues:
        rows = self.rows
        npk = len(rows)
        if npk == 0:
            return PeakValues(np.zeros(0, np.complex64), np.zeros(0, np.float32), np.zeros(0), np.zeros(0, np.int64))
        st = self.stage
        ho, hz, Eu = self.fetch.get()
        noff = self.cand.shape[1]
        pw_all = np.empty((npk, noff), np.float32); pw_all[self.order] = ho.reshape(npk, noff)
        z_all = np.empty((npk, noff), np.complex64); z_all[self.order] = hz.reshape(npk, noff)
        if noff > 1:
            j = np.argmax(np.abs(z_all.astype(np.complex128)), axis=1)  # first index: the earliest sample on ties
            r = np.arange(npk)
            samples, pw, z = self.cand[r, j], pw_all[r, j], z_all[r, j]
            if self.climb is not None:
                # a best on the edge of its range, inside its bounds, climbs on from there (the new range holds
                # it, so the result is at least as large)
                lo, hi, again = self.climb
                e = np.flatnonzero(((j == 0) & (self.cand[:, 0] > lo)) |
                                   ((j == noff - 1) & (self.cand[:, -1] < hi - 1)))
                if len(e):
                    # climbed on, on a stream of its own: the compute stream already holds later batches' work
                    with torch.cuda.stream(_dev.climb_stream()):
                        sub = again(e, samples[e]).result()
                    samples, pw, z = samples.copy(), pw.copy(), z.copy()
                    samples[e], pw[e], z[e] = sub.samples, sub.pw, sub.z
        else:
            samples, pw, z = self.cand[:, 0], pw_all[:, 0], z_all[:, 0]
        E = Eu.astype(np.int64)[self.inv]
        out = PeakValues(z, pw, _fpeak(E, st.tile.nbins[rows], self.seg.df), samples)
        if st.self_check and st.check is None:
            m = getattr(self, 'check_mask', None)
            if m is None:
                st.check = _self_check(st, self.seg, PeakRequest(rows, samples), out, E)
            else:
                st.check = _self_check(st, self.seg, PeakRequest(rows[m], samples[m]),
                                       PeakValues(z[m], pw[m], out.fpeak[m], samples[m]), E[m])
        return out


def _self_check(stage, seg, req, out, E):
    """Up to CHECK_PEAKS peaks of up to CHECK_ROWS rows against a float64 direct sum over the same bins, on the
    device in complex128 (exact_stage.py:269-289). Raises if a tolerance in CHECK_TOL fails."""
    rows_u = np.unique(req.rows)
    pick = rows_u[np.unique(np.linspace(0, len(rows_u) - 1, CHECK_ROWS).astype(int))]
    sel = np.flatnonzero(np.isin(req.rows, pick))[:CHECK_PEAKS]
    H, D, N = stage.tile.H, seg.D, seg.tlen
    z_ref = np.empty(len(sel), np.complex128); pw_ref = np.empty(len(sel))
    for o, i in enumerate(sel):
        e = E[i, :max(int(stage.tile.nbins[req.rows[i]]), 1) + 1]
        k = torch.arange(int(e[0]), int(e[-1]), device=H.device, dtype=torch.int64)
        q = torch.conj(H[int(req.rows[i]), e[0]:e[-1]].to(torch.complex128)) * D[e[0]:e[-1]].to(torch.complex128)
        ph = ((k * int(req.samples[i])) % N).to(torch.float64) * (2 * np.pi / N)
        x = q * torch.exp(1j * ph)
        b = torch.as_tensor(np.searchsorted(e, np.arange(e[0], e[-1]), side='right') - 1, device=H.device)
        Zj = torch.zeros(len(e) - 1, dtype=torch.complex128, device=H.device).index_add_(0, b, x)
        z_ref[o] = complex(Zj.sum().item()); pw_ref[o] = float((Zj.abs() ** 2).sum().item())
    rep = dict(n_peaks=int(len(sel)), n_rows=int(len(np.unique(req.rows[sel]))),
               **compare_values(out.z[sel], out.pw[sel], z_ref, pw_ref, stage.tile.nbins[req.rows[sel]]))
    t = CHECK_TOL
    ok = rep['snr_rel_max'] <= t['snr_rel'] and rep['phase_max'] <= t['phase_rad'] and rep['chisq_rel_max'] <= t['chisq_rel']
    if stage.tile.coarse is not None:
        # the edges from moments against the float64 cumsum of the dense rows' |h|^2 IP (exact_stage.py:283-287)
        rs = np.asarray(req.rows[sel], np.int64)
        nbe = stage._nbe(rs)
        E_ref = edges_rows_dev_multi(rs, np.zeros(len(rs), np.int64), nbe, stage.tile.kmin, stage.tile.kmax[rs],
                                     stage.tile.H, np.array([seg.IP.data_ptr()], np.int64)).cpu().numpy()
        E_ref[stage.tile.nbins[rs] == 0, 0] = stage.tile.kmin
        inside = np.arange(E_ref.shape[1])[None, :] <= nbe[:, None]
        dif = np.abs(E[sel][:, :E_ref.shape[1]].astype(np.int64) - E_ref)[inside]
        rep.update(edge_max_diff=int(dif.max()), edge_frac_diff=float((dif > 0).mean()))
        ok = ok and rep['edge_max_diff'] <= t['edge_bins'] and rep['edge_frac_diff'] <= t['edge_frac']
    assert ok, f'exact stage self-check failed: {rep}'
    return rep


def compare_values(z, pw, z_ref, pw_ref, nbins) -> dict:
    """Relative |z| error, phase error (rad) and relative chi-squared error, max and p99 (exact_stage.py:309-322)."""
    z = np.asarray(z, np.complex128); z_ref = np.asarray(z_ref, np.complex128)
    snr = np.abs(np.abs(z) - np.abs(z_ref)) / np.abs(z_ref)
    ph = np.abs(np.angle(z * np.conj(z_ref)))
    nb = np.asarray(nbins); m = nb > 0
    q = nb[m] * np.asarray(pw, np.float64)[m] - np.abs(z[m]) ** 2
    q_ref = nb[m] * np.asarray(pw_ref, np.float64)[m] - np.abs(z_ref[m]) ** 2
    chi = np.abs(q - q_ref) / np.abs(q_ref) if m.any() else np.zeros(1)
    return dict(snr_rel_max=float(snr.max()), snr_rel_p99=float(np.percentile(snr, 99)),
                phase_max=float(ph.max()), phase_p99=float(np.percentile(ph, 99)),
                chisq_rel_max=float(chi.max()), chisq_rel_p99=float(np.percentile(chi, 99)))
