"""Fused |SNR| peak finding (the reference's peak_kernel.py:1-47): for each (template row, cluster window) one
block reads the complex SNR series once and returns the max |z|^2 and its index in the window, first index on ties
(csrc/peaks.hip win_peak_k, wave64 reductions)."""
from __future__ import annotations

import torch

import _dev
import _lib

THREADS = 256


def window_peaks(z, a0, cw, nwin):
    """z: (rows, N) complex64 on the GPU, C-contiguous. Returns (max |z|^2, argmax within window), each
    (rows, nwin), float32 / int32."""
    assert z.dtype == torch.complex64 and z.is_contiguous()
    rows = z.shape[0]
    pmax = torch.empty((rows, nwin), dtype=torch.float32, device=z.device)
    pidx = torch.empty((rows, nwin), dtype=torch.int32, device=z.device)
    _lib.call('gwgs_window_peaks', z.data_ptr(), z.shape[1], rows, int(a0), int(cw), int(nwin), pmax.data_ptr(),
              pidx.data_ptr(), _dev.stream())
    return pmax, pidx


def block_window_max(V, I, nw, per):
    """fp_filter.py:745-746 on the device: per window the max of `per` consecutive block maxima and the sample of
    that block, first block on ties. V, I (n, >= nw per)."""
    n = V.shape[0]
    vw = torch.empty((n, nw), dtype=torch.float32, device=V.device)
    iw = torch.empty((n, nw), dtype=torch.int32, device=V.device)
    _lib.call('gwgs_block_window_max', V.data_ptr(), I.data_ptr(), V.stride(0), n, nw, per, vw.data_ptr(),
              iw.data_ptr(), _dev.stream())
    return vw, iw


def select_above(vw, cut):
    """Row-major flat indices of vw > cut (strict), as torch.nonzero orders them. Returns an int64 host array."""
    import numpy as np
    total = vw.numel()
    if total == 0:
        return np.zeros(0, np.int64)
    nwork = _lib.load().gwgs_select_workspace(total)
    work = torch.empty(int(nwork) + 1, dtype=torch.int32, device=vw.device)
    cnt = work[-1:]
    cap = max(1024, total // 64)
    while True:
        out = torch.empty(cap, dtype=torch.int32, device=vw.device)
        _lib.call('gwgs_select_above', vw.data_ptr(), total, float(cut), out.data_ptr(), cap, work.data_ptr(),
                  cnt.data_ptr(), _dev.stream())
        n = int(cnt.item())
        if n <= cap:
            return out[:n].cpu().numpy().astype(np.int64)
        cap = n


class Selection:
    """Candidate list being compacted on the device (select_peaks); result() reads it back. extra: a per-candidate
    device array written in the same pass (e.g. each candidate's alternate peaks), read back with the list."""

    def __init__(self, idx, samp, val, count, after, extra=None):
        self.idx, self.samp, self.val, self.after, self.extra = idx, samp, val, after, extra
        self.count = _dev.Fetch(count, after=after)

    def result(self):
        """(flat index g, window sample iw[g], vw[g]) of every vw > cut, in increasing g, as numpy arrays (and the
        extra array's entries, or None, when made with extra). Waits only for the selection's own work (and the
        copies), not for work enqueued after it."""
        import numpy as np
        n = int(self.count.get()[0][0])
        if n == 0:
            out = (np.zeros(0, np.int64), np.zeros(0, np.int32), np.zeros(0, np.float32))
            return out if self.extra is None else out + (np.zeros((0,) + tuple(self.extra.shape[1:]), np.int32),)
        g, s, v, e = _dev.Fetch(self.idx[:n], self.samp[:n], self.val[:n],
                                None if self.extra is None else self.extra[:n], after=self.after).get()
        out = (g.astype(np.int64), s, v)
        return out if self.extra is None else out + (e,)


def select_launch(vw, iw, cut, out=None):
    """Enqueue the one-pass compaction of vw > cut (fp_filter.py:747-757): returns device (idx, samp, val, count)
    with capacity vw.numel(); out: preallocated (idx, samp, val, work) to write into (graph capture)."""
    total = vw.numel()
    nwork = _lib.load().gwgs_select_workspace(total)
    cap = max(total, 1)
    if out is None:
        work = torch.empty(int(nwork) + 1, dtype=torch.int32, device=vw.device)
        idx = torch.empty(cap, dtype=torch.int32, device=vw.device)
        samp = torch.empty(cap, dtype=torch.int32, device=vw.device)
        val = torch.empty(cap, dtype=torch.float32, device=vw.device)
    else:
        idx, samp, val, work = (x[:n] for x, n in zip(out, (cap, cap, cap, int(nwork) + 1)))
    cnt = work[int(nwork):int(nwork) + 1]
    _lib.call('gwgs_select_peaks', vw.data_ptr(), iw.data_ptr(), total, float(cut), idx.data_ptr(), samp.data_ptr(),
              val.data_ptr(), cap, work.data_ptr(), cnt.data_ptr(), _dev.stream())
    return idx, samp, val, cnt


def select_peaks(vw, iw, cut):
    """fp_filter.py:747-757 on the device: entries of vw > cut (strict, row-major order) with their window samples
    iw and values, compacted in one pass (capacity: every entry). Returns a Selection without waiting."""
    idx, samp, val, cnt = select_launch(vw, iw, cut)
    return Selection(idx, samp, val, cnt, _dev.mark())
