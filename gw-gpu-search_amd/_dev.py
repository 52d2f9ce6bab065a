"""Device plumbing for the host modules: torch tensors carry device memory and the current HIP stream; every
computation goes through libgwgs_hip.so (or hipFFT through torch.fft)."""
from __future__ import annotations

import numpy as np
import torch

import _lib

DEV = 'cuda'


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError('gw-gpu-search_amd needs a ROCm GPU (torch.cuda.is_available() is False)')
    _lib.load()


def ptr(t):
    """Device address of a tensor (None -> NULL)."""
    return None if t is None else t.data_ptr()


def stream():
    return torch.cuda.current_stream().cuda_stream


def as_dev(x, dtype=None):
    """numpy / torch -> contiguous torch tensor on the GPU."""
    if isinstance(x, torch.Tensor):
        t = x if x.is_cuda else x.to(DEV)
        if dtype is not None and t.dtype != dtype:
            t = t.to(dtype)
        return t.contiguous()
    a = np.ascontiguousarray(x)
    t = torch.from_numpy(a).to(DEV)
    return t if dtype is None else t.to(dtype)


def host(t):
    return t.detach().cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t)


_SIDE = None
_READY = None


def side_stream():
    """The stream that device-to-host copies run on (default priority), so a copy waits only for the work it
    depends on, not for everything queued on the compute stream."""
    global _SIDE
    if _SIDE is None:
        _SIDE = torch.cuda.Stream()
    return _SIDE


def ready_stream():
    """The stream for copies whose inputs are already computed (default priority). The side stream is in order, so
    a copy queued there also waits for every earlier copy's dependency (work enqueued later on the compute stream);
    a copy that needs no wait goes here instead."""
    global _READY
    if _READY is None:
        _READY = torch.cuda.Stream()
    return _READY


_BASIS = None
_CLIMB = None


def climb_stream():
    """The stream of the exact stage's follow-up passes (default priority): they are launched after later batches'
    work is already queued on the compute stream, and need none of it."""
    global _CLIMB
    if _CLIMB is None:
        _CLIMB = torch.cuda.Stream()
    return _CLIMB


def basis_stream():
    """The stream the filter's basis transforms run on (default priority), so that a batch's transform overlaps the
    previous batch's GEMM (instruction- and memory-bound work beside matrix-core work)."""
    global _BASIS
    if _BASIS is None:
        _BASIS = torch.cuda.Stream()
    return _BASIS


def to_dev(a, dtype=None):
    """Host array -> device tensor without a host-side wait: staged in pinned memory and copied on the current
    stream (the pinned block is kept by torch's host allocator until the copy has run)."""
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.pin_memory().to(DEV, non_blocking=True)


def to_dev_many(*arrays):
    """Several host arrays -> device tensors with one pinned staging buffer and one copy on the current stream
    (each array 16-byte aligned in the buffer). Returns the device tensors in order, with their shapes."""
    arrs = [np.ascontiguousarray(a) for a in arrays]
    offs, total = [], 0
    for a in arrs:
        offs.append(total)
        total += (a.nbytes + 15) // 16 * 16
    host = torch.empty(max(total, 16), dtype=torch.uint8, pin_memory=True)
    hn = host.numpy()
    for a, o in zip(arrs, offs):
        hn[o:o + a.nbytes] = a.view(np.uint8).ravel()
    dev = host.to(DEV, non_blocking=True)
    out = []
    for a, o in zip(arrs, offs):
        t = dev[o:o + a.nbytes].view(_TORCH_DTYPE[a.dtype.str[1:]])
        out.append(t.view(a.shape))
    return out


_TORCH_DTYPE = {'i4': torch.int32, 'i8': torch.int64, 'f4': torch.float32, 'f8': torch.float64,
                'c8': torch.complex64, 'u1': torch.uint8}


def mark():
    """An event at the current point of the current stream."""
    e = torch.cuda.Event()
    e.record()
    return e


class Fetch:
    """Device -> host copies of tensors on the side stream, ordered after `after` (an event; default: everything
    enqueued so far on the current stream); when `after` has already completed, on the ready stream, so that the
    copies do not queue behind the side stream's waits. get() waits for these copies only and returns numpy
    arrays."""

    def __init__(self, *tensors, after=None, here=False):
        if here:
            s = torch.cuda.current_stream()      # after the current stream's own work, in order
        elif after is not None and after.query():
            s = ready_stream()
        else:
            ev = after if after is not None else mark()
            s = side_stream()
            s.wait_event(ev)
        self.host = []
        with torch.cuda.stream(s):
            for t in tensors:
                if t is None:
                    self.host.append(None)
                    continue
                h = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
                h.copy_(t, non_blocking=True)
                t.record_stream(s)
                self.host.append(h)
        self.done = torch.cuda.Event()
        self.done.record(s)

    def get(self):
        self.done.synchronize()
        return [None if h is None else h.numpy() for h in self.host]


def i32(x):
    return as_dev(np.asarray(x, np.int32))


def i64(x):
    return as_dev(np.asarray(x, np.int64))


class Meter:
    """HIP events on the current stream around launches of named kernels, each with its algorithmic work (bytes
    for HBM-bound kernels, flops for the GEMM). Off unless `on` is set; stats() synchronises."""

    def __init__(self):
        self.on = False
        self.rec = {}
        self.notes = {}

    def start(self):
        if not self.on or torch.cuda.is_current_stream_capturing():
            return None
        e = torch.cuda.Event(enable_timing=True)
        e.record()
        return e

    def stop(self, name, e0, work):
        if e0 is None:
            return
        e1 = torch.cuda.Event(enable_timing=True)
        e1.record()
        self.rec.setdefault(name, []).append((e0, e1, float(work)))

    def note(self, name, work):
        """A quantity beside the timed launches (no events), e.g. the bytes a GEMM launch must read."""
        if self.on and not torch.cuda.is_current_stream_capturing():
            n, w = self.notes.get(name, (0, 0.0))
            self.notes[name] = (n + 1, w + float(work))

    def reset(self):
        self.rec = {}
        self.notes = {}

    def stats(self):
        """name -> (launches, kernel seconds, work)."""
        torch.cuda.synchronize()
        return {k: (len(v), sum(a.elapsed_time(b) for a, b, _ in v) / 1e3, sum(w for _, _, w in v))
                for k, v in self.rec.items()}


METER = Meter()
