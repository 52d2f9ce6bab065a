"""The C ABI (CPU): the library loads, exports every symbol include/gwgs.h declares (and the ctypes table covers
them), and rejects bad arguments before touching a GPU."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, 'include', 'gwgs.h')).read()
    return sorted(set(re.findall(r'^(?:int|long)\s+(gwgs_\w+)\(', src, re.M)))


def test_header_symbols_exported(lib):
    import _lib
    names = _declared()
    assert len(names) >= 25
    out = subprocess.run(['nm', '-D', '--defined-only', _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r'\sT\s(gwgs_\w+)', out))
    assert set(names) == exported
    assert set(names) == set(_lib.SIGNATURES)
    for n in names:
        assert getattr(lib, n) is not None


def test_bad_arguments_rejected(lib):
    import _lib
    p = ctypes.c_void_p(256)
    # null pointers: rejected before any HIP call
    assert lib.gwgs_gemm_peaks(1, 0, None, 64, None, None, 64, 10, 64, 1024, 0, 1, 0, None, None, 8, None, None) == -1
    assert lib.gwgs_chisq_points(None, 10, None, 16, None, None, None, None, None, 2, 1, None, None, None, None) == -1
    assert lib.gwgs_chisq_points(p, 10, None, 16, p, p, p, p, p, 17, 3, p, None, None, None) == -1   # no workspace
    import chisq_kernel
    pg = lib.gwgs_chisq_peaks_per_group()
    assert pg == chisq_kernel.P                       # the host groups peaks as the kernel sizes its registers
    assert lib.gwgs_chisq_points_workspace(3, 17) == 3 * 16 * pg * 8
    assert lib.gwgs_chisq_points_workspace(0, 17) == 0 and lib.gwgs_chisq_points_workspace(3, 1) == 0
    assert lib.gwgs_edges_fused(None, 10, None, None, None, None, 2, 0, 1, None, None) == -1
    assert lib.gwgs_welch_reduce(ctypes.c_void_p(16), 1, 20000, 8, 1.0, 0, 1.0, 1.0, ctypes.c_void_p(16), None) == -2
    assert lib.gwgs_gemm_max_k(7) == 0
    assert lib.gwgs_select_workspace(4097) == 3
    # the fused basis transform: power-of-two lengths 2^13..2^22 only
    assert lib.gwgs_basis_series_workspace(10, 1 << 20) == 10 * (1 << 20) * 8
    assert lib.gwgs_basis_series_workspace(10, 1 << 13) == 10 * (1 << 13) * 8
    for n in (1 << 12, 3 << 16, 1 << 23, 0):
        assert lib.gwgs_basis_series_workspace(10, n) == 0
    p = ctypes.c_void_p(256)
    assert lib.gwgs_basis_series(1, p, 100, 10, 16, p, 0, 50, 3 << 16, 0, 10, None, 0, None, p, p, 32, None) == -3
    assert lib.gwgs_basis_series(1, p, 100, 10, 16, p, 0, 50, 1 << 12, 0, 10, None, 0, None, p, p, 32, None) == -3
    assert lib.gwgs_basis_series(1, p, 100, 10, 16, p, 0, 101, 1 << 16, 0, 10, None, 0, None, p, p, 32, None) == -1
    assert lib.gwgs_basis_series(1, p, 100, 10, 16, p, 0, 50, 1 << 16, 0, 10, None, 0, None, p, p, 30, None) == -1
    assert lib.gwgs_basis_series(9, p, 100, 10, 16, p, 0, 50, 1 << 16, 0, 10, None, 0, None, p, p, 32, None) == -3
    assert lib.gwgs_basis_series(1, p, 100, 10, 16, p, 0, 50, 1 << 16, 0, 10, None, 2, None, p, p, 32, None) == -3
    assert lib.gwgs_vd_power(None, 100, 10, p, 0, 50, p, None) == -1
    # the batched form: 1..16 segments, each segment's span inside the transform length
    arr = lambda ct, *v: (ct * len(v))(*v)
    bs = lambda nseg, a0, tn: lib.gwgs_basis_series_batch(
        1, p, 100, 10, 16, nseg, arr(ctypes.c_void_p, *([256] * max(nseg, 1))), 0, 50, 1 << 16,
        arr(ctypes.c_long, *([a0] * max(nseg, 1))), arr(ctypes.c_long, *([tn] * max(nseg, 1))),
        arr(ctypes.c_long, *([0] * max(nseg, 1))), None, 0, None, p, p, 32, None)
    assert bs(0, 0, 10) == -2 and bs(17, 0, 10) == -2
    assert bs(2, (1 << 16) - 5, 10) == -1 and bs(2, -1, 10) == -1
    # the second list pass: no buffers, a misaligned item table or segment table, and list-mode limits
    q = ctypes.c_void_p(260)                                                    # 4-byte aligned only
    assert lib.gwgs_hot_items(None, p, p, 10, 8, 1.0, p, 1, 16, p, p, p, None) == -1
    assert lib.gwgs_hot_items(p, p, p, 10, 0, 1.0, p, 1, 16, p, p, p, None) == -1
    assert lib.gwgs_hot_items(p, p, p, 10, 8, 1.0, q, 1, 16, p, p, p, None) == -1
    assert lib.gwgs_hot_items(p, p, p, 10, 8, 1.0, p, 1, 16, p, q, p, None) == -1
    dyn = lambda items, nit, sep, ldl: lib.gwgs_gemm_peaks_multi_lm_dyn(
        1, 0, p, 64, p, p, 64, 96, 64, items, nit, 8, 1, 16, sep, 0.5, 0.0, p, p, 8, p, p, ldl, None, None)
    assert dyn(None, p, 0, 32) == -1 and dyn(p, None, 0, 32) == -1 and dyn(q, p, 0, 32) == -1
    assert dyn(p, p, 16 * 128, 32) == -1 and dyn(p, p, 0, 31) == -1     # sep beyond the window, VL rows too short
    # pre-consolidation: empty batches return at once; bad flags, no segment table, or ranking without a workspace
    assert lib.gwgs_pre_keep_workspace(100, 5) == 8 * 100 + 4 * 16
    pk = lambda n, nseg, flags, work, wb: lib.gwgs_pre_keep(p, p, n, 10, nseg, p, p, p, p, 2048.0, 2195.456, 3, 1.0,
                                                            5.28, 4.0, 4.0, 5.0, flags, 5, work, wb, None)
    assert pk(0, 1, 2, None, 0) == 0
    assert pk(10, 1, 4, None, 0) == -1 and pk(10, 0, 0, None, 0) == -1 and pk(10, 1, 2, None, 0) == -1
    assert pk(10, 1, 2, p, 100) == -1                 # workspace smaller than gwgs_pre_keep_workspace
    # sine-Gaussian tiles: no peaks returns at once; no tile, a spectrum longer than N / 2 + 1 or null tables rejected
    sg = lambda npk, ntile, flen, ptr: lib.gwgs_sg_tiles(ptr, ptr, flen, 1 << 20, 1 / 512, ptr, ptr, ptr, ptr, ptr,
                                                         ptr, ptr, npk, ntile, ptr, ptr, None)
    assert sg(0, 8, 524289, p) == 0
    assert sg(4, 0, 524289, p) == -1 and sg(4, 8, 524290, p) == -1 and sg(4, 8, 524289, None) == -1
    with pytest.raises(AssertionError):
        _lib.check(-1, 'x')
    with pytest.raises(RuntimeError):
        _lib.check(700, 'x')


def test_no_cpu_fallback_when_library_missing(tmp_path, monkeypatch):
    """The product path fails loudly without the HIP library: load() raises, no other path exists."""
    import _lib
    monkeypatch.setattr(_lib, 'LIB_PATH', str(tmp_path / 'missing.so'))
    monkeypatch.setattr(_lib, '_lib', None)
    with pytest.raises(FileNotFoundError):
        _lib.load()


def test_product_never_imports_oracle():
    """Only tests/, smoke() and bench.py's CPU-baseline leg touch oracle/: no package module imports it."""
    pkg = os.path.join(ROOT, 'gw-gpu-search_amd')
    for f in os.listdir(pkg):
        if f.endswith('.py'):
            src = open(os.path.join(pkg, f)).read()
            assert not re.search(r'^\s*(from|import)\s+oracle\b', src, re.M), f
