"""Build and load libgwgs_hip.so, the C-ABI library of hand-written gfx950 kernels, through ctypes.

The library is built in-tree (hipcc --offload-arch=gfx950) by build(); load() only loads what is built and raises
if it is missing -- there is no fallback path. Return codes: 0 ok; negative = argument error (raised here as
AssertionError, like the reference's asserts); positive = hipError_t (raised as RuntimeError).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import threading
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB_PATH = os.path.join(HERE, 'libgwgs_hip.so')
BUILD_DIR = os.path.join(HERE, '_build')
ARCH = 'gfx950'
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
HIP_FLAGS = ['-O3', '-fPIC', '-std=c++17', f'--offload-arch={ARCH}', '-Xclang', '-target-feature', '-Xclang',
             '-packed-fp32-ops', '-fvisibility=hidden']

_C = ctypes
_p, _i, _l, _f, _d = _C.c_void_p, _C.c_int, _C.c_long, _C.c_float, _C.c_double
_ll = _C.c_longlong

# name -> argtypes (restype int unless listed in _RESTYPE)
SIGNATURES = {
    'gwgs_gemm_peaks': [_i, _i, _p, _l, _p, _p, _l, _i, _i, _i, _i, _i, _i, _p, _p, _l, _p, _p],
    'gwgs_gemm_max_k': [_i],
    'gwgs_gemm_peaks_multi': [_i, _i, _p, _l, _p, _p, _l, _i, _i, _p, _i, _i, _i, _p, _p, _l, _p, _p],
    'gwgs_gemm_peaks_multi_lm': [_i, _i, _p, _l, _p, _p, _l, _i, _i, _p, _i, _i, _i, _i, _f, _f, _p, _p, _l, _p,
                                 _p, _l, _p, _p],
    'gwgs_gemm_lm_count': [],
    'gwgs_gemm_item_bytes': [],
    'gwgs_window_peaks': [_p, _l, _i, _i, _i, _i, _p, _p, _p],
    'gwgs_block_window_max': [_p, _p, _l, _i, _i, _i, _p, _p, _p],
    'gwgs_select_workspace': [_l],
    'gwgs_select_above': [_p, _l, _f, _p, _i, _p, _p, _p],
    'gwgs_select_peaks': [_p, _p, _l, _f, _p, _p, _p, _i, _p, _p, _p],
    'gwgs_select_peaks_noalloc': [_p, _p, _l, _f, _p, _p, _p, _i, _p, _p],
    'gwgs_rebuilt_z': [_i, _p, _l, _p, _p, _l, _i, _p, _p, _p, _i, _p, _p, _p],
    'gwgs_gemm_peaks_multi_lm_dyn': [_i, _i, _p, _l, _p, _p, _l, _i, _i, _p, _p, _i, _i, _i, _i, _f, _f, _p, _p, _l,
                                     _p, _p, _l, _p, _p],
    'gwgs_hot_items': [_p, _p, _p, _i, _i, _f, _p, _i, _i, _p, _p, _p, _p],
    'gwgs_window_alternates': [_p, _p, _l, _i, _p, _p, _p, _p, _i, _i, _i, _f, _f, _i, _i, _p, _p],
    'gwgs_chisq_points': [_p, _l, _p, _ll, _p, _p, _p, _p, _p, _i, _i, _p, _p, _p, _p],
    'gwgs_chisq_points_workspace': [_i, _i],
    'gwgs_chisq_peaks_per_group': [],
    'gwgs_spa_nodes': [_p, _p, _p, _i, _p, _i, _d, _p, _p, _p],
    'gwgs_interp_templates': [_p, _p, _i, _i, _p, _p, _p, _l, _l, _l, _p, _l, _p],
    'gwgs_edges_moments': [_p, _i, _p, _p, _p, _p, _p, _p, _p, _i, _i, _i, _p, _p],
    'gwgs_regen_points_multi': [_p, _p, _i, _p, _p, _p, _i, _p, _p, _ll, _p, _p, _p, _p, _p, _i, _i, _p, _p, _p, _p],
    'gwgs_climb_points': [_p, _l, _p, _ll, _p, _p, _p, _i, _p, _p, _p, _p, _i, _i, _i, _p, _p],
    'gwgs_edges_fused': [_p, _l, _p, _p, _p, _p, _i, _i, _i, _p, _p],
    'gwgs_edges_fused_multi': [_p, _l, _p, _p, _p, _p, _p, _i, _i, _i, _p, _p],
    'gwgs_chisq_points_multi': [_p, _l, _p, _p, _ll, _p, _p, _p, _p, _p, _i, _i, _p, _p, _p, _p],
    'gwgs_form_q': [_p, _i, _l, _p, _l, _p, _p],
    'gwgs_pack_zt': [_i, _p, _l, _i, _l, _l, _l, _p, _p, _p],
    'gwgs_stack_a': [_i, _p, _i, _i, _p, _p, _i, _i, _p, _p, _p],
    'gwgs_stack_a_fp8': [_p, _i, _i, _p, _p, _i, _p, _p, _p, _p],
    'gwgs_stack_a_f16': [_p, _i, _i, _p, _p, _i, _p, _p, _p, _p],
    'gwgs_basis_power': [_p, _l, _i, _l, _l, _p, _p],
    'gwgs_basis_series_workspace': [_i, _l],
    'gwgs_basis_series': [_i, _p, _l, _i, _i, _p, _l, _l, _l, _l, _l, _p, _i, _p, _p, _p, _l, _p],
    'gwgs_basis_series_batch': [_i, _p, _l, _i, _i, _i, _p, _l, _l, _l, _p, _p, _p, _p, _i, _p, _p, _p, _l, _p],
    'gwgs_vd_power': [_p, _l, _i, _p, _l, _l, _p, _p],
    'gwgs_spa_templates': [_p, _p, _p, _i, _l, _d, _d, _d, _p, _l, _p, _p],
    'gwgs_sigmasq_workspace': [_i, _l, _l, _i],
    'gwgs_sigmasq': [_p, _l, _l, _l, _i, _p, _l, _i, _d, _p, _p, _l, _p],
    'gwgs_welch_segments': [_p, _l, _i, _l, _i, _i, _i, _p, _p],
    'gwgs_welch_reduce': [_p, _i, _i, _i, _d, _i, _d, _d, _p, _p],
    'gwgs_interp': [_p, _i, _i, _d, _d, _l, _p, _p],
    'gwgs_isqrt_band': [_p, _i, _l, _l, _l, _p, _p],
    'gwgs_trunc': [_p, _i, _l, _l, _i, _d, _p],
    'gwgs_inv_abs2': [_p, _i, _l, _d, _p, _p],
    'gwgs_whiten': [_p, _f, _p, _p, _i, _l, _l, _p, _p, _p],
    'gwgs_gate_pad': [_p, _l, _l, _l, _l, _p, _p],
    'gwgs_gate_whiten': [_p, _p, _l, _l, _d, _d, _p],
    'gwgs_gate_mag': [_p, _l, _l, _l, _d, _p, _p],
    'gwgs_gate_apply': [_p, _l, _l, _l, _l, _l, _l, _p],
    'gwgs_fir_decimate': [_p, _l, _p, _i, _i, _p, _p],
    'gwgs_phasor_data': [_p, _l, _l, _p, _p, _ll, _i, _p, _p, _l, _p],
    'gwgs_snr_grad': [_p, _p, _l, _p, _l, _p, _l, _i, _i, _d, _p, _p, _p, _p, _p, _p, _l, _p],
    'gwgs_slide_count': [_p, _p, _i, _p, _p, _p, _l, _d, _d, _i, _d, _p, _p, _p],
    'gwgs_slide_write': [_p, _p, _i, _p, _p, _p, _l, _d, _d, _i, _d, _p, _p, _p, _p, _p, _p],
    'gwgs_pre_keep_workspace': [_l, _l],
    'gwgs_sg_tiles': [_p, _p, _l, _l, _d, _p, _p, _p, _p, _p, _p, _p, _i, _i, _p, _p, _p],
    'gwgs_pre_keep': [_p, _p, _i, _i, _i, _p, _p, _p, _p, _d, _d, _i, _f, _f, _f, _f, _f, _i, _i, _p, _l, _p],
}
_RESTYPE = {'gwgs_select_workspace': _l, 'gwgs_basis_series_workspace': _l, 'gwgs_sigmasq_workspace': _l,
            'gwgs_chisq_points_workspace': _l, 'gwgs_pre_keep_workspace': _l}

_lib = None
_lock = threading.Lock()


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.hip'))


# per-file flags: the peak kernels compute |z|^2 without fused multiply-add so that they are bit-exact against
# numpy's float32 arithmetic on the same input, and the pre-consolidation computes its end times and windows as the
# host does in float64
FILE_FLAGS = {'peaks.hip': ['-ffp-contract=off'], 'consolidate.hip': ['-ffp-contract=off']}


def _compile(src):
    obj = os.path.join(BUILD_DIR, os.path.basename(src) + '.o')
    cmd = [HIPCC] + HIP_FLAGS + FILE_FLAGS.get(os.path.basename(src), []) + ['-c', src, '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f'hipcc failed on {src}:\n{r.stderr[-4000:]}')
    return obj


def build(jobs=8):
    """Compile every csrc/*.hip for gfx950 and link libgwgs_hip.so in this directory (always recompiles)."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, sources()))
    tmp = LIB_PATH + '.tmp'
    cmd = [HIPCC, '-shared', f'--offload-arch={ARCH}', '-o', tmp] + objs + ['-lamdhip64']
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f'link failed:\n{r.stderr[-4000:]}')
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def load():
    """The loaded library (ctypes.CDLL). Raises FileNotFoundError if libgwgs_hip.so has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise FileNotFoundError(f'{LIB_PATH} is not built: run __graft_entry__.build() (hipcc gfx950)')
            lib = ctypes.CDLL(LIB_PATH)
            for name, args in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.argtypes = args
                fn.restype = _RESTYPE.get(name, _i)
            _lib = lib
    return _lib


def check(rc, what):
    if rc < 0:
        raise AssertionError(f'{what}: invalid argument (code {rc})')
    if rc > 0:
        raise RuntimeError(f'{what}: HIP error {rc}')


def call(name, *args):
    rc = getattr(load(), name)(*args)
    check(rc, name)
    return rc
