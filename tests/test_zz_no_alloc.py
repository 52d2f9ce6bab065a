"""gwgs_select_peaks_noalloc (the candidate compaction with its block counts in the library's static device array,
for callers that cannot allocate) through the C ABI: on two calls in a row, the second with a smaller input, its
indices, window samples, values and count equal gwgs_select_peaks' on the same input; sizes past the array's limit
are refused before any launch."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

CH = 4096                       # entries per block count
MAX_TOTAL = (1 << 16) * CH      # the static array holds 65,536 block counts


def _inputs(rng, n, nw):
    vw = rng.uniform(0.0, 30.0, (n, nw)).astype(np.float32)
    vw[rng.random((n, nw)) < 0.02] = np.float32(24.0)               # ties at a value above the cut
    iw = rng.integers(0, 2048, (n, nw)).astype(np.int32)
    return torch.from_numpy(vw).cuda(), torch.from_numpy(iw).cuda()


def _select(lib, vw, iw, cut, noalloc):
    total = vw.numel()
    out = torch.full((total,), -7, dtype=torch.int32, device='cuda')
    out_i = torch.full((total,), -7, dtype=torch.int32, device='cuda')
    out_v = torch.full((total,), -7.0, dtype=torch.float32, device='cuda')
    cnt = torch.full((1,), -7, dtype=torch.int32, device='cuda')
    s = torch.cuda.current_stream().cuda_stream
    if noalloc:
        rc = lib.gwgs_select_peaks_noalloc(vw.data_ptr(), iw.data_ptr(), total, ctypes.c_float(cut), out.data_ptr(),
                                           out_i.data_ptr(), out_v.data_ptr(), total, cnt.data_ptr(), s)
    else:
        work = torch.empty(int(lib.gwgs_select_workspace(total)) + 1, dtype=torch.int32, device='cuda')
        rc = lib.gwgs_select_peaks(vw.data_ptr(), iw.data_ptr(), total, ctypes.c_float(cut), out.data_ptr(),
                                   out_i.data_ptr(), out_v.data_ptr(), total, work.data_ptr(), cnt.data_ptr(), s)
    assert rc == 0, rc
    torch.cuda.synchronize()
    n = int(cnt.item())
    return n, out[:n].cpu().numpy(), out_i[:n].cpu().numpy(), out_v[:n].cpu().numpy()


def test_noalloc_select_equals_workspace_form(lib):
    rng = np.random.default_rng(7)
    cut = 20.25
    for shape in ((512, 1800), (64, 37)):              # 921,600 then 2,368 entries: the second call is smaller
        print('noalloc select', shape, flush=True)
        vw, iw = _inputs(rng, *shape)
        exp = _select(lib, vw, iw, cut, noalloc=False)
        got = _select(lib, vw, iw, cut, noalloc=True)
        assert exp[0] > 0 and got[0] == exp[0]
        for a, b in zip(got[1:], exp[1:]):
            np.testing.assert_array_equal(a, b)
        flat = vw.cpu().numpy().ravel()
        np.testing.assert_array_equal(got[1], np.flatnonzero(flat > np.float32(cut)))


def test_noalloc_select_refuses_past_limit(lib):
    vw, iw = _inputs(np.random.default_rng(1), 4, 4)
    cnt = torch.zeros(1, dtype=torch.int32, device='cuda')
    s = torch.cuda.current_stream().cuda_stream
    rc = lib.gwgs_select_peaks_noalloc(vw.data_ptr(), iw.data_ptr(), MAX_TOTAL + 1, ctypes.c_float(1.0), None, None,
                                       None, 0, cnt.data_ptr(), s)
    assert rc == -2
    assert lib.gwgs_select_peaks_noalloc(vw.data_ptr(), iw.data_ptr(), -1, ctypes.c_float(1.0), None, None, None, 0,
                                         cnt.data_ptr(), s) == -1
