"""sa_traces_l2_noalloc (the misfit with its partial sums in the library's static device array,
for callers that cannot allocate device memory) gives the same residual and misfit, bit for
bit, as sa_traces_l2 (which allocates its partial sums per call), called through the C ABI
twice in a row, the second time with fewer rows; it rejects more rows than its maximum."""

import ctypes

import numpy as np
import pytest

from stride_amd._native import Solver, lib

pytestmark = pytest.mark.gpu

MAX_ROWS = 1 << 16


def _l2(fn, solver, mod, obs, res, ntr, nt):
    fun = ctypes.c_double(-1.0)
    rc = fn(solver.handle, ctypes.c_void_p(mod.data_ptr()), ctypes.c_void_p(obs.data_ptr()),
            ctypes.c_void_p(res.data_ptr()), ctypes.c_int(ntr), ctypes.c_int(nt), ctypes.byref(fun))
    return rc, fun.value


def test_l2_noalloc_matches_l2(gpu):
    import torch
    L = lib()
    s = Solver(2, (32, 32), (1e-3, 1e-3), (4, 4), 1e-7, 10)
    dev = torch.device('cuda', 0)
    rng = np.random.default_rng(41)
    for ntr, nt in ((1024, 2500), (200, 999)):
        print('l2 noalloc: %d rows x %d samples' % (ntr, nt), flush=True)
        mod = torch.from_numpy(rng.standard_normal((ntr, nt)).astype(np.float32)).to(dev)
        obs = torch.from_numpy(rng.standard_normal((ntr, nt)).astype(np.float32)).to(dev)
        res_a = torch.zeros((ntr, nt), dtype=torch.float32, device=dev)
        res_b = torch.full((ntr, nt), 7.0, dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        rc_a, fun_a = _l2(L.sa_traces_l2, s, mod, obs, res_a, ntr, nt)
        rc_b, fun_b = _l2(L.sa_traces_l2_noalloc, s, mod, obs, res_b, ntr, nt)
        assert rc_a == 0 and rc_b == 0, L.sa_last_error().decode()
        assert fun_b == fun_a and fun_a > 0
        assert torch.equal(res_b, res_a)
        want = 0.5 * np.sum((mod.cpu().numpy().astype(np.float64) - obs.cpu().numpy()) ** 2)
        assert abs(fun_a - want) <= 1e-9 * want
    # more rows than the static array holds: rejected before any launch
    big = MAX_ROWS + 1
    tiny = torch.zeros(1, dtype=torch.float32, device=dev)
    rc, _ = _l2(L.sa_traces_l2_noalloc, s, tiny, tiny, tiny, big, 1)
    assert rc == -1 and 'exceed the maximum of 65536' in L.sa_last_error().decode()
    s.close()
