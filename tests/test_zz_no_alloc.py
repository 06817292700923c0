"""The allocation-free applyZ!/applyZinv! entry point against lhl_applyZ /
lhl_applyZinv on the same device columns: identical results (the same
kernels), on two calls in a row, the second with a smaller input; sizes
beyond the scratch array and host pointers are rejected."""
import ctypes

import numpy as np
import pytest

from inputs import randn

pytestmark = pytest.mark.gpu


def _noalloc(L, ws, inverse, x):
    import torch
    n, m = x.shape
    code = {torch.float64: (0, 0), torch.float32: (1, 0), torch.complex128: (0, 1), torch.complex64: (1, 1)}[x.dtype]
    return L.lhl_applyZ_noalloc(ws._h, int(inverse), ctypes.c_void_p(x.data_ptr()), n, m, x.stride(1), code[0],
                                code[1])


def test_noalloc_applyZ_matches_applyZ(lhl_hip):
    import torch
    L = lhl_hip.load_library()
    L.lhl_applyZ_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                     ctypes.c_int64, ctypes.c_int, ctypes.c_int]
    L.lhl_applyZ_noalloc.restype = ctypes.c_int
    rng = np.random.default_rng(71)
    for dtype, tdt in ((np.float64, torch.float64), (np.float32, torch.float32)):
        for n, m, cplx in ((700, 9, False), (300, 3, True), (700, 9, False), (130, 2, False)):
            ws = lhl_hip.lhl(randn(700 + n, n, dtype=dtype))
            for inverse in (0, 1):
                print(f"case: applyZ no-alloc {np.dtype(dtype).name} n={n} nrhs={m} complex={cplx} inverse={inverse}",
                      flush=True)
                X0 = rng.standard_normal((n, m)) + (1j * rng.standard_normal((n, m)) if cplx else 0)
                t = tdt if not cplx else (torch.complex128 if dtype == np.float64 else torch.complex64)
                xa = torch.from_numpy(np.asfortranarray(X0)).to(t).cuda().t().contiguous().t()
                xb = xa.clone().t().contiguous().t()
                (lhl_hip.applyZinv_ if inverse else lhl_hip.applyZ_)(xa, ws)
                assert _noalloc(L, ws, inverse, xb) == 0
                torch.cuda.synchronize()
                assert torch.equal(xa, xb), (n, m, inverse)
    # beyond the 16 MiB array, and a host pointer: rejected, nothing written
    ws = lhl_hip.lhl(randn(5, 512))
    big = torch.zeros((512, 4097), dtype=torch.float64, device="cuda").t().contiguous().t()
    assert _noalloc(L, ws, 0, big) == lhl_hip.LHL_ERR_ARG
    assert "scratch" in L.lhl_last_error(ws._h).decode()
    h = np.zeros((512, 2), order="F")
    code = L.lhl_applyZ_noalloc(ws._h, 0, ctypes.c_void_p(h.ctypes.data), 512, 2, 512, 0, 0)
    assert code == lhl_hip.LHL_ERR_ARG
    # and the entry point still works afterwards
    x = torch.from_numpy(np.asfortranarray(rng.standard_normal((512, 2)))).cuda().t().contiguous().t()
    y = x.clone().t().contiguous().t()
    assert _noalloc(L, ws, 1, x) == 0
    lhl_hip.applyZinv_(y, ws)
    assert torch.equal(x, y)
