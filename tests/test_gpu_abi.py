"""The C ABI's argument checks and reported outcomes on the GPU: batched
right-hand-side offsets, a tolerance that is not reached, the shift-state
getter behind the reference's property forwarding (src:182-191), and an FP64
system matrix handed to the refinement of an FP32 solution."""
import ctypes

import numpy as np
import pytest

from inputs import randn, stiff

pytestmark = pytest.mark.gpu


def test_batched_offset_checks(lhl_hip):
    n = 40
    ws = lhl_hip.lhl(randn(21, n))
    B = np.asfortranarray(np.random.default_rng(1).standard_normal((n, 6)))
    with pytest.raises(lhl_hip.DimensionMismatch):      # decreasing
        lhl_hip.shift_solve_batched(ws, [1, 1], [-0.1, -0.2], B, [0, 4, 2])
    with pytest.raises(lhl_hip.DimensionMismatch):      # negative
        lhl_hip.shift_solve_batched(ws, [1, 1], [-0.1, -0.2], B, [-1, 2, 4])
    with pytest.raises(lhl_hip.DimensionMismatch):      # past B's columns
        lhl_hip.shift_solve_batched(ws, [1, 1], [-0.1, -0.2], B, [0, 3, 7])
    import torch
    Xd = torch.zeros((n, 6), dtype=torch.float64, device="cuda").t().contiguous().t()
    with pytest.raises(lhl_hip.ArgumentError):          # B on the host, X on the GPU
        lhl_hip.shift_solve_batched(ws, [1, 1], [-0.1, -0.2], B, [0, 3, 6], X=Xd)
    # the C ABI itself rejects decreasing offsets (a caller that skips the wrapper)
    L = lhl_hip.load_library()
    X = np.zeros_like(B)
    s2 = np.array([1.0, 0, 1.0, 0])
    t2 = np.array([-0.1, 0, -0.2, 0])
    offs = np.array([0, 5, 2], dtype=np.int64)
    info = np.zeros(2, dtype=np.int64)
    code = L.lhl_shift_solve_batched(ws._h, 2, s2.ctypes.data, t2.ctypes.data, 0, offs.ctypes.data,
                                     ctypes.c_void_p(B.ctypes.data), n, ctypes.c_void_p(X.ctypes.data), n, 0, 0,
                                     0.0, 0, info.ctypes.data, None)
    assert code == lhl_hip.LHL_ERR_DIM
    # and the valid call still works after the rejected ones
    Xv, info, _ = lhl_hip.shift_solve_batched(ws, [1, 1], [-0.1, -0.2], B, [0, 3, 6])
    J = randn(21, n)
    assert np.allclose(Xv[:, :3], np.linalg.solve(np.eye(n) - 0.1 * J, B[:, :3]), rtol=1e-9)


def test_not_converged_is_reported(lhl_hip):
    """FP32 factors of a stiff J cannot reach 1e-12 without refinement rounds:
    max_iter = 0 leaves every column above tol, which the call reports
    (LHLNotConverged, outputs written) instead of returning as converged."""
    n, S, R = 300, 3, 2
    J = stiff(31, n)
    ws = lhl_hip.lhl(J, dtype=np.float32)
    gam = np.logspace(-4, -1, S)
    offs = np.arange(0, S * R + 1, R, dtype=np.int64)
    B = np.asfortranarray(np.random.default_rng(2).standard_normal((n, S * R)))
    with pytest.raises(lhl_hip.LHLNotConverged) as ei:
        lhl_hip.shift_solve_batched(ws, np.ones(S), -gam, B, offs, precision="refine64", tol=1e-12, max_iter=0)
    X0, info0, it0 = ei.value.result
    assert np.all(info0 == 0) and np.all(it0 == 0)
    W = np.eye(n) - gam[0] * J
    assert np.linalg.norm(X0[:, 0] - np.linalg.solve(W, B[:, 0])) <= 1e-3 * np.linalg.norm(X0[:, 0])
    X1, _, it1 = lhl_hip.shift_solve_batched(ws, np.ones(S), -gam, B, offs, precision="refine64", tol=1e-12,
                                            max_iter=0, strict=False)
    assert np.array_equal(X1, X0)
    # with rounds enough the same call converges and reports how many it took
    X2, _, it2 = lhl_hip.shift_solve_batched(ws, np.ones(S), -gam, B, offs, precision="refine64", tol=1e-12,
                                            max_iter=30)
    assert it2.max() >= 1
    for s in range(S):
        W = np.eye(n) - gam[s] * J
        c = offs[s]
        assert np.abs(B[:, c] - W @ X2[:, c]).max() <= 1e-12 * np.abs(B[:, c]).max()
    # lhl_refine! with a tolerance: one step is not enough for an FP32 solve at 1e-15
    lhl_hip.lhl_shift_(ws, 1, -gam[2])
    b = B[:, 4].astype(np.float32)
    x = lhl_hip.lhl_ldiv_(b.copy(), ws)
    with pytest.raises(lhl_hip.LHLNotConverged):
        lhl_hip.lhl_refine_(x, (np.eye(n) - gam[2] * J).astype(np.float32), b, ws, 1, tol=1e-15)


def test_shift_info_getter(lhl_hip):
    """ws.info / ws.sigma / ws.tau come from the handle (the fields the Julia
    property forwarding reads), for ws.shift and a held LHLShift."""
    L = lhl_hip.load_library()
    ws = lhl_hip.lhl(np.array([[1.0, 0.0], [0.0, 2.0]]))
    sh = lhl_hip.LHLShift(ws, np.complex128)

    def info_of(p):
        info = ctypes.c_int64(-7)
        sg = (ctypes.c_double * 2)()
        tu = (ctypes.c_double * 2)()
        loaded = ctypes.c_int(-1)
        assert L.lhl_get_shift_info(ws._h, p, ctypes.byref(info), sg, tu, ctypes.byref(loaded)) == 0
        return info.value, complex(sg[0], sg[1]), complex(tu[0], tu[1]), loaded.value

    assert info_of(None)[3] == 0
    lhl_hip.lhl_shift_(ws, 1.0, -1.0)            # pivot 1 is exactly zero (core_tests.jl:235-239)
    assert info_of(None) == (1, 1.0, -1.0, 1)
    assert ws.info == 1
    lhl_hip.lhl_shift_(ws, 1.0, -0.25)
    assert info_of(None) == (0, 1.0, -0.25, 1)
    lhl_hip.lhl_shift_(sh, ws, -1j, 1j)
    assert info_of(sh._ptr) == (1, -1j, 1j, 1)
    assert sh.info == 1
    lhl_hip.lhl_(ws, np.array([[3.0, 1.0], [0.5, 2.0]]))
    assert info_of(None)[3] == 0                 # a new reduction unloads the shift


def test_refine_torch_A_of_another_precision(lhl_hip):
    """An FP64 GPU A with an FP32 GPU x: A is cast to x's type, not
    reinterpreted; the result equals passing an FP32 A."""
    import torch
    n = 120
    J = randn(23, n)
    ws = lhl_hip.lhl(J, dtype=np.float32)
    g = 0.05
    W = np.eye(n) - g * J
    lhl_hip.lhl_shift_(ws, 1, -g)
    b = torch.from_numpy(np.random.default_rng(3).standard_normal(n).astype(np.float32)).cuda()
    x0 = lhl_hip.lhl_ldiv_(b.clone(), ws)
    A64 = torch.from_numpy(np.asfortranarray(W)).cuda()
    A32 = torch.from_numpy(np.asfortranarray(W.astype(np.float32))).cuda()
    xa = lhl_hip.lhl_refine_(x0.clone(), A64, b, ws, 2)
    xb = lhl_hip.lhl_refine_(x0.clone(), A32, b, ws, 2)
    assert torch.equal(xa, xb)
    r = np.abs(b.cpu().numpy() - W.astype(np.float32) @ xa.cpu().numpy()).max()
    assert r <= 1e-4 * np.abs(b.cpu().numpy()).max()
