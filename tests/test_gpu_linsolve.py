"""The LinearSolve adapter on the GPU (README.md:57-70): W x = b for the
operator's concrete form, one reduction across update_gamma_ calls, results
bitwise equal to the drop-in calls for the same (sigma, tau), a complex Radau
stage beside the real one on one reduction, and refinement of FP32 factors.
Parity with LinearSolve.jl itself is unpinned (not in the reference)."""
import numpy as np
import pytest

from inputs import randn, stiff

pytestmark = pytest.mark.gpu


def rel_resid(W, u, b):
    return np.abs(b - W @ u).max() / np.abs(b).max()


def bwd(W, u, b):
    """Normwise backward error of an unrefined FP64 solve: |b - W u| / (|W| |u| + |b|)
    in the infinity norm.  The solves go through the explicit inverse of L',
    whose rounding a relative residual bound ties to W's condition; the
    backward error does not depend on it (the oracle's own solve of the n = 150
    case below: 1.8e-15, the GPU's about 2e-14)."""
    Wn = np.abs(W).sum(axis=1).max()
    return np.abs(b - W @ u).max() / (Wn * np.abs(u).max() + np.abs(b).max())


def test_update_gamma_reuses_reduction(lhl_hip):
    from lhl_hip import linsolve as LS
    n = 300
    J = randn(71, n)
    b = np.random.default_rng(1).standard_normal(n)
    W = LS.WOperator(None, 0.01, J)
    cache = LS.init(LS.LinearProblem(W, b), LS.LHLFactorization())
    ws = lhl_hip.lhl(J)
    for g in (0.01, 0.013, 0.021, 0.5):
        if g != 0.01:
            LS.update_gamma_(cache, g)
        sol = LS.solve_(cache)
        assert sol.retcode == "Success"
        assert bwd(J - np.eye(n) / g, sol.u, b) <= 1e-12
        lhl_hip.lhl_shift_(ws, -1.0 / g, 1.0)
        x = lhl_hip.lhl_ldiv_(b.copy(), ws)
        assert np.array_equal(sol.u, x)
    assert cache.reductions == 1


def test_radau_complex_stage_and_mass(lhl_hip):
    from lhl_hip import linsolve as LS
    n = 200
    J = stiff(72, n)
    b = np.random.default_rng(2).standard_normal(n)
    W = LS.WOperator(2.0 * np.eye(n), 0.05, J, transform=False)      # W = gamma J - 2 I
    cache = LS.init(LS.LinearProblem(W, b), LS.LHLFactorization())
    u = LS.solve_(cache).u.copy()
    assert bwd(0.05 * J - 2 * np.eye(n), u, b) <= 1e-12
    gc = 0.03 + 0.02j                                                  # the complex stage
    LS.update_gamma_(cache, gc)
    uc = LS.solve_(cache).u.copy()
    assert np.iscomplexobj(uc)
    assert bwd(gc * J - 2 * np.eye(n), uc, b) <= 1e-12
    LS.update_gamma_(cache, 0.05)                                      # back to the real stage
    ur = LS.solve_(cache).u
    assert np.array_equal(ur.real, u) and not np.any(ur.imag)
    assert cache.reductions == 1


def test_new_operator_and_rhs(lhl_hip):
    from lhl_hip import linsolve as LS
    n = 150
    J1, J2 = randn(73, n), randn(74, n)
    rng = np.random.default_rng(3)
    b1, b2 = rng.standard_normal(n), rng.standard_normal(n)
    cache = LS.init(LS.LinearProblem(LS.WOperator(None, 0.1, J1), b1), LS.LHLFactorization())
    assert bwd(J1 - 10 * np.eye(n), LS.solve_(cache).u, b1) <= 1e-12
    cache.b = b2
    assert bwd(J1 - 10 * np.eye(n), LS.solve_(cache).u, b2) <= 1e-12
    assert cache.reductions == 1
    cache.A = LS.WOperator(None, 0.2, J2)
    assert bwd(J2 - 5 * np.eye(n), LS.solve_(cache).u, b2) <= 1e-12
    assert cache.reductions == 2


def test_device_rhs_and_fp32_refinement(lhl_hip):
    import torch
    from lhl_hip import linsolve as LS
    n = 400
    J = stiff(75, n)
    b = np.random.default_rng(4).standard_normal(n)
    bd = torch.from_numpy(b).cuda()
    alg = LS.LHLFactorization(dtype=np.float32, refine=20, tol=1e-12)
    cache = LS.init(LS.LinearProblem(LS.WOperator(None, 1e-3, J), bd), alg)
    for g in (1e-3, 1e-2):
        LS.update_gamma_(cache, g)
        u = LS.solve_(cache).u
        assert u.is_cuda and u.dtype == torch.float64
        assert rel_resid(J - np.eye(n) / g, u.cpu().numpy(), b) <= 1e-12
    assert cache.reductions == 1
