"""The LinearSolve adapter's operator algebra (no GPU): W's concrete form and
the (sigma, tau) it maps to, checked against W built directly with numpy.
The mapping follows README.md:57-70 (parity unpinned: LinearSolve.jl is not in
the reference)."""
import numpy as np
import pytest


def test_shift_of_matches_dense(lhl_hip):
    from lhl_hip import linsolve as LS
    rng = np.random.default_rng(3)
    n = 7
    J = rng.standard_normal((n, n))
    I = np.eye(n)
    for M, lam in ((None, 1.0), ("I", 1.0), (2.5, 2.5), (3.0 * I, 3.0)):
        for g in (0.01, 0.25, 0.02 + 0.01j):
            s, t = LS.shift_of(LS.WOperator(M, g, J))
            assert np.allclose(s * I + t * J, J - lam * I / g, rtol=1e-15, atol=0)
            s, t = LS.shift_of(LS.WOperator(M, g, J, transform=False))
            assert np.allclose(s * I + t * J, g * J - lam * I, rtol=1e-15, atol=0)
    # the README's own statement: (-1/(dt*gamma), 1) is J - I/(dt*gamma)
    assert LS.shift_of(LS.WOperator(None, 0.5, J)) == (-2.0, 1.0)


def test_mass_matrix_checks(lhl_hip):
    from lhl_hip import linsolve as LS
    n = 5
    M = np.eye(n)
    M[2, 2] = 2.0
    with pytest.raises(lhl_hip.ArgumentError):
        LS.mass_scalar(M)
    with pytest.raises(lhl_hip.ArgumentError):
        LS.mass_scalar(np.ones((2, 3)))
    with pytest.raises(lhl_hip.ArgumentError):
        LS.shift_of(LS.WOperator(None, 0.0, np.eye(n)))
    assert LS.mass_scalar(4.0 * np.eye(n)) == 4.0
    with pytest.raises(lhl_hip.ArgumentError):
        LS.init(LS.LinearProblem(np.eye(n), np.ones(n)), LS.LHLFactorization())
