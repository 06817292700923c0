"""Pins for the CPU oracle (no GPU).  Each check is chosen so that a plausible
mistake in the restatement -- a dropped term, a wrong sign or index, a
transposed operand -- fails it; none retypes the oracle's formulas.

Reference tests restated: test/core_tests.jl:5-21 (reconstruction), 23-36
(shifted solves), 61-92 (blocked == unblocked), 115-212 (rows == fused shift
state, bit for bit), 235-239 (singular shift), 337-380 (pivoting families and
refinement)."""
import numpy as np
import pytest

from inputs import (PRECOMPILE_J, bwd, graded, hilbert_plus_I, near_nilpotent, randn, stiff, tie_matrix,
                    wilkinson)

SIZES = [1, 2, 3, 4, 17, 80, 128, 200, 301]


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("balance", [True, False])
def test_reconstruction(oracle, n, balance):
    J = randn(n, n)
    red = oracle.reduce(J, balance=balance)
    Z = red.Z()
    Zi = np.linalg.inv(Z) if n else Z
    assert np.allclose(Z @ Zi, np.eye(n), atol=1e-10)
    assert np.linalg.norm(Z @ red.H @ Zi - J) <= 1e-12 * max(1, n) * max(np.linalg.norm(J), 1)
    if n > 2:
        assert np.abs(np.tril(red.A, -2)).max() <= 1.0


@pytest.mark.parametrize("n", [40, 130])
def test_eigenvalues_preserved(oracle, n):
    J = randn(100 + n, n)
    red = oracle.reduce(J)
    ej = np.sort_complex(np.linalg.eigvals(J))
    eh = np.sort_complex(np.linalg.eigvals(red.H))
    assert np.allclose(ej, eh, atol=1e-9 * np.abs(ej).max())


def test_balance_exact_and_powers_of_two(oracle):
    J = graded(3, 60)
    B, d, isc = oracle.balance(J)
    assert np.all(np.frexp(d)[0] == 0.5)            # powers of two
    assert np.array_equal(d * isc, np.ones_like(d))
    # D (D^-1 J D) D^-1 == J bit for bit
    assert np.array_equal((B * d[:, None]) / d[None, :], J)
    assert np.array_equal(B, (J * isc[:, None]) * d[None, :])
    # the balanced matrix is a fixed point
    B2, d2, _ = oracle.balance(B)
    assert np.array_equal(d2, np.ones_like(d2))
    assert np.array_equal(B2, B)


def test_balance_equalizes_norms(oracle):
    J = graded(5, 40)
    B, d, _ = oracle.balance(J)
    A = np.abs(B)
    np.fill_diagonal(A, 0)
    c, r = A.sum(0), A.sum(1)
    ok = (c > 0) & (r > 0)
    # each converged pair is within a factor of 2 of balance (src:317-326) or was not worth moving
    assert np.all(np.maximum(c[ok] / r[ok], r[ok] / c[ok]) < 4.0)


def test_upper_hessenberg_closed_form(oracle):
    n = 30
    J = np.triu(randn(7, n), -1)
    red = oracle.reduce(J, balance=False)
    assert np.array_equal(red.A, J)
    assert np.array_equal(red.ipiv, np.arange(1, n - 1))
    assert np.array_equal(red.perm, np.arange(n))


def test_diagonal_closed_form(oracle):
    d = np.array([1.0, -3.0, 2.5, 7.0, -0.25, 4.0])
    J = np.diag(d)
    red = oracle.reduce(J)
    assert np.array_equal(red.H, J)
    b = np.arange(1.0, 7.0)
    sh = oracle.shift(red.Ht, 3.0, -0.5)
    x = oracle.ldiv(red, sh, b)
    assert np.allclose(x, b / (3.0 - 0.5 * d), rtol=1e-15)


@pytest.mark.parametrize("n", [1, 2])
def test_no_steps_for_tiny_n(oracle, n):
    J = randn(11, n)
    red = oracle.reduce(J, balance=False)
    assert np.array_equal(red.A, J)
    assert len(red.ipiv) == 0


@pytest.mark.parametrize("n", [160, 163, 200, 257])
@pytest.mark.parametrize("balance", [True, False])
def test_blocked_matches_unblocked(oracle, n, balance):
    J = randn(n, n)
    rb = oracle.reduce(J, balance=balance, blocked=1)
    ru = oracle.reduce(J, balance=balance, blocked=0)
    assert np.array_equal(rb.ipiv, ru.ipiv)
    assert np.linalg.norm(rb.A - ru.A) <= 1e-12 * n * np.linalg.norm(ru.A)


def test_blocked_matches_unblocked_ties(oracle):
    J = tie_matrix(9, 150, -1, 1)
    rb = oracle.reduce(J, blocked=1)
    ru = oracle.reduce(J, blocked=0)
    assert np.array_equal(rb.ipiv, ru.ipiv)
    assert np.linalg.norm(rb.A - ru.A) <= 1e-12 * 150 * np.linalg.norm(ru.A)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("tie", [False, True])
def test_shift_rows_equal_fused(oracle, dtype, tie):
    """test/core_tests.jl:192-211: every fused R gives the same state bit for bit."""
    for n in list(range(1, 41)) + [77]:
        J = tie_matrix(n, n, dtype=dtype) if tie else randn(n, n, dtype=dtype)
        red = oracle.reduce(J)
        for s, t in ((1, -0.1), (0, 1), (2, 0), (1, -3)):
            ref = oracle.shift(red.Ht, s, t)
            for R in (1, 2, 3, 4):
                f = oracle.shift(red.Ht, s, t, R=R)
                assert f.info == ref.info
                assert np.array_equal(f.swap, ref.swap)
                assert np.array_equal(f.U(), ref.U(), equal_nan=True)
                assert np.array_equal(f.mult(), ref.mult(), equal_nan=True)
                assert np.array_equal(f.rdiag, ref.rdiag, equal_nan=True)


def test_singular_shift_known_answers(oracle):
    red = oracle.reduce(np.array([[1.0, 0.0], [0.0, 2.0]]))
    assert oracle.shift(red.Ht, 1, -1.0).info == 1           # core_tests.jl:235-239
    assert oracle.shift_c(red.Ht, -1j, 1j).info != 0         # core_tests.jl:321-323
    assert oracle.shift(red.Ht, 0, 0).info == 1


@pytest.mark.parametrize("n", [1, 2, 3, 5, 40, 129, 260])
def test_shifted_solves_vs_lapack(oracle, n):
    """test/core_tests.jl:23-36."""
    J = randn(n, n)
    b = np.random.default_rng(n + 7).standard_normal(n)
    red = oracle.reduce(J)
    for s, t in ((1.0, 0.0), (1.0, -1e-8), (1.0, -0.05), (1.0, 1.7), (0.0, 1.0), (-3.0, 2.0)):
        W = s * np.eye(n) + t * J
        if np.linalg.matrix_rank(W) < n:
            continue
        sh = oracle.shift(red.Ht, s, t)
        x = oracle.ldiv(red, sh, b)
        xr = np.linalg.solve(W, b)
        assert np.linalg.norm(x - xr) <= 1e-9 * np.linalg.norm(xr)
        if n <= 260:
            assert bwd(W, x, b) < 50 * np.finfo(float).eps * max(1, n / 100)


@pytest.mark.parametrize("n", [3, 5, 40, 129])
def test_complex_shifts_vs_lapack(oracle, n):
    """test/core_tests.jl:251-303."""
    J = randn(n, n)
    rng = np.random.default_rng(n + 7)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    red = oracle.reduce(J)
    for s, t in ((0.4 + 0.7j, -1.0), (1.0, -0.05 + 0.3j), (0.0, 1j), (-3.0 + 2j, 2.0 - 1j)):
        W = s * np.eye(n) + t * J
        sh = oracle.shift_c(red.Ht, s, t)
        x = oracle.ldiv(red, sh, b)
        xr = np.linalg.solve(W, b)
        assert np.linalg.norm(x - xr) <= 1e-9 * np.linalg.norm(xr)
        y = oracle.refine(red, sh, W, b, x, 1)
        assert bwd(W, y, b) < 50 * np.finfo(float).eps


def test_mpmath_tiny(oracle):
    """Brute force in 50-digit arithmetic on tiny inputs (core_tests.jl:252, 270-273)."""
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 50
    for n in (3, 6, 11):
        J = randn(200 + n, n)
        b = np.random.default_rng(n).standard_normal(n)
        red = oracle.reduce(J)
        for s, t in ((1.0, -0.3), (0.5, 2.0)):
            sh = oracle.shift(red.Ht, s, t)
            x = oracle.ldiv(red, sh, b)
            W = mp.matrix([[mp.mpf(s) * (i == j) + mp.mpf(t) * mp.mpf(J[i, j]) for j in range(n)] for i in range(n)])
            xe = mp.lu_solve(W, mp.matrix([mp.mpf(v) for v in b]))
            xe = np.array([float(v) for v in xe])
            assert np.linalg.norm(x - xe) <= 1e-12 * np.linalg.norm(xe)


@pytest.mark.parametrize("name", ["randn", "wilkinson", "graded"])
def test_pivoting_families_refined(oracle, name):
    """test/core_tests.jl:337-363: one refinement step gives bwd < 1e-13."""
    n = 100
    J = {"randn": randn(1, n), "wilkinson": wilkinson(n), "graded": graded(3, n)}[name]
    b = np.random.default_rng(42).standard_normal(n)
    red = oracle.reduce(J)
    assert np.abs(np.tril(red.A, -2)).max() <= 1 + np.finfo(float).eps
    for g in (1e-3, 1.0, 1e3):
        sh = oracle.shift(red.Ht, 1.0, -g)
        if sh.info != 0:
            continue
        W = np.eye(n) - g * J
        x = oracle.ldiv(red, sh, b)
        x = oracle.refine(red, sh, W, b, x, 1)
        assert bwd(W, x, b) < 1e-13


def test_refinement_near_nilpotent(oracle):
    """test/core_tests.jl:365-380: kappa(Z) ~ 1e10, one step recovers bwd < 1e-13."""
    n = 60
    J = near_nilpotent(17, n)
    b = np.random.default_rng(18).standard_normal(n)
    W = np.eye(n) - 0.5 * J
    red = oracle.reduce(J)
    sh = oracle.shift(red.Ht, 1.0, -0.5)
    raw = oracle.ldiv(red, sh, b)
    ref = oracle.refine(red, sh, W, b, raw, 1)
    assert bwd(W, ref, b) <= bwd(W, raw, b)
    assert bwd(W, ref, b) < 1e-13


def test_precompile_workload_inputs(oracle):
    """The reference's hard-coded 4x4 precompile J (src:3225-3233)."""
    J = PRECOMPILE_J
    b = np.ones(4)
    red = oracle.reduce(J)
    sh = oracle.shift(red.Ht, 1.0, -0.1)
    x = oracle.ldiv(red, sh, b)
    W = np.eye(4) - 0.1 * J
    assert np.allclose(x, np.linalg.solve(W, b), rtol=1e-14)
    shc = oracle.shift_c(red.Ht, 1.0, 0.1j)
    xc = oracle.ldiv(red, shc, b.astype(complex))
    assert np.allclose(xc, np.linalg.solve(np.eye(4) + 0.1j * J, b), rtol=1e-14)


@pytest.mark.parametrize("dtype,n", [(np.float64, 48), (np.float64, 508), (np.float32, 48)])
def test_hilbert_plus_identity(oracle, dtype, n):
    """src:3204-3215: J = [1/(i+j) + (i==j)] at n = 48 and block_min + 8."""
    J = hilbert_plus_I(n, dtype)
    red = oracle.reduce(J)
    b = np.ones(n, dtype=dtype)
    sh = oracle.shift(red.Ht, 1.0, -1.0 / 8)
    x = oracle.ldiv(red, sh, b)
    W = np.eye(n) - J.astype(np.float64) / 8
    rtol = 1e-12 if dtype == np.float64 else 1e-4
    xr = np.linalg.solve(W, b.astype(np.float64))
    assert np.linalg.norm(x - xr) <= rtol * np.linalg.norm(xr)


def test_float32_reduction(oracle):
    J = randn(4, 150, dtype=np.float32)
    red = oracle.reduce(J)
    Z = red.Z().astype(np.float64)
    err = np.linalg.norm(Z @ red.H.astype(np.float64) @ np.linalg.inv(Z) - J) / np.linalg.norm(J)
    assert err < 1e-4
    assert np.array_equal(red.ipiv, oracle.reduce(J.astype(np.float64), blocked=0).ipiv) or err < 1e-4


def test_stiff_family_blocked(oracle):
    n = 600
    J = stiff(2000, n)
    rb = oracle.reduce(J)           # blocked (n >= 500)
    ru = oracle.reduce(J, blocked=0)
    assert np.array_equal(rb.ipiv, ru.ipiv)
    assert np.linalg.norm(rb.A - ru.A) <= 1e-12 * n * np.linalg.norm(ru.A)
