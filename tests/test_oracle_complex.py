"""Pins for the CPU oracle's complex-J functions (no GPU).

Reference tests restated: test/core_tests.jl:241-249 (the complex J solve),
5-21 (reconstruction), 235-239 and 321-323 (a singular shift), 337-380
(refinement to the bwd bound).  Each check is one a plausible slip in the
restatement (a conjugated product, a swapped plane, a dropped term of the
column-pair update) fails; none retypes the oracle's formulas."""
import numpy as np
import pytest

from inputs import graded, randn


def crandn(seed, n, m=None, dtype=np.complex128):
    rng = np.random.default_rng(seed)
    shape = (n, n if m is None else m)
    return np.asfortranarray((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype))


SIZES = [1, 2, 3, 5, 17, 40, 129, 260]


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("balance", [True, False])
def test_zreconstruction(oracle, n, balance):
    J = crandn(n, n)
    red = oracle.zreduce(J, balance=balance)
    Z = red.Z()
    Zi = np.linalg.inv(Z) if n else Z
    assert np.allclose(Z @ Zi, np.eye(n), atol=1e-10)
    assert np.linalg.norm(Z @ red.H @ Zi - J) <= 1e-12 * max(1, n) * max(np.linalg.norm(J), 1)
    if n > 2:
        # the pivot has the largest modulus of its column
        assert np.abs(np.tril(red.A, -2)).max() <= 1.0 + 4e-16
    # Ht is H^T
    assert np.array_equal(red.Ht, np.triu(red.A, -1).T)


@pytest.mark.parametrize("n", [40, 130])
def test_zeigenvalues_preserved(oracle, n):
    J = crandn(300 + n, n)
    red = oracle.zreduce(J)
    ej = np.sort_complex(np.linalg.eigvals(J))
    eh = np.sort_complex(np.linalg.eigvals(red.H))
    assert np.allclose(ej, eh, atol=1e-9 * np.abs(ej).max())


def test_zabs_against_modulus(oracle):
    rng = np.random.default_rng(5)
    z = rng.standard_normal(2000) * 10.0 ** rng.integers(-200, 200, 2000) \
        + 1j * rng.standard_normal(2000) * 10.0 ** rng.integers(-200, 200, 2000)
    for v in z:
        a = oracle.zabs(v)
        assert abs(a - abs(v)) <= 2.5 * np.spacing(abs(v)), v
    assert oracle.zabs(0j) == 0.0
    assert oracle.zabs(complex(-3, 0)) == 3.0          # a real entry keeps its magnitude exactly
    assert oracle.zabs(complex(0, -2.5)) == 2.5
    assert oracle.zabs(complex(np.inf, np.nan)) == np.inf
    assert np.isnan(oracle.zabs(complex(np.nan, 1.0)))


def test_zbalance_exact_and_powers_of_two(oracle):
    G = graded(3, 50)
    J = np.asfortranarray(G + 1j * graded(4, 50))
    B, d, isc = oracle.zbalance(J)
    assert np.all(np.frexp(d)[0] == 0.5)
    assert not np.all(d == 1.0)
    assert np.array_equal(d * isc, np.ones_like(d))
    assert np.array_equal((B * d[:, None]) / d[None, :], J)
    B2, d2, _ = oracle.zbalance(B)
    assert np.array_equal(d2, np.ones_like(d2))


def test_zbalance_uses_the_modulus(oracle):
    # a purely imaginary copy balances with the same scales as the real matrix
    J = graded(7, 40)
    _, d_re, _ = oracle.balance(J)
    _, d_im, _ = oracle.zbalance(1j * J)
    assert np.array_equal(d_re, d_im)


@pytest.mark.parametrize("n", [3, 12, 61, 200])
def test_real_J_as_complex_matches_real_reduction(oracle, n):
    J = randn(40 + n, n)
    rr = oracle.reduce(J, blocked=0)
    zr = oracle.zreduce(J.astype(np.complex128))
    assert np.array_equal(zr.ipiv, rr.ipiv)
    assert np.array_equal(zr.perm, rr.perm)
    assert np.array_equal(zr.scale, rr.scale)
    assert np.abs(zr.A.imag).max(initial=0.0) == 0.0
    assert np.linalg.norm(zr.A.real - rr.A) <= 1e-13 * n * np.linalg.norm(rr.A)


@pytest.mark.parametrize("n", [1, 2, 6, 33])
def test_zupper_hessenberg_closed_form(oracle, n):
    J = np.asfortranarray(np.triu(crandn(9, n), -1))
    red = oracle.zreduce(J, balance=False)
    assert np.array_equal(red.A, J)
    assert np.array_equal(red.ipiv, np.arange(1, n - 1))
    assert np.array_equal(red.perm, np.arange(n))


def test_zreference_complex_testset(oracle):
    # core_tests.jl:241-249: n = 30, gamma = 0.2 + 0.3i, rtol 1e-9 against a dense solve
    n = 30
    J = crandn(13, n)
    b = crandn(14, n, 1)[:, 0]
    g = 0.2 + 0.3j
    red = oracle.zreduce(J)
    sh = oracle.zshift(red.Ht, 1, -g)
    x = oracle.zldiv(red, sh, b)
    xd = np.linalg.solve(np.eye(n) - g * J, b)
    assert np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd)


@pytest.mark.parametrize("n", [2, 7, 50, 200])
@pytest.mark.parametrize("st", [(0.4 + 0.7j, -1.0), (1.0, -0.05 + 0.3j), (0.0, 1j), (-3.0 + 2j, 2.0 - 1j)])
def test_zshift_is_an_lu_of_the_shifted_hessenberg(oracle, n, st):
    s, t = st
    J = crandn(60 + n, n)
    red = oracle.zreduce(J)
    sh = oracle.zshift(red.Ht, s, t)
    G = s * np.eye(n) + t * red.H
    # rebuild P L U from the stored pivot choices: step k takes row k (pending)
    # or row k+1 (fresh) as pivot row, l in Gt[k, k+1]
    U = sh.U()
    rows = [G[k].copy() for k in range(n)]
    r = rows[0]
    for k in range(n - 1):
        a, bb = r, rows[k + 1]
        piv, other = (bb, a) if sh.swap[k] else (a, bb)
        assert np.allclose(piv[k:], U[k, k:], rtol=1e-11, atol=1e-11 * np.abs(G).max())
        l = sh.mult()[k]
        r = other - l * piv
        assert abs(r[k]) <= 1e-10 * max(1.0, np.abs(G).max())
    assert np.allclose(r[n - 1], U[n - 1, n - 1], rtol=1e-10, atol=1e-10)
    assert np.allclose(sh.rdiag + 1j * sh.rdiag_im, 1.0 / np.diag(U), rtol=1e-14)
    assert sh.info == 0


def test_zshift_singular_known_answers(oracle):
    # core_tests.jl:235-239 and 321-323: I - J and -i I + i J of diag(1, 2) are singular in
    # the first coordinate
    J = np.array([[1, 0], [0, 2]], dtype=np.complex128, order="F")
    red = oracle.zreduce(J)
    assert oracle.zshift(red.Ht, 1, -1.0).info == 1
    assert oracle.zshift(red.Ht, -1j, 1j).info == 1
    assert oracle.zshift(red.Ht, 0, 0).info == 1


@pytest.mark.parametrize("n", [5, 64, 257])
def test_zldiv_and_refinement(oracle, n):
    J = crandn(500 + n, n)
    b = crandn(600 + n, n, 1)[:, 0]
    for s, t in ((1.0, -0.05 + 0.3j), (0.4 + 0.7j, -1.0)):
        red = oracle.zreduce(J)
        sh = oracle.zshift(red.Ht, s, t)
        W = s * np.eye(n) + t * J
        x = oracle.zldiv(red, sh, b)
        xd = np.linalg.solve(W, b)
        assert np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd)
        bwd = np.linalg.norm(b - W @ x, np.inf) / (np.linalg.norm(W, np.inf) * np.linalg.norm(x, np.inf)
                                                   + np.linalg.norm(b, np.inf))
        assert bwd < 50 * np.finfo(float).eps * max(1, n / 100)
        x2 = x + oracle.zldiv(red, sh, b - W @ x)
        bwd2 = np.linalg.norm(b - W @ x2, np.inf) / (np.linalg.norm(W, np.inf) * np.linalg.norm(x2, np.inf)
                                                     + np.linalg.norm(b, np.inf))
        assert bwd2 < 1e-13


def test_zldiv_high_precision(oracle):
    mp = pytest.importorskip("mpmath")
    mp.mp.prec = 200
    n = 9
    J = crandn(77, n)
    b = crandn(78, n, 1)[:, 0]
    s, t = 0.3 - 0.2j, -1.0 + 0.5j
    red = oracle.zreduce(J)
    x = oracle.zldiv(red, oracle.zshift(red.Ht, s, t), b)
    W = mp.matrix(n, n)
    for i in range(n):
        for j in range(n):
            W[i, j] = mp.mpc(s if i == j else 0) + mp.mpc(t) * mp.mpc(complex(J[i, j]))
    xe = mp.lu_solve(W, mp.matrix([mp.mpc(complex(v)) for v in b]))
    xe = np.array([complex(xe[i]) for i in range(n)])
    assert np.linalg.norm(x - xe) <= 1e-13 * np.linalg.norm(xe)


def test_zapplyZ_roundtrip_and_matrix(oracle):
    n = 47
    J = crandn(88, n)
    red = oracle.zreduce(J)
    x = crandn(89, n, 1)[:, 0]
    y = oracle.zapplyZ(red, x.copy())
    L = red.L
    D = np.diag(red.scale)
    P = np.eye(n)[:, red.perm] if False else None
    # Z = D P L with P the row permutation perm: (P y)[perm[i]] = y[i]
    Zd = np.zeros((n, n), dtype=complex)
    Zd[red.perm, :] = L
    Zd = D @ Zd
    assert np.linalg.norm(y - Zd @ x) <= 1e-13 * np.linalg.norm(y)
    back = oracle.zapplyZinv(red, y)
    assert np.linalg.norm(back - x) <= 1e-13 * np.linalg.norm(x)
