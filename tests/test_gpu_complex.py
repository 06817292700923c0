"""GPU parity for a complex J (LHLWorkspace{ComplexF64}): liblhl_hip.so through
the host mirror and its C ABI against the CPU oracle's complex restatement on
the same seeded inputs.

  * reduction: identical pivots (ipiv, perm) and balancing scales; factors bit
    for bit where the trailing update fits one column chunk (n <= 258), within
    1e-12*n beyond; J == Z H Z^-1 at n = 1024;
  * shift LU of the complex H: U, multipliers, swap, rdiag, info bit for bit
    against the oracle on the same Ht;
  * solves: the reference's complex testset (core_tests.jl:241-249, rtol 1e-9),
    the oracle's solve on the GPU's factors within 1e-12*n, one refinement step
    to bwd < 1e-13, a singular shift reported in info (core_tests.jl:235-239,
    321-323), applyZ / applyZinv, batched shifts refined to 1e-12, argument
    errors.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHIFTS = ((0.4 + 0.7j, -1.0), (1.0, -0.05 + 0.3j), (0.0, 1j), (-3.0 + 2j, 2.0 - 1j), (1.0, -0.05))


def say(*a):
    print(*a, flush=True)


def crandn(seed, n, m=None, dtype=np.complex128):
    rng = np.random.default_rng(seed)
    shape = (n, n if m is None else m)
    return np.asfortranarray((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype))


def relerr(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def zbwd(W, x, b):
    r = b - W @ x
    return np.linalg.norm(r, np.inf) / (np.linalg.norm(W, np.inf) * np.linalg.norm(x, np.inf)
                                        + np.linalg.norm(b, np.inf))


@pytest.mark.parametrize("n", [1, 2, 3, 5, 40, 129, 258, 260, 600])
@pytest.mark.parametrize("balance", [True, False])
def test_zreduction_matches_oracle(lhl_hip, oracle, n, balance):
    say(f"case: complex reduction n={n} balance={balance}")
    J = crandn(n, n)
    ws = lhl_hip.lhl(J, balance=balance)
    assert ws.dtype == np.complex128 and ws.shift.dtype == np.complex128
    red = oracle.zreduce(J, balance=balance)
    assert np.array_equal(ws.ipiv - 1, red.ipiv)
    assert np.array_equal(ws.perm - 1, red.perm)
    assert np.array_equal(ws.iperm - 1, red.iperm)
    assert np.array_equal(ws.scale, red.scale)
    F = ws.factors
    if n <= 258:
        assert np.array_equal(F, red.A), f"factors differ at n={n}: {relerr(F, red.A):.3e}"
    else:
        assert relerr(F, red.A) <= 1e-12 * n
    assert np.array_equal(ws.Ht, np.triu(F, -1).T)


def test_zreduction_graded_balancing(lhl_hip, oracle):
    # a graded complex matrix: balancing changes the scales
    n = 120
    rng = np.random.default_rng(3)
    g = 10.0 ** np.linspace(-6, 6, n)
    J = np.asfortranarray((g[:, None] / g[None, :]) * crandn(31, n))
    say("case: complex graded balancing")
    ws = lhl_hip.lhl(J)
    red = oracle.zreduce(J)
    assert not np.all(red.scale == 1.0)
    assert np.array_equal(ws.scale, red.scale)
    assert np.array_equal(ws.ipiv - 1, red.ipiv)
    assert np.array_equal(ws.factors, red.A)


def test_zreduction_reconstructs_n1024(lhl_hip, oracle):
    n = 1024
    say(f"case: complex reduction n={n}")
    J = crandn(1024, n)
    ws = lhl_hip.lhl(J)
    red = oracle.zreduce(J)
    assert np.array_equal(ws.ipiv - 1, red.ipiv)
    assert relerr(ws.factors, red.A) <= 1e-12 * n
    # J Z = Z H with the GPU's own factors (Z = D P L)
    F = ws.factors
    L = np.eye(n, dtype=complex)
    for c in range(n - 2):
        L[c + 2:, c + 1] = F[c + 2:, c]
    Z = np.zeros((n, n), dtype=complex)
    Z[ws.perm - 1, :] = L
    Z = ws.scale[:, None] * Z
    H = np.triu(F, -1)
    assert np.linalg.norm(J @ Z - Z @ H) <= 1e-12 * n * np.linalg.norm(J) * np.linalg.norm(Z)
    assert np.abs(np.tril(F, -2)).max() <= 1.0 + 1e-15


@pytest.mark.parametrize("n", [1, 2, 5, 40, 129, 511, 1100])
def test_zshift_bitwise(lhl_hip, oracle, n):
    say(f"case: complex-H shift bitwise n={n}")
    ws = lhl_hip.lhl(crandn(100 + n, n))
    Ht = ws.Ht
    for s, t in SHIFTS + ((0, 0),):
        lhl_hip.lhl_shift_(ws, s, t)
        ref = oracle.zshift(Ht, s, t)
        st = ws.shift
        assert st.info == ref.info
        assert np.array_equal(st.swap, ref.swap.astype(bool))
        assert np.array_equal(st.U, ref.U(), equal_nan=True)
        assert np.array_equal(st.multipliers, ref.mult(), equal_nan=True)
        assert np.array_equal(st.rdiag, ref.rdiag + 1j * ref.rdiag_im, equal_nan=True)


def test_zreference_complex_testset(lhl_hip):
    # core_tests.jl:241-249: n = 30, gamma = 0.2 + 0.3i, rtol 1e-9
    say("case: complex testset n=30")
    n = 30
    J = crandn(13, n)
    b = crandn(14, n, 1)[:, 0].copy()
    g = 0.2 + 0.3j
    ws = lhl_hip.lhl(J)
    lhl_hip.lhl_shift_(ws, 1, -g)
    x = lhl_hip.lhl_ldiv_(b.copy(), ws)
    xd = np.linalg.solve(np.eye(n) - g * J, b)
    assert np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd)


@pytest.mark.parametrize("n", [3, 40, 260, 600, 1100])
def test_zsolves_match_oracle(lhl_hip, oracle, n):
    say(f"case: complex solves n={n}")
    J = crandn(700 + n, n)
    ws = lhl_hip.lhl(J)
    red = oracle.zreduce(J)
    B = crandn(800 + n, n, 3)
    for s, t in SHIFTS:
        lhl_hip.lhl_shift_(ws, s, t)
        X = lhl_hip.lhl_ldiv_(B.copy(order="F"), ws)
        W = s * np.eye(n) + t * J
        sh = oracle.zshift(red.Ht, s, t)
        for c in range(B.shape[1]):
            xo = oracle.zldiv(red, sh, B[:, c])
            assert relerr(X[:, c], xo) <= 1e-12 * max(n, 10) * max(1.0, np.linalg.cond(W) / 1e3)
            assert zbwd(W, X[:, c], B[:, c]) < 50 * np.finfo(float).eps * max(1, n / 100)
        # one refinement step with the dense W, then with the resident J
        Xr = X.copy(order="F")
        lhl_hip.lhl_refine_(Xr, np.asfortranarray(W), B, ws, 1)
        Xj = X.copy(order="F")
        lhl_hip.lhl_refine_(Xj, None, B, ws, 1)
        for c in range(B.shape[1]):
            assert zbwd(W, Xr[:, c], B[:, c]) < 1e-13
            assert zbwd(W, Xj[:, c], B[:, c]) < 1e-13


def test_zsingular_shift_reported(lhl_hip):
    J = np.array([[1.0, 0.0], [0.0, 2.0]], dtype=np.complex128)
    ws = lhl_hip.lhl(J)
    lhl_hip.lhl_shift_(ws, 1, -1.0)
    assert ws.info == 1
    lhl_hip.lhl_shift_(ws, -1j, 1j)
    assert ws.info == 1


@pytest.mark.parametrize("n", [7, 300])
def test_zapplyZ(lhl_hip, oracle, n):
    say(f"case: complex applyZ n={n}")
    J = crandn(900 + n, n)
    ws = lhl_hip.lhl(J)
    red = oracle.zreduce(J)
    X = crandn(901, n, 2)
    Y = lhl_hip.applyZ_(X.copy(order="F"), ws)
    for c in range(2):
        assert relerr(Y[:, c], oracle.zapplyZ(red, X[:, c])) <= 1e-13 * n
    back = lhl_hip.applyZinv_(Y.copy(order="F"), ws)
    assert relerr(back, X) <= 1e-13 * n


def test_zbatched_refined(lhl_hip):
    n = 500
    say(f"case: complex batched n={n}")
    J = crandn(55, n)
    ws = lhl_hip.lhl(J)
    sig = np.array([s for s, _ in SHIFTS], dtype=complex)
    tau = np.array([t for _, t in SHIFTS], dtype=complex)
    offs = np.array([0, 2, 3, 6, 7, 9], dtype=np.int64)
    B = crandn(56, n, int(offs[-1]))
    X, info, iters = lhl_hip.shift_solve_batched(ws, sig, tau, B, offs, precision="refine64", tol=1e-12, max_iter=5)
    assert np.all(np.asarray(info) == 0)
    for s in range(len(sig)):
        W = sig[s] * np.eye(n) + tau[s] * J
        for c in range(offs[s], offs[s + 1]):
            r = B[:, c] - W @ X[:, c]
            assert np.abs(r).max() <= 1e-12 * np.abs(B[:, c]).max()
    # the same batch one shift at a time through the drop-in calls
    X1, _, _ = lhl_hip.shift_solve_batched(ws, sig, tau, B, offs)
    for s in range(len(sig)):
        lhl_hip.lhl_shift_(ws, sig[s], tau[s])
        Xs = lhl_hip.lhl_ldiv_(B[:, offs[s]:offs[s + 1]].copy(order="F"), ws)
        assert relerr(X1[:, offs[s]:offs[s + 1]], Xs) <= 1e-13 * n


def test_zcomplex64(lhl_hip, oracle):
    n = 200
    say(f"case: complex64 n={n}")
    J = crandn(66, n, dtype=np.complex64)
    ws = lhl_hip.lhl(J)
    assert ws.dtype == np.complex64
    b = crandn(67, n, 1, dtype=np.complex64)[:, 0].copy()
    lhl_hip.lhl_shift_(ws, 1, -0.05 + 0.3j)
    x = lhl_hip.lhl_ldiv_(b.copy(), ws)
    W = np.eye(n) - (0.05 - 0.3j) * J.astype(np.complex128)
    xd = np.linalg.solve(W, b.astype(np.complex128))
    # the reduction's own single-precision error (the oracle in complex64 loses
    # as much: about 2e-4 here, against 2.5e-7 for LAPACK's complex64 LU of W)
    red = oracle.zreduce(J, dtype=np.complex64)
    xo = oracle.zldiv(red, oracle.zshift(red.Ht, 1, -0.05 + 0.3j), b).astype(np.complex128)
    assert relerr(x.astype(np.complex128), xd) <= 4 * relerr(xo, xd)
    # one refinement step in complex64 with the resident J
    lhl_hip.lhl_refine_(x, None, b, ws, 1)
    assert relerr(x.astype(np.complex128), xd) <= relerr(xo, xd) / 3


def test_zargument_errors(lhl_hip):
    n = 6
    J = crandn(1, n)
    ws = lhl_hip.lhl(J)
    with pytest.raises(lhl_hip.ArgumentError):
        lhl_hip.LHLShift(ws, np.float64)          # TG must be the complex T
    lhl_hip.lhl_shift_(ws, 1, -0.1)
    with pytest.raises(lhl_hip.ArgumentError):
        lhl_hip.lhl_ldiv_(np.ones(n), ws)          # a real x cannot hold a complex solution
    wr = lhl_hip.LHLWorkspace(n, np.float64)
    with pytest.raises(lhl_hip.ArgumentError):
        lhl_hip.lhl_(wr, J)                        # a complex J on a real workspace
    # a real J on a complex workspace: the same reduction as its complex copy
    Jr = np.real(J).copy(order="F")
    wz = lhl_hip.LHLWorkspace(n, np.complex128)
    lhl_hip.lhl_(wz, Jr)
    wc = lhl_hip.lhl(Jr.astype(np.complex128))
    assert np.array_equal(wz.factors, wc.factors)
