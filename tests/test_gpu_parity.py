"""GPU parity: liblhl_hip.so (through the host mirror and its C ABI) against the
CPU oracle on the same seeded inputs.

  * reduction: identical pivots (ipiv, perm) and balancing scales, factors within
    1e-12*n relative (Frobenius) of the oracle's;
  * shift LU: the GPU state (U, multipliers, swap, rdiag, info) equals the
    oracle's bit for bit when both factor the same Ht (the reference's own
    bitwise spec, test/core_tests.jl:115-233);
  * solves and refinement: against the oracle and dense LAPACK within the
    reference's tolerances (rtol 1e-9 solves, bwd bounds).
"""
import numpy as np
import pytest

from inputs import PRECOMPILE_J, bwd, hilbert_plus_I, near_nilpotent, randn, stiff, tie_matrix, wilkinson

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def relerr(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def check_reduction(L, O, J, balance=True, blocked=-1, dtype=np.float64, tol_n=1e-12):
    n = J.shape[0]
    ws = L.lhl(J, balance=balance, dtype=dtype)
    red = O.reduce(J.astype(dtype), balance=balance, blocked=blocked)
    assert np.array_equal(ws.ipiv - 1, red.ipiv), f"ipiv differs at n={n}"
    assert np.array_equal(ws.perm - 1, red.perm)
    assert np.array_equal(ws.scale, red.scale)
    F = ws.factors
    assert relerr(F.astype(np.float64), red.A.astype(np.float64)) <= tol_n * max(n, 1)
    Ht = ws.Ht
    assert np.array_equal(Ht, np.triu(F, -1).T)
    return ws, red


def test_smoke_precompile_inputs(lhl_hip, oracle):
    say("case: precompile 4x4")
    J = PRECOMPILE_J
    ws, red = check_reduction(lhl_hip, oracle, J)
    lhl_hip.lhl_shift_(ws, 1.0, -0.1)
    x = lhl_hip.lhl_ldiv_(np.ones(4), ws)
    assert np.allclose(x, np.linalg.solve(np.eye(4) - 0.1 * J, np.ones(4)), rtol=1e-14)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 17, 80, 128, 200, 301, 400])
@pytest.mark.parametrize("balance", [True, False])
def test_reduction_matches_oracle_unblocked(lhl_hip, oracle, n, balance):
    say(f"case: reduction n={n} balance={balance}")
    check_reduction(lhl_hip, oracle, randn(n, n), balance=balance, blocked=0)


@pytest.mark.parametrize("n", [150, 200])
def test_reduction_ties(lhl_hip, oracle, n):
    say(f"case: tie reduction n={n}")
    check_reduction(lhl_hip, oracle, tie_matrix(9, n, -1, 1), blocked=0)


@pytest.mark.parametrize("n", [600, 1030])
def test_reduction_blocked_sizes(lhl_hip, oracle, n):
    say(f"case: reduction n={n} (oracle blocked)")
    check_reduction(lhl_hip, oracle, randn(n, n), blocked=1)


def test_reduction_float32(lhl_hip, oracle):
    say("case: float32 reduction n=300")
    # Two Float32 reductions differ by summation order alone (~n eps in the
    # factors), so each is compared to the Float64 reduction of the same J, as
    # test/core_tests.jl:428-449 does: the GPU must be as accurate as the oracle.
    n = 300
    J = randn(4, n, dtype=np.float32)
    ws = lhl_hip.lhl(J)
    red = oracle.reduce(J, blocked=0)
    r64 = oracle.reduce(J.astype(np.float64), blocked=0)
    assert np.array_equal(ws.ipiv - 1, red.ipiv)
    assert np.array_equal(red.ipiv, r64.ipiv)
    e_gpu = relerr(ws.factors.astype(np.float64), r64.A)
    e_orc = relerr(red.A.astype(np.float64), r64.A)
    assert e_gpu <= 4 * e_orc + 100 * n * np.finfo(np.float32).eps


def shift_state_equal(lhl_hip, oracle, ws, s, t, sh=None):
    Ht = ws.Ht
    if sh is None:
        lhl_hip.lhl_shift_(ws, s, t)
        st = ws.shift
    else:
        lhl_hip.lhl_shift_(sh, ws, s, t)
        st = sh
    if st.is_complex:
        ref = oracle.shift_c(Ht, s, t)
        refU = ref.U()
        refl = ref.mult()
        refrd = ref.rdiag + 1j * ref.rdiag_im
    else:
        ref = oracle.shift(Ht, s, t)
        refU, refl, refrd = ref.U(), ref.mult(), ref.rdiag
    assert st.info == ref.info
    assert np.array_equal(st.swap, ref.swap.astype(bool))
    assert np.array_equal(st.U, refU, equal_nan=True)
    assert np.array_equal(st.multipliers, refl, equal_nan=True)
    assert np.array_equal(st.rdiag, refrd, equal_nan=True)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("tie", [False, True])
def test_shift_bitwise(lhl_hip, oracle, dtype, tie):
    # The pipelined kernel's last block is full at 64, 127, 190 (FP32, 63-step
    # blocks) and 65, 129, 193 (FP64, 64-step blocks), partial at the others
    for n in [1, 2, 3, 5, 17, 64, 65, 126, 127, 128, 129, 130, 190, 193, 255, 256, 257, 511, 512, 520, 1000, 3000]:
        say(f"case: shift bitwise n={n} {np.dtype(dtype).name} tie={tie}")
        J = tie_matrix(n, n, dtype=dtype) if tie else randn(n, n, dtype=dtype)
        ws = lhl_hip.lhl(J)
        for s, t in ((1, -0.1), (0, 1), (1, -3), (2, 0), (0, 0)):
            shift_state_equal(lhl_hip, oracle, ws, s, t)


@pytest.mark.parametrize("dtype,n", [(np.float32, n) for n in [2500, 4096, 6000, 8192, 8513, 9601, 13440, 13441]]
                         + [(np.float64, n) for n in [2500, 4096, 4480, 4481, 5761, 9600, 9601, 13440, 13441]])
def test_shift_bitwise_large(lhl_hip, oracle, dtype, n):
    # the pipelined shift kernel at 6, 10, 14 and 19 columns (FP32) and 3 and 5
    # column pairs (FP64) per bulk thread, both sides of the switch from 448 to
    # 960 bulk threads (FP32 8512 / 8513, FP64 4480 / 4481) and of the group
    # counts of the 960-thread form, up to its size limit (13440), and the
    # one-step kernel for real shifts beyond it (13441)
    say(f"case: shift bitwise large n={n} {np.dtype(dtype).name}")
    ws = lhl_hip.lhl(randn(n, n, dtype=dtype))
    for s, t in ((1, -0.1), (0.5, -3)):
        shift_state_equal(lhl_hip, oracle, ws, s, t)


@pytest.mark.parametrize("tie", [False, True])
def test_complex_shift_bitwise(lhl_hip, oracle, tie):
    for n in [1, 2, 5, 40, 100, 256, 511]:
        say(f"case: complex shift bitwise n={n} tie={tie}")
        J = tie_matrix(n, n) if tie else randn(n, n)
        ws = lhl_hip.lhl(J, shift=np.complex128)
        for s, t in ((0.4 + 0.7j, -1), (1, -0.1 + 0.2j), (0, 1j), (2j, 0), (0, 0)):
            shift_state_equal(lhl_hip, oracle, ws, s, t)


def test_singular_shift_reported(lhl_hip):
    ws = lhl_hip.lhl(np.array([[1.0, 0.0], [0.0, 2.0]]))
    lhl_hip.lhl_shift_(ws, 1, -1.0)
    assert ws.info != 0
    wd = lhl_hip.lhl(np.array([[1.0, 0.0], [0.0, 2.0]]), shift=np.complex128)
    lhl_hip.lhl_shift_(wd, -1j, 1j)
    assert wd.info != 0


# 513, 1025 and 1100: the L sweeps' GEMMs split once (halves below 512 rows),
# twice, and twice with uneven halves
@pytest.mark.parametrize("n", [1, 2, 3, 5, 40, 129, 260, 400, 513, 1025, 1100])
def test_shifted_solves(lhl_hip, oracle, n):
    say(f"case: shifted solves n={n}")
    J = randn(n, n)
    b = np.random.default_rng(n + 7).standard_normal(n)
    ws = lhl_hip.lhl(J)
    red = oracle.reduce(J, blocked=0)
    for s, t in ((1.0, 0.0), (1.0, -1e-8), (1.0, -0.05), (1.0, 1.7), (0.0, 1.0), (-3.0, 2.0)):
        W = s * np.eye(n) + t * J
        if np.linalg.matrix_rank(W) < n:
            continue
        lhl_hip.lhl_shift_(ws, s, t)
        x = lhl_hip.lhl_ldiv_(b.copy(), ws)
        assert relerr(x, np.linalg.solve(W, b)) <= 1e-9
        xo = oracle.ldiv(red, oracle.shift(red.Ht, s, t), b)
        # two backward-stable solves of W differ by O(cond(W) eps): 1e-12 n, or
        # 20 cond(W) eps where W is that ill-conditioned
        eps = np.finfo(float).eps
        cw = np.linalg.cond(W)
        e, bw = relerr(x, xo), bwd(W, x, b)
        say(f"  shift ({s}, {t}): vs oracle {e:.2e} (cond(W) {cw:.1e}), bwd {bw / eps:.1f} eps")
        assert e <= max(1e-12 * max(n, 1), 20 * eps * cw)
        # test/core_tests.jl:54: 50 eps up to n = 500
        assert bw < 50 * eps * (1 if n <= 500 else n / 100)


@pytest.mark.parametrize("n", [3, 40, 129, 300])
def test_complex_solves(lhl_hip, n):
    say(f"case: complex solves n={n}")
    J = randn(n, n)
    rng = np.random.default_rng(n + 7)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ws = lhl_hip.lhl(J, shift=np.complex128)
    for s, t in ((0.4 + 0.7j, -1.0), (1.0, -0.05 + 0.3j), (0.0, 1j), (-3.0 + 2j, 2.0 - 1j), (1.0, -0.05)):
        W = s * np.eye(n) + t * J
        lhl_hip.lhl_shift_(ws, s, t)
        x = lhl_hip.lhl_ldiv_(b.copy(), ws)
        assert relerr(x, np.linalg.solve(W, b)) <= 1e-9
        y = lhl_hip.lhl_refine_(x.copy(), W, b, ws, 1)
        assert bwd(W, y, b) <= 2 * bwd(W, x, b) + np.finfo(float).eps
        assert bwd(W, y, b) < 50 * np.finfo(float).eps


def test_radau_real_and_complex_held_together(lhl_hip):
    n = 80
    J = randn(90, n)
    br = np.random.default_rng(91).standard_normal(n)
    rng = np.random.default_rng(92)
    bc = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ws = lhl_hip.lhl(J)
    sh = lhl_hip.LHLShift(ws, np.complex128)
    for g, gc in ((0.1, 0.05 + 0.1j), (0.2, 0.1 - 0.3j)):
        assert lhl_hip.lhl_shift_(ws, 1, -g) is ws
        assert lhl_hip.lhl_shift_(sh, ws, 1, -gc) is sh
        assert relerr(lhl_hip.lhl_ldiv_(br.copy(), ws), np.linalg.solve(np.eye(n) - g * J, br)) <= 1e-9
        assert relerr(lhl_hip.lhl_ldiv_(bc.copy(), sh, ws), np.linalg.solve(np.eye(n) - gc * J, bc)) <= 1e-9
        # complex rhs, real shift
        assert relerr(lhl_hip.lhl_ldiv_(bc.copy(), ws), np.linalg.solve(np.eye(n) - g * J, bc)) <= 1e-9
        Wc = np.eye(n) - gc * J
        xr = lhl_hip.lhl_refine_(lhl_hip.lhl_ldiv_(bc.copy(), sh, ws), Wc, bc, sh, ws, 1)
        assert relerr(xr, np.linalg.solve(Wc, bc)) <= 1e-12
        assert ws.info == 0 and sh.info == 0 and sh.sigma == 1 and sh.tau == -gc


def test_errors(lhl_hip):
    n = 20
    ws = lhl_hip.lhl(randn(93, n))
    with pytest.raises(lhl_hip.ArgumentError):
        lhl_hip.lhl_shift_(ws, 1, -0.3j)
    wc = lhl_hip.lhl(randn(93, n), shift=np.complex128)
    lhl_hip.lhl_shift_(wc, 1, -0.3j)
    with pytest.raises(lhl_hip.ArgumentError):
        lhl_hip.lhl_ldiv_(np.ones(n), wc)
    lhl_hip.lhl_shift_(ws, 1, -0.3)
    with pytest.raises(lhl_hip.DimensionMismatch):
        lhl_hip.lhl_ldiv_(np.ones(n + 1), ws)
    with pytest.raises(lhl_hip.DimensionMismatch):
        lhl_hip.lhl_ldiv_(np.ones(n, dtype=complex), lhl_hip.LHLShift(n // 2, np.complex128), ws)
    with pytest.raises(lhl_hip.DimensionMismatch):
        lhl_hip.applyZ_(np.ones(n - 1), ws)
    w32 = lhl_hip.lhl(randn(95, 3, dtype=np.float32))
    with pytest.raises(lhl_hip.ArgumentError):
        lhl_hip.lhl_shift_(lhl_hip.LHLShift(3, np.complex128), w32, 1, 1)
    # complex with zero imaginary parts is fine
    lhl_hip.lhl_shift_(ws, 1 + 0j, -0.3 + 0j)


def test_refinement_families(lhl_hip):
    n = 100
    b = np.random.default_rng(42).standard_normal(n)
    for name, J in (("randn", randn(1, n)), ("wilkinson", wilkinson(n))):
        say(f"case: refine family {name}")
        ws = lhl_hip.lhl(J)
        assert np.abs(np.tril(ws.factors, -2)).max() <= 1 + np.finfo(float).eps
        for g in (1e-3, 1.0, 1e3):
            lhl_hip.lhl_shift_(ws, 1, -g)
            if ws.info != 0:
                continue
            W = np.eye(n) - g * J
            x = lhl_hip.lhl_ldiv_(b.copy(), ws)
            assert lhl_hip.lhl_refine_(x, W, b, ws, 1) is x
            assert bwd(W, x, b) < 1e-13


def test_refinement_near_nilpotent(lhl_hip):
    n = 60
    J = near_nilpotent(17, n)
    b = np.random.default_rng(18).standard_normal(n)
    W = np.eye(n) - 0.5 * J
    ws = lhl_hip.lhl(J)
    lhl_hip.lhl_shift_(ws, 1, -0.5)
    raw = lhl_hip.lhl_ldiv_(b.copy(), ws)
    ref = lhl_hip.lhl_refine_(raw.copy(), W, b, ws, 1)
    assert bwd(W, ref, b) <= bwd(W, raw, b)
    assert bwd(W, ref, b) < 1e-13


def test_unrefined_solve_large_kappa_z(lhl_hip, oracle):
    """Unrefined solves where Z = D P L is ill-conditioned (the reference's
    families, test/core_tests.jl:343-380, where kappa(Z) reaches 1e10): the
    drop-in solve's backward error stays within 10x the oracle's own unrefined
    solve on the same factors, a batched solve (the product with the explicit
    L'^-1) within 30x; one refinement step then reaches 1e-13."""
    from inputs import graded
    eps = np.finfo(float).eps
    n = 64
    b = np.random.default_rng(77).standard_normal(n)
    for name, J, gams in (("near-nilpotent", near_nilpotent(17, n), (0.5, 1e-3)),
                          ("graded", graded(5, n), (1e-3, 1.0)),
                          ("wilkinson", wilkinson(n), (1e-3, 1.0))):
        ws = lhl_hip.lhl(J)
        red = oracle.reduce(J, blocked=0)
        Z = lhl_hip.applyZ_(np.asfortranarray(np.eye(n)), ws)
        kz = np.linalg.cond(Z)
        for g in gams:
            W = np.eye(n) - g * J
            lhl_hip.lhl_shift_(ws, 1, -g)
            if ws.info != 0:
                continue
            x = lhl_hip.lhl_ldiv_(b.copy(), ws)
            xo = oracle.ldiv(red, oracle.shift(red.Ht, 1.0, -g), b)
            X, info, _ = lhl_hip.shift_solve_batched(ws, np.ones(1), np.array([-g]), b.reshape(n, 1).copy(),
                                                     np.array([0, 1]))
            bx, bo, bb = bwd(W, x, b), bwd(W, xo, b), bwd(W, X[:, 0], b)
            say(f"  {name} gamma={g}: kappa(Z) {kz:.1e}, bwd drop-in {bx:.1e}, batched {bb:.1e}, oracle {bo:.1e}")
            assert bx <= 10 * bo + 100 * eps
            # the explicit inverse's error grows with cond(L'): 9.9x the oracle's
            # at kappa(Z) = 4.7e8 (near-nilpotent, gamma = 1e-3)
            assert bb <= 30 * bo + 100 * eps
            assert bwd(W, lhl_hip.lhl_refine_(x, W, b, ws, 1), b) < 1e-13


def test_applyZ_roundtrip(lhl_hip, oracle):
    n = 130
    J = randn(130, n)
    ws = lhl_hip.lhl(J)
    red = oracle.reduce(J, blocked=0)
    x0 = np.random.default_rng(5).standard_normal(n)
    z = lhl_hip.applyZ_(x0.copy(), ws)
    zo = oracle.applyZ(red, np.asfortranarray(x0.copy()))
    assert relerr(z, zo) <= 1e-12 * n
    back = lhl_hip.applyZinv_(z.copy(), ws)
    assert relerr(back, x0) <= 1e-12 * n


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_reduction_two_pass_rows_properties(lhl_hip, dtype):
    # n = 4500 takes the step loop's two-passes-per-wave shape, beyond the
    # sizes the oracle finishes in seconds: check what the mathematics fixes.
    # J Z = Z H with Z = D P L; partial pivoting bounds every multiplier by 1
    # (src:470-491); the pivots are rows below the diagonal.
    n = 4500
    say(f"case: reduction properties n={n} {np.dtype(dtype).name}")
    J = stiff(4500, n)
    ws = lhl_hip.lhl(J, dtype=dtype)
    F = ws.factors.astype(np.float64)
    ipiv = ws.ipiv - 1
    assert np.all(ipiv >= np.arange(1, n - 1)) and np.all(ipiv < n)
    assert np.array_equal(np.sort(ws.perm - 1), np.arange(n))
    L = np.tril(F, -2)
    assert np.all(np.abs(L) <= 1.0)
    H = np.triu(F, -1)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(n)
    zx = lhl_hip.applyZ_(x.copy(), ws)
    zhx = lhl_hip.applyZ_(H @ x, ws)
    rel = np.linalg.norm(J @ zx - zhx) / (np.linalg.norm(J, np.inf) * np.linalg.norm(zx))
    say(f"  |J Z x - Z H x| / (|J| |Z x|) = {rel:.2e}")
    assert rel <= (1e-13 if dtype == np.float64 else 1e-5)


def test_workspace_reuse_and_resize(lhl_hip):
    n = 50
    J = randn(3, n)
    b = np.random.default_rng(4).standard_normal(n)
    ws = lhl_hip.lhl(J)
    for g in (0.02, 1e-6, 5.0, 0.02):
        lhl_hip.lhl_shift_(ws, 1, -g)
        assert relerr(lhl_hip.lhl_ldiv_(b.copy(), ws), np.linalg.solve(np.eye(n) - g * J, b)) <= 1e-9
    J2 = randn(5, n)
    assert lhl_hip.lhl_(ws, J2) is ws
    lhl_hip.lhl_shift_(ws, 1, -0.3)
    assert relerr(lhl_hip.lhl_ldiv_(b.copy(), ws), np.linalg.solve(np.eye(n) - 0.3 * J2, b)) <= 1e-9
    J3 = randn(6, 2 * n)
    lhl_hip.lhl_(ws, J3)
    lhl_hip.lhl_shift_(ws, 1, -0.3)
    b3 = np.random.default_rng(7).standard_normal(2 * n)
    assert relerr(lhl_hip.lhl_ldiv_(b3.copy(), ws), np.linalg.solve(np.eye(2 * n) - 0.3 * J3, b3)) <= 1e-9


def test_config1_readme_workflow(lhl_hip, oracle):
    """J = randn(400,400), b = randn(400); lhl, then for gamma in (0.01, 0.013,
    0.021): shift (1,-gamma), ldiv, refine 1 (README.md:40-52)."""
    say("case: config 1")
    n = 400
    J = randn(400, n)
    b = np.random.default_rng(401).standard_normal(n)
    ws = lhl_hip.lhl(J)
    red = oracle.reduce(J)
    assert np.array_equal(ws.ipiv - 1, red.ipiv)
    assert np.array_equal(ws.perm - 1, red.perm)
    for g in (0.01, 0.013, 0.021):
        W = np.eye(n) - g * J
        lhl_hip.lhl_shift_(ws, 1, -g)
        x = lhl_hip.lhl_ldiv_(b.copy(), ws)
        sh = oracle.shift(red.Ht, 1.0, -g)
        xo = oracle.ldiv(red, sh, b)
        assert relerr(x, xo) <= 1e-12 * n
        x = lhl_hip.lhl_refine_(x, W, b, ws, 1)
        xo = oracle.refine(red, sh, W, b, xo, 1)
        assert relerr(x, xo) <= 1e-12 * n
        assert bwd(W, x, b) < 1e-15


@pytest.mark.parametrize("n", [2000])
def test_config2_stiff_fp64(lhl_hip, oracle, n):
    say(f"case: config 2 stiff n={n}")
    J = stiff(2000 + n, n)
    ws = lhl_hip.lhl(J)
    red = oracle.reduce(J)          # blocked oracle, as the reference at n >= 500
    assert np.array_equal(ws.ipiv - 1, red.ipiv)
    assert np.array_equal(ws.scale, red.scale)
    assert relerr(ws.factors, red.A) <= 1e-12 * n
    rng = np.random.default_rng(n)
    for g in np.logspace(-4, -1, 8):
        b = rng.standard_normal(n)
        lhl_hip.lhl_shift_(ws, 1, -g)
        x = lhl_hip.lhl_ldiv_(b.copy(), ws)
        xo = oracle.ldiv(red, oracle.shift(red.Ht, 1.0, -g), b)
        assert relerr(x, xo) <= 1e-12 * n
        W = np.eye(n) - g * J
        assert np.linalg.norm(b - W @ x, np.inf) / np.linalg.norm(b, np.inf) <= 1e-10


def test_batched_reports_singular_shifts(lhl_hip):
    # J diagonal with integer entries: H = J exactly, so sigma = J[i,i],
    # tau = -1 makes pivot i of the shift LU exactly zero (info = i + 1, src:2450).
    # The batch must read each shift's info after its LU has run.
    n = 64
    J = np.diag(np.arange(1.0, n + 1))
    ws = lhl_hip.lhl(J, balance=False)
    sig = np.array([1.7, 5.0, 2.0, 40.0, 0.5])
    tau = np.array([-0.5, -1.0, -0.13, -1.0, 1.0])
    B = np.asfortranarray(np.random.default_rng(2).standard_normal((n, 5)))
    _, info, _ = lhl_hip.shift_solve_batched(ws, sig, tau, B, np.arange(6))
    single = []
    for s, t in zip(sig, tau):
        lhl_hip.lhl_shift_(ws, s, t)
        single.append(ws.info)
    assert list(info) == single
    assert single == [0, 5, 0, 40, 0]


def test_batched_two_chunks(lhl_hip, monkeypatch):
    # 130 FP64 shifts at n = 4096 exceed an 8 GiB bound on one chunk's packed
    # U's (127 shifts), so the batch runs as two chunks of 65 that reuse the
    # shift pool; every shift must match its single-shift solve.
    monkeypatch.setenv("LHL_SHIFT_POOL_BYTES", str(8 << 30))
    n, S = 4096, 130
    say(f"case: batched two chunks n={n} shifts={S}")
    J = stiff(130, n)
    ws = lhl_hip.lhl(J)
    gam = np.logspace(-4, -1, S)
    B = np.asfortranarray(np.random.default_rng(5).standard_normal((n, S)))
    X, info, _ = lhl_hip.shift_solve_batched(ws, np.ones(S), -gam, B, np.arange(S + 1))
    assert np.all(info == 0)
    for s in (0, 1, 64, 65, 128, 129):
        lhl_hip.lhl_shift_(ws, 1, -gam[s])
        x = lhl_hip.lhl_ldiv_(B[:, s].copy(), ws)
        say(f"  shift {s}: batched vs single {relerr(X[:, s], x):.2e}")
        assert relerr(X[:, s], x) <= 1e-13


def test_batched_one_shift_per_chunk(lhl_hip, monkeypatch):
    # a 1-byte bound on a chunk's packed U's leaves one shift per chunk; with
    # uneven right-hand-side counts every chunk's column offsets differ, and
    # the result must equal the one-chunk call's (rocBLAS may pick other GEMM
    # kernels for other widths, hence the tolerance)
    n = 300
    J = stiff(77, n)
    gam = np.logspace(-4, -1, 6)
    offs = np.concatenate([[0], np.cumsum([1, 2, 3, 1, 2, 4])]).astype(np.int64)
    B = np.asfortranarray(np.random.default_rng(9).standard_normal((n, int(offs[-1]))))
    for dtype, precision in ((np.float64, "working"), (np.float32, "refine64")):
        say(f"case: batched one shift per chunk {np.dtype(dtype).name} {precision}")
        ws = lhl_hip.lhl(J, dtype=dtype)
        X1, info1, _ = lhl_hip.shift_solve_batched(ws, np.ones(6), -gam, B, offs, precision=precision, tol=1e-12,
                                                   max_iter=30)
        monkeypatch.setenv("LHL_SHIFT_POOL_BYTES", "1")
        X6, info6, _ = lhl_hip.shift_solve_batched(ws, np.ones(6), -gam, B, offs, precision=precision, tol=1e-12,
                                                   max_iter=30)
        monkeypatch.delenv("LHL_SHIFT_POOL_BYTES")
        assert np.array_equal(info1, info6) and np.all(info1 == 0)
        assert relerr(X6, X1) <= 1e-12


def test_batched_matches_single(lhl_hip):
    n = 300
    J = stiff(11, n)
    ws = lhl_hip.lhl(J)
    gam = np.logspace(-4, -1, 6)
    sig = np.ones(6)
    tau = -gam
    offs = np.array([0, 2, 3, 5, 6, 8, 10])
    rng = np.random.default_rng(12)
    B = np.asfortranarray(rng.standard_normal((n, 10)))
    X, info, _ = lhl_hip.shift_solve_batched(ws, sig, tau, B, offs)
    assert np.all(info == 0)
    for s in range(6):
        lhl_hip.lhl_shift_(ws, 1, -gam[s])
        for c in range(offs[s], offs[s + 1]):
            x = lhl_hip.lhl_ldiv_(B[:, c].copy(), ws)
            say(f"  column {c}: batched vs single {relerr(X[:, c], x):.2e}")
            assert relerr(X[:, c], x) <= 1e-13
    # complex shifts s = i*omega on a real reduction, tau = -1
    om = np.logspace(-2, 2, 4)
    Bc = np.asfortranarray(rng.standard_normal((n, 8)) + 1j * rng.standard_normal((n, 8)))
    Xc, info, _ = lhl_hip.shift_solve_batched(ws, 1j * om, -np.ones(4), Bc, np.arange(0, 9, 2))
    for s in range(4):
        W = 1j * om[s] * np.eye(n) - J
        for c in range(2 * s, 2 * s + 2):
            assert relerr(Xc[:, c], np.linalg.solve(W, Bc[:, c])) <= 1e-9


def test_precision_form_refines_to_1e12(lhl_hip):
    """FP32 factors of an FP64 J, FP64-residual refinement until ||r||/||b|| <= 1e-12."""
    n = 1024
    J = stiff(4096, n)
    ws = lhl_hip.lhl(J, dtype=np.float32)
    gam = np.logspace(-4, -1, 4)
    offs = np.arange(0, 4 * 4 + 1, 4)
    B = np.asfortranarray(np.random.default_rng(3).standard_normal((n, 16)))
    X, info, iters = lhl_hip.shift_solve_batched(ws, np.ones(4), -gam, B, offs, precision="refine64", tol=1e-12,
                                                 max_iter=30)
    for s in range(4):
        W = np.eye(n) - gam[s] * J
        for c in range(offs[s], offs[s + 1]):
            r = np.linalg.norm(B[:, c] - W @ X[:, c], np.inf) / np.linalg.norm(B[:, c], np.inf)
            assert r <= 1e-12, (s, c, r, iters)


def test_precision_form_n8192(lhl_hip):
    """The precision form at n = 8192 (the larger size of its workload): FP32
    factors of the stiff FP64 J, FP64-residual refinement to 1e-12 per column,
    and agreement with the FP64 path's solutions."""
    n, S, R = 8192, 4, 2
    say(f"case: precision form n={n} shifts={S} rhs={R}")
    J = stiff(8192, n)
    gam = np.logspace(-4, -1, S)
    offs = np.arange(0, S * R + 1, R, dtype=np.int64)
    B = np.asfortranarray(np.random.default_rng(8).standard_normal((n, S * R)))
    ws32 = lhl_hip.lhl(J, dtype=np.float32)
    X32, info, iters = lhl_hip.shift_solve_batched(ws32, np.ones(S), -gam, B, offs, precision="refine64",
                                                   tol=1e-12, max_iter=30)
    assert np.all(info == 0)
    del ws32
    ws64 = lhl_hip.lhl(J)
    X64, info64, _ = lhl_hip.shift_solve_batched(ws64, np.ones(S), -gam, B, offs)
    assert np.all(info64 == 0)
    del ws64
    JX = J @ X32
    for s in range(S):
        for c in range(offs[s], offs[s + 1]):
            r = np.abs(B[:, c] - (X32[:, c] - gam[s] * JX[:, c])).max() / np.abs(B[:, c]).max()
            assert r <= 1e-12, (s, c, r, iters)
            assert relerr(X32[:, c], X64[:, c]) <= 1e-9, (s, c)


def test_solves_above_packed_column_limit(lhl_hip):
    """n = 8200, above the n = 8192 up to which U is also kept in packed
    columns: the back substitution runs from U's packed rows (real and complex
    shifts).  Residuals of FP64 solves, FP32 solves refined to 1e-12, and a
    complex-shift solve on the real reduction."""
    n, S, R = 8200, 2, 2
    say(f"case: solves above the packed-column limit n={n}")
    J = stiff(8200, n)
    gam = np.array([1e-3, 1e-1])
    offs = np.arange(0, S * R + 1, R, dtype=np.int64)
    B = np.asfortranarray(np.random.default_rng(12).standard_normal((n, S * R)))

    def resid(X, s, cols, shift):
        W_X = X[:, cols] - shift * (J @ X[:, cols])
        return np.abs(B[:, cols] - W_X).max() / np.abs(B[:, cols]).max()

    for dtype, precision, bound in ((np.float64, "working", 1e-10), (np.float32, "refine64", 1e-12)):
        ws = lhl_hip.lhl(J, dtype=dtype)
        X, info, _ = lhl_hip.shift_solve_batched(ws, np.ones(S), -gam, B, offs, precision=precision, tol=1e-12,
                                                 max_iter=30)
        assert np.all(info == 0)
        for s in range(S):
            assert resid(X, s, slice(offs[s], offs[s + 1]), gam[s]) <= bound, (np.dtype(dtype).name, s)
        del ws
    ws = lhl_hip.lhl(J, shift=np.complex128)
    om = 0.5
    bc = (B[:, 0] + 1j * B[:, 1]).copy()
    lhl_hip.lhl_shift_(ws, 1j * om, -1.0)
    x = lhl_hip.lhl_ldiv_(bc.copy(), ws)
    r = np.abs(bc - (1j * om * x - J @ x)).max() / np.abs(bc).max()
    assert r <= 1e-10, r


def _rel_resid(W, X, B):
    R = B - W @ X
    return float(np.max(np.abs(R).max(axis=0) / np.abs(B).max(axis=0)))


@pytest.mark.parametrize("dtype,precision", [(np.float64, "working"), (np.float32, "refine64")])
def test_config4_complex_and_real_sweep(lhl_hip, dtype, precision):
    """The batched shift sweep of one real reduction (scaled down from n=4096,
    256+256 shifts): complex shifts (i*omega, -1) and real shifts (1, -gamma),
    8 right-hand sides each, A = G/sqrt(n) - 2I."""
    n, S, R = 512, 32, 8
    say(f"case: config-4 sweep n={n} {np.dtype(dtype).name} {precision}")
    rng = np.random.default_rng(404)
    A = np.asfortranarray(rng.standard_normal((n, n)) / np.sqrt(n) - 2.0 * np.eye(n))
    ws = lhl_hip.lhl(A, dtype=dtype, shift=np.complex128 if dtype == np.float64 else np.complex64)
    om = np.logspace(-2, 2, S)
    offs = np.arange(0, S * R + 1, R, dtype=np.int64)
    Bc = np.asfortranarray(rng.standard_normal((n, S * R)) + 1j * rng.standard_normal((n, S * R)))
    Xc, info, iters = lhl_hip.shift_solve_batched(ws, 1j * om, -np.ones(S), Bc, offs, precision=precision,
                                                  tol=1e-12, max_iter=30)
    assert np.all(info == 0)
    bound = 1e-12 if precision == "refine64" else 1e-11
    for s in range(0, S, 5):
        W = 1j * om[s] * np.eye(n) - A
        cols = slice(offs[s], offs[s + 1])
        assert _rel_resid(W, Xc[:, cols], Bc[:, cols]) <= bound, (s, iters[s])
    gam = np.logspace(-4, -1, S)
    Br = np.asfortranarray(rng.standard_normal((n, S * R)))
    Xr, info, iters = lhl_hip.shift_solve_batched(ws, np.ones(S), -gam, Br, offs, precision=precision, tol=1e-12,
                                                  max_iter=30)
    assert np.all(info == 0)
    for s in range(0, S, 5):
        W = np.eye(n) - gam[s] * A
        cols = slice(offs[s], offs[s + 1])
        assert _rel_resid(W, Xr[:, cols], Br[:, cols]) <= bound, (s, iters[s])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_handles_on_host_threads(lhl_hip, dtype):
    """Handles driven from separate host threads at once (each handle on its
    own stream, as the uneven batch of independent matrices runs) give the
    results of the same calls made one after another: identical pivots, and
    solutions equal to rocBLAS's run-to-run rounding."""
    from concurrent.futures import ThreadPoolExecutor

    sizes = [700, 300, 1200, 129, 520, 64]
    say(f"case: handles on host threads {np.dtype(dtype).name} n={sizes}")
    mats = [stiff(600 + i, n) for i, n in enumerate(sizes)]

    def run(i):
        n, S = sizes[i], 5
        ws = lhl_hip.lhl(mats[i], dtype=dtype)
        B = np.asfortranarray(np.random.default_rng(i).standard_normal((n, 2 * S)))
        X, info, iters = lhl_hip.shift_solve_batched(ws, np.ones(S), -np.logspace(-4, -1, S), B,
                                                     np.arange(0, 2 * S + 1, 2), precision="refine64", tol=1e-12,
                                                     max_iter=30)
        assert np.all(info == 0)
        return ws.ipiv.copy(), X, iters

    one = [run(i) for i in range(len(sizes))]
    with ThreadPoolExecutor(3) as ex:
        many = list(ex.map(run, range(len(sizes))))
    for n, (p1, x1, i1), (p2, x2, i2) in zip(sizes, one, many):
        for name, x, it in (("sequential", x1, i1), ("threaded", x2, i2)):
            bad = np.flatnonzero(~np.all(np.isfinite(x), axis=0))
            if bad.size:
                say(f"  n={n} {name}: non-finite columns {bad.tolist()}, refinement rounds per shift {it.tolist()}")
        assert np.array_equal(p1, p2)
        assert relerr(x2, x1) <= 1e-13


def test_config5_uneven_batches(lhl_hip):
    """Independent stiff matrices of uneven n (scaled down from 64 matrices,
    n in [128, 3000]), each with its own shifts and right-hand-side counts,
    FP64 and FP32+refine to 1e-12."""
    rng = np.random.default_rng(505)
    for m in range(6):
        n = int(rng.integers(128, 701))
        S = int(rng.integers(4, 13))
        counts = rng.integers(1, 9, size=S)
        offs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        say(f"case: config-5 matrix {m} n={n} shifts={S} rhs={int(offs[-1])}")
        J = stiff(1000 + m, n)
        gam = np.logspace(-4, -1, S)
        B = np.asfortranarray(rng.standard_normal((n, int(offs[-1]))))
        for dtype, precision, bound in ((np.float64, "working", 1e-10), (np.float32, "refine64", 1e-12)):
            ws = lhl_hip.lhl(J, dtype=dtype)
            X, info, iters = lhl_hip.shift_solve_batched(ws, np.ones(S), -gam, B, offs, precision=precision,
                                                         tol=1e-12, max_iter=30)
            assert np.all(info == 0)
            for s in range(S):
                W = np.eye(n) - gam[s] * J
                cols = slice(offs[s], offs[s + 1])
                assert _rel_resid(W, X[:, cols], B[:, cols]) <= bound, (m, n, s, precision, iters[s])
