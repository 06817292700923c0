"""GPU parity at the full sizes of the benchmark workloads.

  * n = 4096 stiff J (the headline workload's matrix): the reduction against
    the oracle's (identical pivots and scales, factors within 1e-12*n), the
    plain FP64 drop-in calls (8 shifts x 1 right-hand side through
    lhl_shift!/lhl_ldiv!) against the oracle's solves, and the batched
    FP32+refine solve of 64 shifts x 16 right-hand sides: every column's FP64
    residual ||b - (I - gamma J) x||_inf / ||b||_inf <= 1e-12, and four shifts
    spread over gamma against the oracle's FP64 solves within 1e-12*n;
  * the complex + real shift sweep at n = 4096 (256 + 256 shifts x 8
    right-hand sides), FP32+refine and FP64 refined to 1e-12: every column;
  * the uneven batch of 64 independent stiff J (n in [128, 3000]): every
    column <= 1e-12 in both precisions, and for n <= 1000 the oracle's pivots
    and sampled solves.

Residuals are computed on the GPU in FP64 with torch (dense products, no
library code involved).
"""
import numpy as np
import pytest

from inputs import stiff

pytestmark = pytest.mark.gpu

N = 4096


def say(*a):
    print(*a, flush=True)


def relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available()
    return torch, torch.device("cuda:0")


@pytest.fixture(scope="module")
def big(lhl_hip, oracle):
    """The headline matrix (the benchmark's seed) and the oracle's reduction of it."""
    J = stiff(4096, N)
    say(f"case: oracle reduction n={N} (blocked, OpenMP)")
    red = oracle.reduce(J)
    return J, red


def col_resid(torch, dev, J, X, B, sig, tau, R):
    """max over columns of ||b - (sigma x + tau J x)||_inf / ||b||_inf, on the GPU in FP64."""
    Jd = torch.from_numpy(np.asfortranarray(J)).to(dev)
    Xd = torch.from_numpy(np.asfortranarray(X)).to(dev)
    Bd = torch.from_numpy(np.asfortranarray(B)).to(dev)
    if Xd.is_complex():
        Jd = Jd.to(torch.complex128)
    s = torch.from_numpy(np.repeat(np.asarray(sig), R)).to(dev)
    t = torch.from_numpy(np.repeat(np.asarray(tau), R)).to(dev)
    Rd = Bd - (Xd * s[None, :] + (Jd @ Xd) * t[None, :])
    r = (Rd.abs().amax(0) / Bd.abs().amax(0)).cpu().numpy()
    return r


def test_reduction_n4096_matches_oracle(lhl_hip, big):
    J, red = big
    say(f"case: reduction n={N} against the oracle")
    ws = lhl_hip.lhl(J)
    assert np.array_equal(ws.ipiv - 1, red.ipiv), "pivot sequence differs from the oracle at n=4096"
    assert np.array_equal(ws.perm - 1, red.perm)
    assert np.array_equal(ws.scale, red.scale)
    e = relerr(ws.factors, red.A)
    say(f"  factors rel err {e:.2e}")
    assert e <= 1e-12 * N


def test_config2_dropin_n4096(lhl_hip, oracle, big):
    """The plain FP64 drop-in at n = 4096: one lhl(), then per shift
    lhl_shift!(ws, 1, -gamma) and lhl_ldiv! on one N(0,1) right-hand side,
    against the oracle's shift and solve on the oracle's reduction."""
    J, red = big
    ws = lhl_hip.lhl(J)
    rng = np.random.default_rng(N)
    for g in np.logspace(-4, -1, 8):
        say(f"case: config 2 n={N} gamma={g:.3g}")
        b = rng.standard_normal(N)
        assert lhl_hip.lhl_shift_(ws, 1.0, -g) is ws
        assert ws.info == 0
        x = lhl_hip.lhl_ldiv_(b.copy(), ws)
        xo = oracle.ldiv(red, oracle.shift(red.Ht, 1.0, -g), b)
        assert relerr(x, xo) <= 1e-12 * N
        r = np.abs(b - (x - g * (J @ x))).max() / np.abs(b).max()
        assert r <= 1e-10, r


def test_config3_headline_n4096(lhl_hip, oracle, big, torch_dev):
    """The headline workload: FP32 factors, 64 shifts x 16 right-hand sides,
    FP64-residual refinement to 1e-12 for every column; the FP32+refine
    solutions of four shifts spread over gamma against the oracle's FP64
    solves (oracle shift and solve on the oracle's FP64 reduction)."""
    torch, dev = torch_dev
    J, red = big
    S, R = 64, 16
    gam = np.logspace(-4, -1, S)
    offs = np.arange(0, S * R + 1, R, dtype=np.int64)
    B = np.asfortranarray(np.random.default_rng(7).standard_normal((N, S * R)))
    say(f"case: config 3 headline n={N} {S}x{R} FP32+refine")
    ws = lhl_hip.lhl(J, dtype=np.float32)
    X, info, iters = lhl_hip.shift_solve_batched(ws, np.ones(S), -gam, B, offs, precision="refine64", tol=1e-12,
                                                 max_iter=30)
    assert np.all(info == 0)
    r = col_resid(torch, dev, J, X, B, np.ones(S), -gam, R)
    say(f"  max col rel resid {r.max():.2e}, refine rounds {iters.min()}-{iters.max()}")
    assert np.all(r <= 1e-12), np.argmax(r)
    for s in (0, 21, 42, 63):
        sh = oracle.shift(red.Ht, 1.0, -gam[s])
        c = offs[s]
        xo = oracle.ldiv(red, sh, B[:, c])
        e = relerr(X[:, c], xo)
        say(f"  shift {s} gamma={gam[s]:.3g}: FP32+refine vs oracle FP64 {e:.2e}")
        assert e <= 1e-12 * N


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_config4_full_sweep(lhl_hip, torch_dev, dtype):
    """n = 4096, A = G/sqrt(n) - 2I: 256 complex shifts (i*omega, -1) and 256
    real shifts (1, -gamma), 8 right-hand sides each, refined to 1e-12."""
    torch, dev = torch_dev
    S, R = 256, 8
    rng = np.random.default_rng(404)
    A = np.asfortranarray(rng.standard_normal((N, N)) / np.sqrt(N) - 2.0 * np.eye(N))
    om = np.logspace(-2, 2, S)
    gam = np.logspace(-4, -1, S)
    offs = np.arange(0, S * R + 1, R, dtype=np.int64)
    Bc = np.asfortranarray(rng.standard_normal((N, S * R)) + 1j * rng.standard_normal((N, S * R)))
    Br = np.asfortranarray(rng.standard_normal((N, S * R)))
    say(f"case: config 4 full sweep n={N} {np.dtype(dtype).name}")
    ws = lhl_hip.lhl(A, dtype=dtype)
    Xc, info, itc = lhl_hip.shift_solve_batched(ws, 1j * om, -np.ones(S), Bc, offs, precision="refine64",
                                                tol=1e-12, max_iter=30)
    assert np.all(info == 0)
    rc = col_resid(torch, dev, A, Xc, Bc, 1j * om, -np.ones(S), R)
    Xr, info, itr = lhl_hip.shift_solve_batched(ws, np.ones(S), -gam, Br, offs, precision="refine64", tol=1e-12,
                                                max_iter=30)
    assert np.all(info == 0)
    rr = col_resid(torch, dev, A, Xr, Br, np.ones(S), -gam, R)
    say(f"  complex max {rc.max():.2e} (rounds {itc.min()}-{itc.max()}), real max {rr.max():.2e} "
        f"(rounds {itr.min()}-{itr.max()})")
    assert np.all(rc <= 1e-12) and np.all(rr <= 1e-12)


def test_config5_full_uneven(lhl_hip, oracle, torch_dev):
    """64 independent stiff J, n uniform in [128, 3000], 4-32 shifts x 1-8
    right-hand sides each (the benchmark's seeds), FP64 and FP32+refine to
    1e-12 for every column; for n <= 1000 the oracle's pivots, and one solve
    per matrix against the oracle's."""
    torch, dev = torch_dev
    rng = np.random.default_rng(505)
    M = 64
    ns = rng.integers(128, 3001, M)
    nsh = rng.integers(4, 33, M)
    nrhs = rng.integers(1, 9, M)
    worst = {np.float64: 0.0, np.float32: 0.0}
    checked = 0
    for m in range(M):
        n, S, R = int(ns[m]), int(nsh[m]), int(nrhs[m])
        J = stiff(5000 + m, n)
        B = np.asfortranarray(rng.standard_normal((n, S * R)))
        gam = np.logspace(-4, -1, S)
        offs = np.arange(0, S * R + 1, R, dtype=np.int64)
        say(f"case: config 5 matrix {m} n={n} shifts={S} rhs={R}")
        for dtype in (np.float64, np.float32):
            ws = lhl_hip.lhl(J, dtype=dtype)
            X, info, _ = lhl_hip.shift_solve_batched(ws, np.ones(S), -gam, B, offs, precision="refine64",
                                                     tol=1e-12, max_iter=30)
            assert np.all(info == 0)
            r = col_resid(torch, dev, J, X, B, np.ones(S), -gam, R)
            worst[dtype] = max(worst[dtype], float(r.max()))
            assert np.all(r <= 1e-12), (m, n, np.dtype(dtype).name, float(r.max()))
            if dtype == np.float64 and n <= 1000:
                red = oracle.reduce(J)
                assert np.array_equal(ws.ipiv - 1, red.ipiv), (m, n)
                sh = oracle.shift(red.Ht, 1.0, -gam[S // 2])
                xo = oracle.ldiv(red, sh, B[:, offs[S // 2]])
                assert relerr(X[:, offs[S // 2]], xo) <= 1e-12 * n, (m, n)
                checked += 1
    say(f"  worst col rel resid fp64 {worst[np.float64]:.2e} fp32+refine {worst[np.float32]:.2e}; "
        f"{checked} matrices checked against the oracle")
    assert checked >= 10
