"""The solves' dense products on the library's MFMA kernels (gemm_impl = 1)
against rocBLAS (gemm_impl = 0, the default): the L sweeps (an explicit
triangular inverse, and I + L' read from the factors' strict lower triangle)
and the refinement residual's J x.  Batched solves of real and complex shifts
in both precisions agree between the two within rounding, and both meet the
residual bar; sizes straddle the 128 x 64 tiles (n = 129, 700, 1100)."""
import numpy as np
import pytest

from inputs import randn, stiff

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_mfma_products_match_rocblas(lhl_hip, dtype):
    for n, S, R in ((129, 3, 5), (700, 5, 3), (1100, 4, 16)):
        J = stiff(90 + n, n)
        rng = np.random.default_rng(n)
        gam = np.logspace(-4, -1, S)
        offs = np.arange(0, S * R + 1, R, dtype=np.int64)
        B = np.asfortranarray(rng.standard_normal((n, S * R)))
        Bc = np.asfortranarray(B + 1j * rng.standard_normal((n, S * R)))
        om = np.logspace(-1, 1, S)
        ws = lhl_hip.lhl(J, dtype=dtype)
        out = {}
        for impl in (0, 1):
            say(f"case: gemm_impl={impl} n={n} {np.dtype(dtype).name}")
            lhl_hip.tune_set(ws, "gemm_impl", impl)
            Xw, _, _ = lhl_hip.shift_solve_batched(ws, np.ones(S), -gam, B, offs)
            Xr, _, _ = lhl_hip.shift_solve_batched(ws, np.ones(S), -gam, B, offs, precision="refine64", tol=1e-12,
                                                   max_iter=30)
            Xc, _, _ = lhl_hip.shift_solve_batched(ws, 1j * om, -np.ones(S), Bc, offs, precision="refine64",
                                                   tol=1e-12, max_iter=30)
            out[impl] = (Xw, Xr, Xc)
            for s in range(S):
                c = slice(offs[s], offs[s + 1])
                W = np.eye(n) - gam[s] * J
                r = np.abs(B[:, c] - W @ Xr[:, c]).max() / np.abs(B[:, c]).max()
                assert r <= 1e-12, (impl, s, r)
                Wc = 1j * om[s] * np.eye(n) - J
                r = np.abs(Bc[:, c] - Wc @ Xc[:, c]).max() / np.abs(Bc[:, c]).max()
                assert r <= 1e-12, (impl, s, r)
        tol = 1e-10 if dtype == np.float64 else 1e-3
        e = np.linalg.norm(out[0][0] - out[1][0]) / np.linalg.norm(out[0][0])
        say(f"  unrefined solves differ by {e:.2e}")
        assert e <= tol
        for k in (1, 2):
            e = np.linalg.norm(out[0][k] - out[1][k]) / np.linalg.norm(out[0][k])
            assert e <= 1e-11, (k, e)
