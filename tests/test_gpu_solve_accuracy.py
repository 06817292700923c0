"""The reference's solve-accuracy testset on the GPU path, with its thresholds
(test/core_tests.jl:37-57, "explicit-vector solve kernels agree with the
generic ones"): for T in (Float64, Float32) and n in (3, 7, 33, 130, 331, 500,
725, 1030), J = randn(T, n, n), one shift (1, -0.05), W = I - 0.05 J:

  * the solve agrees with a dense LU solve of W to rtol 1e-9 (Float64) or 1e-3
    (Float32, n <= 500);
  * the normwise backward error |b - W x| / (|W| |x| + |b|) is below 50 eps(T)
    for n <= 500; beyond, the other right-hand-side paths stay within twice the
    real vector's backward error (plus eps(T)), as the reference requires of
    its three paths;
  * a vector of the wrong length is a DimensionMismatch for lhl_ldiv! and
    applyZ!.

The reference's three paths (a Vector{T}, a view and a complex vector) map to a
real vector of the workspace's type, an FP64 vector on an FP32 workspace
(FP32 factors, the right-hand side and solution held in FP64) and a complex
vector.  The matrices come from this suite's seeded generator, not Julia's
MersenneTwister.  Each case prints its errors.
"""
import numpy as np
import pytest

from inputs import bwd, randn

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [3, 7, 33, 130, 331, 500, 725, 1030])
def test_reference_solve_accuracy(lhl_hip, n, dtype):
    T = np.dtype(dtype)
    eps = float(np.finfo(T).eps)
    J = randn(n, n, dtype=T)
    b = randn(n + 7, n, 1, dtype=T)[:, 0].copy()
    ws = lhl_hip.lhl(J)
    assert ws.factors.dtype == T
    lhl_hip.lhl_shift_(ws, 1.0, -0.05)
    W = np.eye(n) - np.float64(T.type(0.05)) * J.astype(np.float64)
    W = W.astype(T).astype(np.float64)   # W as the reference forms it, in T
    b64 = b.astype(np.float64)
    x = lhl_hip.lhl_ldiv_(b.copy(), ws)
    y = lhl_hip.lhl_ldiv_(b64.copy(), ws)
    z = lhl_hip.lhl_ldiv_(b.astype(np.complex128), ws)
    assert np.abs(z.imag).max() == 0.0
    ref = np.linalg.solve(W, b64)
    paths = {"real": x.astype(np.float64), "fp64 rhs": np.asarray(y, dtype=np.float64), "complex": z.real}
    bx = bwd(W, paths["real"], b64)
    for name, v in paths.items():
        e = np.linalg.norm(v - ref) / np.linalg.norm(ref)
        bw = bwd(W, v, b64)
        say(f"case: {T.name} n={n} {name}: relerr {e:.2e}, bwd {bw:.2e} = {bw / eps:.1f} eps")
        if T == np.float64 or n <= 500:
            assert e <= (1e-3 if T == np.float32 else 1e-9)
        assert bw < (50 * eps if n <= 500 else 2 * bx + eps)
    with pytest.raises(lhl_hip.DimensionMismatch):
        lhl_hip.lhl_ldiv_(np.zeros(n + 1, dtype=T), ws)
    with pytest.raises(lhl_hip.DimensionMismatch):
        lhl_hip.applyZ_(np.zeros(n - 1, dtype=T), ws)
