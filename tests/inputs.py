"""Seeded test inputs shared by the CPU and GPU tests (the reference's own
inputs come from Julia's MersenneTwister, which is not available here, so the
matrix families are rebuilt with numpy's PCG64 and fixed seeds)."""
import numpy as np


def randn(seed, n, m=None, dtype=np.float64):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.standard_normal((n, n if m is None else m)).astype(dtype))


def tie_matrix(seed, n, lo=-2, hi=2, dtype=np.float64):
    """Integer matrix with exact ties and zero pivots (test/core_tests.jl:191)."""
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.integers(lo, hi + 1, size=(n, n)).astype(dtype))


def stiff(seed, n, dtype=np.float64):
    """J = -diag(10^u) + G/sqrt(n), u linearly spaced in [0, 6], G ~ N(0,1)."""
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((n, n))
    u = np.linspace(0.0, 6.0, n)
    J = G / np.sqrt(n) - np.diag(10.0 ** u)
    return np.asfortranarray(J.astype(dtype))


def wilkinson(n):
    """Wilkinson's growth matrix (test/core_tests.jl:343-346)."""
    A = np.where(np.tril(np.ones((n, n)), -1) > 0, -1.0, 0.0)
    np.fill_diagonal(A, 1.0)
    A[:, -1] = 1.0
    return np.asfortranarray(A)


def graded(seed, n):
    """Diagonal(10^range(-8,8)) * G * Diagonal(10^range(8,-8)) (core_tests.jl:347-348)."""
    d = 10.0 ** np.linspace(-8, 8, n)
    return np.asfortranarray(np.diag(d) @ randn(seed, n) @ np.diag(d[::-1]))


def near_nilpotent(seed, n):
    """triu(G, 1) with a 1e-8 corner (core_tests.jl:369-370)."""
    J = np.triu(randn(seed, n), 1)
    J[n - 1, 0] = 1e-8
    return np.asfortranarray(J)


PRECOMPILE_J = np.asfortranarray(np.array([[1.0, 0.2, 0.0, 0.0], [0.0, 2.0, 0.3, 0.0],
                                           [0.0, 0.0, 3.0, 0.4], [0.1, 0.0, 0.0, 4.0]]))


def hilbert_plus_I(n, dtype=np.float64):
    """J = [1/(i+j) + (i == j)], 1-based i, j (src:3205)."""
    i = np.arange(1, n + 1)
    return np.asfortranarray((1.0 / (i[:, None] + i[None, :]) + np.eye(n)).astype(dtype))


def bwd(W, x, b):
    """test/core_tests.jl:3."""
    W = np.asarray(W)
    r = np.asarray(b) - W @ np.asarray(x)
    return np.linalg.norm(r, np.inf) / (np.linalg.norm(W, np.inf) * np.linalg.norm(x, np.inf)
                                        + np.linalg.norm(b, np.inf))
