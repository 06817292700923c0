"""Python mirror of LinearSolve.jl's `LHLFactorization` algorithm over a
SciMLOperators `WOperator`, driving the same C ABI as the drop-in calls.

The reference documents this surface without shipping it (README.md:57-75,
src/LHLFactorization.jl:39-41):

    W = WOperator{true}(I, 0.01, J, similar(b))
    cache = init(LinearProblem(W, b), LHLFactorization())
    u1 = solve!(cache).u
    update_gamma!(cache, 0.013)       # O(n^2): reuses the reduction of J
    u2 = solve!(cache).u

Here:  W = WOperator(None, 0.01, J);  cache = init(LinearProblem(W, b),
LHLFactorization());  solve_(cache).u;  update_gamma_(cache, 0.013).

The operator is the split an implicit ODE solver builds, in its concrete form
    transform=True  (the default):  W = J - M/gamma   (README.md:57, src:40)
    transform=False:                W = gamma J - M
and the mass matrix M must be a multiple lam*I of the identity (None or "I"
for I itself), so that W = sigma I + tau J with
    transform=True:   (sigma, tau) = (-lam/gamma, 1)   (README.md:57: (-1/(dt*gamma), 1))
    transform=False:  (sigma, tau) = (-lam, gamma).
A complex gamma (a Radau stage's complex shift on a real J) gives a complex
shift, loaded into a complex LHLShift held beside the workspace's real one:
the real and the complex stages share one reduction, as README.md's Radau
note describes.  A mass matrix that is
not a multiple of I raises ArgumentError: W is then not a shift of J and the
reduction does not apply.

Caching follows LinearSolve's: `init` reduces J once (O(n^3)) and loads the
shift; `update_gamma_` only reloads the shift (O(n^2)); assigning `cache.A`
a new operator marks the cache fresh, and the next `solve_` reduces its J;
assigning `cache.b` changes only the right-hand side.  `cache.reductions`
counts the reductions run.

Parity is unpinned: LinearSolve.jl and SciMLOperators.jl are not in the
reference, so the mapping above follows the reference's README text.  The GPU
tests check that W x = b holds and that the adapter's results are bitwise
the drop-in calls' (lhl_shift_ then lhl_ldiv_) for the same (sigma, tau).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import (ArgumentError, LHLShift, LHLWorkspace, _is_torch, lhl, lhl_, lhl_ldiv_, lhl_refine_, lhl_shift_)

__all__ = ["WOperator", "LinearProblem", "LHLFactorization", "LinearCache", "LinearSolution", "init", "solve_",
           "update_gamma_", "shift_of", "mass_scalar"]


def mass_scalar(M):
    """lam for a mass matrix lam*I (None and "I" are I); ArgumentError otherwise."""
    if M is None or (isinstance(M, str) and M == "I"):
        return 1.0
    if np.isscalar(M):
        return M
    A = M.detach().cpu().numpy() if _is_torch(M) else np.asarray(M)
    if A.ndim == 0:
        return A.item()
    if A.ndim != 2 or A.shape[0] != A.shape[1]:
        raise ArgumentError("LHLFactorization: the mass matrix must be square")
    lam = A[0, 0] if A.shape[0] else 1.0
    if not np.array_equal(A, lam * np.eye(A.shape[0], dtype=A.dtype)):
        raise ArgumentError("LHLFactorization: the mass matrix must be a multiple of I "
                            "(W = J - M/gamma is then a shift of J)")
    return lam.item() if hasattr(lam, "item") else lam


class WOperator:
    """The W operator of an implicit ODE step (SciMLOperators.WOperator):
    mass matrix M, step scalar gamma, Jacobian J; `concrete()` is the dense W."""

    def __init__(self, mass_matrix, gamma, J, u=None, transform=True):
        self.mass_matrix = mass_matrix
        self.gamma = gamma
        self.J = J
        self.u = u
        self.transform = bool(transform)

    @property
    def n(self):
        return self.J.shape[0]

    def concrete(self):
        """The dense W on the host (FP64 or complex128), for checks."""
        J = self.J.detach().cpu().numpy() if _is_torch(self.J) else np.asarray(self.J)
        s, t = shift_of(self)
        return s * np.eye(J.shape[0]) + t * J


def shift_of(W: WOperator):
    """(sigma, tau) with W = sigma I + tau J for the operator's concrete form."""
    lam = mass_scalar(W.mass_matrix)
    g = W.gamma
    if g == 0:
        raise ArgumentError("LHLFactorization: gamma must be nonzero")
    if W.transform:
        return -lam / g, 1.0
    return -lam, g


@dataclass
class LinearProblem:
    A: WOperator
    b: object


@dataclass
class LHLFactorization:
    """The algorithm.  dtype: the reduction's element type (None: J's; float32
    gives the FP32 factors of the precision form); refine: lhl_refine! rounds
    per solve against sigma I + tau J with the resident FP64 J (0: none); tol:
    stop refining once every column's relative residual is below it."""
    dtype: object = None
    refine: int = 0
    tol: float = 0.0
    balance: bool = True


@dataclass
class LinearSolution:
    u: object
    retcode: str
    cache: "LinearCache" = field(repr=False)


def _is_complex_scalar(z):
    return isinstance(z, complex) or np.iscomplexobj(z)


class LinearCache:
    """LinearSolve's cache for this algorithm: the operator, the right-hand
    side, the solution vector and the workspace (`cacheval`)."""

    def __init__(self, prob: LinearProblem, alg: LHLFactorization):
        self.alg = alg
        self._A = prob.A
        self._b = prob.b
        self.cacheval: LHLWorkspace | None = None
        self._csh: LHLShift | None = None
        self.isfresh = True
        self.reductions = 0
        self._loaded = {}         # (sigma, tau) loaded per shift state
        self.u = self._new_u()

    def _needs_complex(self):
        b = self._b
        bc = b.is_complex() if _is_torch(b) else np.iscomplexobj(b)
        return bc or _is_complex_scalar(self._A.gamma) or _is_complex_scalar(mass_scalar(self._A.mass_matrix))

    def _new_u(self):
        b = self._b
        cplx = self._needs_complex()
        if _is_torch(b):
            import torch
            dt = b.dtype
            if cplx and not b.is_complex():
                dt = torch.complex128 if dt == torch.float64 else torch.complex64
            return torch.zeros(b.shape, dtype=dt, device=b.device)
        dt = np.asarray(b).dtype
        if cplx and not np.iscomplexobj(b):
            dt = np.complex128 if dt == np.float64 else np.complex64
        return np.zeros(np.shape(b), dtype=dt)

    @property
    def A(self):
        return self._A

    @A.setter
    def A(self, W):
        self._A = W
        self.isfresh = True

    @property
    def b(self):
        return self._b

    @b.setter
    def b(self, b):
        self._b = b
        if tuple(np.shape(b)) != tuple(np.shape(self.u)) or self._needs_complex() != self._u_complex():
            self.u = self._new_u()

    def _u_complex(self):
        u = self.u
        return u.is_complex() if _is_torch(u) else np.iscomplexobj(u)

    def _reduce(self):
        W = self._A
        ws = self.cacheval
        T = np.dtype(self.alg.dtype) if self.alg.dtype is not None else None
        if ws is None or ws.n != W.n or (T is not None and ws.dtype != T):
            self.cacheval = lhl(W.J, balance=self.alg.balance, dtype=T)
            self._csh = None
        else:
            lhl_(ws, W.J, balance=self.alg.balance)
        self.reductions += 1
        self.isfresh = False
        self._loaded = {}
        if tuple(np.shape(self.u)) != tuple(np.shape(self._b)) or self._needs_complex() != self._u_complex():
            self.u = self._new_u()

    def _target(self, s, t):
        """The shift state for (s, t): the workspace's own (real) one, or for a
        complex shift a complex LHLShift on the same reduction."""
        ws = self.cacheval
        if not (_is_complex_scalar(s) or _is_complex_scalar(t)):
            return ws.shift
        if self._csh is None:
            self._csh = LHLShift(ws, np.complex128 if ws.dtype == np.float64 else np.complex64)
        return self._csh

    def _load_shift(self):
        s, t = shift_of(self._A)
        sh = self._target(s, t)
        if self._loaded.get(id(sh)) != (s, t):
            if sh is self.cacheval.shift:
                lhl_shift_(self.cacheval, s, t)
            else:
                lhl_shift_(sh, self.cacheval, s, t)
            self._loaded[id(sh)] = (s, t)
        return sh


def init(prob: LinearProblem, alg: LHLFactorization | None = None) -> LinearCache:
    """init(LinearProblem(W, b), LHLFactorization()): reduce J and load the shift."""
    if not isinstance(prob.A, WOperator):
        raise ArgumentError("LHLFactorization consumes a WOperator (the split J - M/gamma)")
    cache = LinearCache(prob, alg or LHLFactorization())
    cache._reduce()
    cache._load_shift()
    return cache


def solve_(cache: LinearCache) -> LinearSolution:
    """solve!(cache): u = W^-1 b with the held reduction and shift (a fresh
    operator is reduced first).  retcode "Failure" when the shift LU met a zero
    pivot (info > 0; u then holds non-finite values), as for a singular W."""
    if cache.isfresh:
        cache._reduce()
    sh = cache._load_shift()
    ws = cache.cacheval
    u, b = cache.u, cache._b

    def fill(dst):
        if _is_torch(dst):
            import torch
            dst.copy_(b if _is_torch(b) else torch.as_tensor(np.asarray(b)))
        else:
            dst[...] = np.asarray(b)
        return dst

    fill(u)
    own = sh is ws.shift
    if own:
        lhl_ldiv_(u, ws)
    else:
        lhl_ldiv_(u, sh, ws)
    if cache.alg.refine > 0 and sh.info == 0:
        rb = fill(u.clone() if _is_torch(u) else np.array(u))
        if own:
            lhl_refine_(u, None, rb, ws, cache.alg.refine, tol=cache.alg.tol)
        else:
            lhl_refine_(u, None, rb, sh, ws, cache.alg.refine, tol=cache.alg.tol)
    return LinearSolution(u, "Success" if sh.info == 0 else "Failure", cache)


def update_gamma_(cache: LinearCache, gamma) -> LinearCache:
    """update_gamma!(cache, gamma): move the shift, reusing the reduction of J
    (one O(n^2) shift LU, no O(n^3) work)."""
    cache._A.gamma = gamma
    if not cache.isfresh:
        if cache._needs_complex() != cache._u_complex():
            cache.u = cache._new_u()
        cache._load_shift()
    return cache
