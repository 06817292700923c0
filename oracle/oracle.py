"""ctypes front end of the CPU oracle (oracle/liblhl_oracle.so) -- test infrastructure only.

The oracle is a plain-C restatement of SciML/LHL.jl's algorithm
(src/LHLFactorization.jl; each C function cites the lines it follows).  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, as the checker or as the timed CPU baseline.  Indices are 0-based here
(ipiv[c] is the 0-based pivot row of step c); the reference's are 1-based.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblhl_oracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
_sp = np.ctypeslib.ndpointer(dtype=np.float32, flags="F_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_bp = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def build():
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        for sfx, fp in (("_d", _dp), ("_s", _sp)):
            R = ctypes.c_double if sfx == "_d" else ctypes.c_float
            c_int = ctypes.c_int
            f = getattr(L, "lhlo_reduce" + sfx)
            f.argtypes = [c_int, fp, c_int, c_int, c_int, c_int, fp, c_int, _ip, _ip, _ip, fp, fp, fp]
            f.restype = None
            f = getattr(L, "lhlo_balance" + sfx)
            f.argtypes = [c_int, fp, c_int, fp, fp]
            f.restype = None
            f = getattr(L, "lhlo_reduce_unblocked" + sfx)
            f.argtypes = [c_int, fp, c_int, _ip]
            f.restype = None
            f = getattr(L, "lhlo_reduce_blocked" + sfx)
            f.argtypes = [c_int, fp, c_int, _ip, c_int]
            f.restype = None
            f = getattr(L, "lhlo_shift" + sfx)
            f.argtypes = [c_int, fp, R, R, fp, _bp, fp]
            f.restype = c_int
            f = getattr(L, "lhlo_shift_fused" + sfx)
            f.argtypes = [c_int, c_int, fp, R, R, fp, _bp, fp]
            f.restype = c_int
            f = getattr(L, "lhlo_shift_c" + sfx)
            f.argtypes = [c_int, fp, R, R, R, R, fp, fp, _bp, fp, fp]
            f.restype = c_int
            f = getattr(L, "lhlo_hess_solve" + sfx)
            f.argtypes = [c_int, fp, _bp, fp, fp]
            f.restype = None
            f = getattr(L, "lhlo_ldiv" + sfx)
            f.argtypes = [c_int, fp, c_int, _ip, _ip, fp, fp, fp, _bp, fp, fp]
            f.restype = None
            f = getattr(L, "lhlo_ldiv_c" + sfx)
            f.argtypes = [c_int, fp, c_int, _ip, _ip, fp, fp, fp, fp, _bp, fp, fp, fp, fp]
            f.restype = None
            f = getattr(L, "lhlo_applyZ" + sfx)
            f.argtypes = [c_int, fp, c_int, _ip, fp, fp]
            f.restype = None
            f = getattr(L, "lhlo_applyZinv" + sfx)
            f.argtypes = [c_int, fp, c_int, _ip, fp, fp]
            f.restype = None
            f = getattr(L, "lhlo_refine" + sfx)
            f.argtypes = [c_int, fp, c_int, fp, fp, c_int, _ip, _ip, fp, fp, fp, _bp, fp, fp, c_int]
            f.restype = None
        for sfx, fp in (("_d", _dp), ("_s", _sp)):
            R = ctypes.c_double if sfx == "_d" else ctypes.c_float
            c_int = ctypes.c_int
            f = getattr(L, "lhlo_zabs" + sfx)
            f.argtypes = [R, R]
            f.restype = R
            f = getattr(L, "lhlo_zbalance" + sfx)
            f.argtypes = [c_int, fp, fp, c_int, fp, fp]
            f.restype = None
            f = getattr(L, "lhlo_zreduce_unblocked" + sfx)
            f.argtypes = [c_int, fp, fp, c_int, _ip]
            f.restype = None
            f = getattr(L, "lhlo_zreduce" + sfx)
            f.argtypes = [c_int, fp, fp, c_int, c_int, fp, fp, c_int, _ip, _ip, _ip, fp, fp, fp, fp]
            f.restype = None
            f = getattr(L, "lhlo_zshift" + sfx)
            f.argtypes = [c_int, fp, fp, R, R, R, R, fp, fp, _bp, fp, fp]
            f.restype = c_int
            f = getattr(L, "lhlo_zldiv" + sfx)
            f.argtypes = [c_int, fp, fp, c_int, _ip, _ip, fp, fp, fp, fp, _bp, fp, fp, fp, fp]
            f.restype = None
            for nm in ("lhlo_zapplyZ", "lhlo_zapplyZinv"):
                f = getattr(L, nm + sfx)
                f.argtypes = [c_int, fp, fp, c_int, _ip, fp, fp, fp]
                f.restype = None
        _lib = L
    return _lib


def _sfx(dtype):
    dtype = np.dtype(dtype)
    if dtype == np.float64:
        return "_d"
    if dtype == np.float32:
        return "_s"
    raise TypeError(f"oracle supports float64/float32, got {dtype}")


def _f(a, dtype):
    return np.asfortranarray(a, dtype=dtype)


def _vec(a, dtype):
    return np.asfortranarray(np.asarray(a, dtype=dtype).reshape(-1))


@dataclass
class Reduction:
    """The oracle's reduction of J: factors A (H in triu(A,-1), step-c multipliers
    in A[c+2:, c]), 0-based ipiv/perm/iperm, balancing scale/iscale, Ht = H^T."""

    A: np.ndarray
    ipiv: np.ndarray
    perm: np.ndarray
    iperm: np.ndarray
    scale: np.ndarray
    iscale: np.ndarray
    Ht: np.ndarray

    @property
    def n(self):
        return self.A.shape[0]

    @property
    def H(self):
        return np.triu(self.A, -1)

    @property
    def L(self):
        """Unit lower triangular L with L[i, c+1] = A[i, c] for i >= c+2."""
        n = self.n
        L = np.eye(n, dtype=self.A.dtype)
        for c in range(max(n - 2, 0)):
            L[c + 2:, c + 1] = self.A[c + 2:, c]
        return L

    def Z(self):
        """Z = D P L as a dense matrix (columns of Z = applyZ on unit vectors)."""
        n = self.n
        Z = np.eye(n, dtype=self.A.dtype, order="F")
        for j in range(n):
            col = np.asfortranarray(Z[:, j].copy())
            applyZ(self, col)
            Z[:, j] = col
        return Z


def reduce(J, balance=True, blocked=-1, nb=16, dtype=None):
    """lhl(J; balance) on the CPU oracle; blocked=-1 picks the reference's
    size switch (src:398, 703-704), 0 forces unblocked, 1 blocked."""
    J = np.asarray(J)
    dtype = np.dtype(dtype or J.dtype)
    n = J.shape[0]
    Jf = _f(J, dtype)
    A = np.zeros((n, n), dtype=dtype, order="F")
    ipiv = np.zeros(max(n - 2, 1), dtype=np.int32)
    perm = np.zeros(max(n, 1), dtype=np.int32)
    iperm = np.zeros(max(n, 1), dtype=np.int32)
    scale = np.zeros(max(n, 1), dtype=dtype, order="F")
    iscale = np.zeros(max(n, 1), dtype=dtype, order="F")
    Ht = np.zeros((max(n, 1), max(n, 1)), dtype=dtype, order="F")
    if n > 0:
        getattr(lib(), "lhlo_reduce" + _sfx(dtype))(
            n, Jf, max(n, 1), int(bool(balance)), int(blocked), int(nb), A, max(n, 1),
            ipiv, perm, iperm, scale, iscale, Ht)
    return Reduction(A, ipiv[: max(n - 2, 0)].copy(), perm[:n].copy(), iperm[:n].copy(),
                     scale[:n].copy(), iscale[:n].copy(), Ht[:n, :n].copy(order="F"))


def balance(J, dtype=None):
    J = np.asarray(J)
    dtype = np.dtype(dtype or J.dtype)
    A = _f(J, dtype).copy(order="F")
    n = A.shape[0]
    d = np.zeros(max(n, 1), dtype=dtype)
    isc = np.zeros(max(n, 1), dtype=dtype)
    if n:
        getattr(lib(), "lhlo_balance" + _sfx(dtype))(n, A, n, d, isc)
    return A, d[:n], isc[:n]


@dataclass
class Shift:
    """The LU of sigma*I + tau*H in transposed form (Gt[j, k] = U[k, j] for
    j >= k, multiplier in Gt[k, k+1]), swap flags, reciprocal pivots, info."""

    Gt: np.ndarray
    swap: np.ndarray
    rdiag: np.ndarray
    info: int
    Gt_im: np.ndarray | None = None
    rdiag_im: np.ndarray | None = None

    @property
    def is_complex(self):
        return self.Gt_im is not None

    def U(self):
        G = self.Gt if not self.is_complex else self.Gt + 1j * self.Gt_im
        return np.triu(G.T)

    def mult(self):
        G = self.Gt if not self.is_complex else self.Gt + 1j * self.Gt_im
        n = G.shape[0]
        return np.array([G[k, k + 1] for k in range(n - 1)])


def shift(Ht, sigma, tau, R=None):
    """lhl_shift! (real) on the oracle; R selects the fused R-step variant."""
    dtype = Ht.dtype
    n = Ht.shape[0]
    Htf = _f(Ht, dtype)
    Gt = np.zeros((max(n, 1), max(n, 1)), dtype=dtype, order="F")
    sw = np.zeros(max(n, 1), dtype=np.uint8)
    rd = np.zeros(max(n, 1), dtype=dtype, order="F")
    info = 0
    if n:
        if R is None:
            info = getattr(lib(), "lhlo_shift" + _sfx(dtype))(n, Htf, sigma, tau, Gt, sw, rd)
        else:
            info = getattr(lib(), "lhlo_shift_fused" + _sfx(dtype))(n, int(R), Htf, sigma, tau, Gt, sw, rd)
    return Shift(Gt[:n, :n].copy(order="F"), sw[:n].copy(), rd[:n].copy(), int(info))


def shift_c(Ht, sigma, tau):
    """lhl_shift! with a complex shift on a real reduction (planar storage)."""
    dtype = Ht.dtype
    n = Ht.shape[0]
    sigma = complex(sigma)
    tau = complex(tau)
    Htf = _f(Ht, dtype)
    Gr = np.zeros((max(n, 1), max(n, 1)), dtype=dtype, order="F")
    Gi = np.zeros_like(Gr)
    sw = np.zeros(max(n, 1), dtype=np.uint8)
    rr = np.zeros(max(n, 1), dtype=dtype, order="F")
    ri = np.zeros_like(rr)
    info = 0
    if n:
        info = getattr(lib(), "lhlo_shift_c" + _sfx(dtype))(
            n, Htf, sigma.real, sigma.imag, tau.real, tau.imag, Gr, Gi, sw, rr, ri)
    return Shift(Gr[:n, :n].copy(order="F"), sw[:n].copy(), rr[:n].copy(), int(info),
                 Gi[:n, :n].copy(order="F"), ri[:n].copy())


def ldiv(red: Reduction, sh: Shift, x):
    """lhl_ldiv!: x <- (sigma I + tau J)^-1 x, returns a new array."""
    n = red.n
    dtype = red.A.dtype
    if sh.is_complex or np.iscomplexobj(x):
        x = np.asarray(x, dtype=np.result_type(dtype, np.complex64))
        xr = _vec(x.real, dtype).copy()
        xi = _vec(x.imag, dtype).copy()
        if sh.is_complex:
            Gi, ri = sh.Gt_im, sh.rdiag_im
        else:
            Gi, ri = np.zeros_like(sh.Gt), np.zeros_like(sh.rdiag)
        if n:
            getattr(lib(), "lhlo_ldiv_c" + _sfx(dtype))(
                n, red.A, n, red.perm, red.iperm, red.scale, red.iscale,
                _f(sh.Gt, dtype), _f(Gi, dtype), sh.swap, _vec(sh.rdiag, dtype), _vec(ri, dtype), xr, xi)
        return xr + 1j * xi
    xx = _vec(x, dtype).copy()
    if n:
        getattr(lib(), "lhlo_ldiv" + _sfx(dtype))(
            n, red.A, n, red.perm, red.iperm, red.scale, red.iscale, sh.Gt, sh.swap, sh.rdiag, xx)
    return xx


def applyZ(red: Reduction, x):
    n = red.n
    if n:
        getattr(lib(), "lhlo_applyZ" + _sfx(red.A.dtype))(n, red.A, n, red.iperm, red.scale, x)
    return x


def applyZinv(red: Reduction, x):
    n = red.n
    if n:
        getattr(lib(), "lhlo_applyZinv" + _sfx(red.A.dtype))(n, red.A, n, red.perm, red.iscale, x)
    return x


def refine(red: Reduction, sh: Shift, W, b, x, steps=1):
    """lhl_refine! with a dense real W; complex systems go through numpy here."""
    dtype = red.A.dtype
    if sh.is_complex or np.iscomplexobj(x) or np.iscomplexobj(W) or np.iscomplexobj(b):
        x = np.array(x, dtype=np.result_type(dtype, np.complex64))
        for _ in range(steps):
            r = np.asarray(b) - np.asarray(W) @ x
            x = x + ldiv(red, sh, r)
        return x
    n = red.n
    xx = _vec(x, dtype).copy()
    if n:
        getattr(lib(), "lhlo_refine" + _sfx(dtype))(
            n, _f(W, dtype), n, _vec(b, dtype), red.A, n, red.perm, red.iperm, red.scale, red.iscale,
            sh.Gt, sh.swap, sh.rdiag, xx, int(steps))
    return xx


def bwd(W, x, b):
    """Backward error of the reference tests (test/core_tests.jl:3)."""
    W = np.asarray(W)
    r = np.asarray(b) - W @ np.asarray(x)
    return np.linalg.norm(r, np.inf) / (np.linalg.norm(W, np.inf) * np.linalg.norm(x, np.inf)
                                        + np.linalg.norm(b, np.inf))


# ---- complex J (LHLWorkspace{Complex{T}}): planar arrays for the C side ----

def _real_of(dtype):
    dtype = np.dtype(dtype)
    if dtype == np.complex128:
        return np.dtype(np.float64)
    if dtype == np.complex64:
        return np.dtype(np.float32)
    raise TypeError(f"complex oracle supports complex128/complex64, got {dtype}")


def zabs(z, dtype=np.complex128):
    """The modulus used by the complex reduction's pivoting and balancing."""
    rt = _real_of(dtype)
    return getattr(lib(), "lhlo_zabs" + _sfx(rt))(float(np.real(z)), float(np.imag(z)))


@dataclass
class ZReduction:
    """The oracle's reduction of a complex J (as Reduction, complex A and Ht;
    scale and iscale real)."""

    A: np.ndarray
    ipiv: np.ndarray
    perm: np.ndarray
    iperm: np.ndarray
    scale: np.ndarray
    iscale: np.ndarray
    Ht: np.ndarray

    @property
    def n(self):
        return self.A.shape[0]

    @property
    def H(self):
        return np.triu(self.A, -1)

    @property
    def L(self):
        n = self.n
        L = np.eye(n, dtype=self.A.dtype)
        for c in range(max(n - 2, 0)):
            L[c + 2:, c + 1] = self.A[c + 2:, c]
        return L

    def Z(self):
        n = self.n
        Z = np.eye(n, dtype=self.A.dtype, order="F")
        for j in range(n):
            Z[:, j] = zapplyZ(self, Z[:, j].copy())
        return Z


def zreduce(J, balance=True, dtype=None):
    """lhl(J; balance) of a complex J on the oracle (unblocked step loop)."""
    J = np.asarray(J)
    dtype = np.dtype(dtype or (J.dtype if np.iscomplexobj(J) else np.complex128))
    rt = _real_of(dtype)
    n = J.shape[0]
    m = max(n, 1)
    Jr, Ji = _f(J.real, rt), _f(J.imag, rt)
    Ar = np.zeros((m, m), dtype=rt, order="F")
    Ai = np.zeros_like(Ar)
    Hr = np.zeros_like(Ar)
    Hi = np.zeros_like(Ar)
    ipiv = np.zeros(max(n - 2, 1), dtype=np.int32)
    perm = np.zeros(m, dtype=np.int32)
    iperm = np.zeros(m, dtype=np.int32)
    scale = np.zeros(m, dtype=rt)
    iscale = np.zeros(m, dtype=rt)
    if n > 0:
        getattr(lib(), "lhlo_zreduce" + _sfx(rt))(
            n, Jr, Ji, m, int(bool(balance)), Ar, Ai, m, ipiv, perm, iperm, scale, iscale, Hr, Hi)
    A = (Ar[:n, :n] + 1j * Ai[:n, :n]).astype(dtype)
    Ht = (Hr[:n, :n] + 1j * Hi[:n, :n]).astype(dtype)
    return ZReduction(np.asfortranarray(A), ipiv[: max(n - 2, 0)].copy(), perm[:n].copy(), iperm[:n].copy(),
                      scale[:n].copy(), iscale[:n].copy(), np.asfortranarray(Ht))


def zbalance(J, dtype=np.complex128):
    rt = _real_of(dtype)
    J = np.asarray(J)
    n = J.shape[0]
    Ar, Ai = _f(J.real, rt).copy(order="F"), _f(J.imag, rt).copy(order="F")
    d = np.zeros(max(n, 1), dtype=rt)
    isc = np.zeros(max(n, 1), dtype=rt)
    if n:
        getattr(lib(), "lhlo_zbalance" + _sfx(rt))(n, Ar, Ai, n, d, isc)
    return (Ar + 1j * Ai).astype(dtype), d[:n], isc[:n]


def zshift(Ht, sigma, tau):
    """lhl_shift! of a complex H (Ht complex, n x n)."""
    dtype = np.dtype(Ht.dtype)
    rt = _real_of(dtype)
    n = Ht.shape[0]
    m = max(n, 1)
    sigma, tau = complex(sigma), complex(tau)
    Gr = np.zeros((m, m), dtype=rt, order="F")
    Gi = np.zeros_like(Gr)
    sw = np.zeros(m, dtype=np.uint8)
    rr = np.zeros(m, dtype=rt)
    ri = np.zeros(m, dtype=rt)
    info = 0
    if n:
        info = getattr(lib(), "lhlo_zshift" + _sfx(rt))(
            n, _f(Ht.real, rt), _f(Ht.imag, rt), sigma.real, sigma.imag, tau.real, tau.imag, Gr, Gi, sw, rr, ri)
    return Shift(Gr[:n, :n].copy(order="F"), sw[:n].copy(), rr[:n].copy(), int(info),
                 Gi[:n, :n].copy(order="F"), ri[:n].copy())


def zldiv(red: ZReduction, sh: Shift, x):
    """lhl_ldiv! on a complex reduction: returns (sigma I + tau J)^-1 x."""
    n = red.n
    rt = _real_of(red.A.dtype)
    x = np.asarray(x)
    xr, xi = _vec(x.real, rt).copy(), _vec(x.imag, rt).copy()
    if n:
        getattr(lib(), "lhlo_zldiv" + _sfx(rt))(
            n, _f(red.A.real, rt), _f(red.A.imag, rt), n, red.perm, red.iperm, _vec(red.scale, rt),
            _vec(red.iscale, rt), _f(sh.Gt, rt), _f(sh.Gt_im, rt), sh.swap, _vec(sh.rdiag, rt),
            _vec(sh.rdiag_im, rt), xr, xi)
    return xr + 1j * xi


def _zapply(name, red, x, idx, sc):
    n = red.n
    rt = _real_of(red.A.dtype)
    x = np.asarray(x)
    xr, xi = _vec(x.real, rt).copy(), _vec(x.imag, rt).copy()
    if n:
        getattr(lib(), name + _sfx(rt))(n, _f(red.A.real, rt), _f(red.A.imag, rt), n, idx, _vec(sc, rt), xr, xi)
    return xr + 1j * xi


def zapplyZ(red: ZReduction, x):
    return _zapply("lhlo_zapplyZ", red, x, red.iperm, red.scale)


def zapplyZinv(red: ZReduction, x):
    return _zapply("lhlo_zapplyZinv", red, x, red.perm, red.iscale)
