"""Python host mirror of SciML/LHL.jl's public interface over liblhl_hip.so.

The names follow the reference (src/LHLFactorization.jl:55-56) with Julia's
`!` spelled as a trailing underscore:

    lhl(J; balance, shift)            -> lhl(J, balance=True, shift=None)          src:3179
    lhl!(ws, J; balance)              -> lhl_(ws, J, balance=True)                  src:3192
    lhl_reduce!(ws, J, balance)       -> lhl_reduce_(ws, J, balance)                src:355
    lhl_shift!(ws, s, t)              -> lhl_shift_(ws, s, t)                       src:2456
    lhl_shift!(sh, ws, s, t)          -> lhl_shift_(sh, ws, s, t)                   src:2461
    lhl_ldiv!(x, ws) / (x, sh, ws)    -> lhl_ldiv_(x, ws) / lhl_ldiv_(x, sh, ws)   src:3052-3123
    lhl_refine!(x, A, b, [sh,] ws, n) -> lhl_refine_(x, A, b, [sh,] ws, steps)      src:3144-3157
    applyZ!(x, ws), applyZinv!(x, ws) -> applyZ_(x, ws), applyZinv_(x, ws)         src:2069-2126
    LHLWorkspace{T}(n; shift)         -> LHLWorkspace(n, dtype, shift=None)         src:164
    LHLShift{TG}(n | ws)              -> LHLShift(n | ws, dtype)                    src:107, 180

Mutating calls return their first argument (the reference's `===` contract,
test/Interface/public_interface.jl).  Errors are ArgumentError /
DimensionMismatch as in the reference; a zero shift pivot is reported in
`.info`, never raised.  Index-valued properties (ipiv, perm, iperm) are 1-based
like the reference's.

Vectors may be numpy arrays (host memory, copied in and out) or torch tensors
on the GPU (used in place).  Every computation runs in the HIP library; there
is no CPU path: importing this module on a machine without the built library
raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

__all__ = [
    "ArgumentError", "DimensionMismatch", "LHLError", "LHLWorkspace", "LHLShift", "lhl", "lhl_",
    "lhl_reduce_", "lhl_shift_", "lhl_ldiv_", "lhl_refine_", "applyZ_", "applyZinv_",
    "shift_solve_batched", "library_path", "load_library", "profile_enable", "profile_reset", "profile_read",
    "profile_read_work", "sync", "LHLNotConverged", "tune_set", "tune_get", "TUNING_KNOBS", "BatchPlan",
    "lhl_group", "lhl_group_",
]

TUNING_KNOBS = ("reduce_blocked_min", "reduce_nb", "reduce_rmin", "reduce_switch", "lsweep_depth",
                "shift_wide_min", "shift_pool_bytes", "h2d_chunk_bytes", "gemm_impl", "group_bytes",
                "shift_split", "utrsm_split")

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "liblhl_hip.so")

LHL_OK, LHL_ERR_ARG, LHL_ERR_DIM, LHL_ERR_HIP, LHL_ERR_STATE, LHL_ERR_NOMEM, LHL_NOT_CONVERGED = range(7)
LHL_F64, LHL_F32, LHL_C64, LHL_C32 = 0, 1, 2, 3
LHL_PREC_WORKING, LHL_PREC_REFINE64 = 0, 1


class LHLError(RuntimeError):
    """A HIP or state error from the library (the reference's ErrorException)."""


class LHLNotConverged(LHLError):
    """A tolerance-driven refinement left a column above tol after the allowed
    rounds.  The outputs are written; `result` holds what the call returns."""

    def __init__(self, msg, result=None):
        super().__init__(msg)
        self.result = result


class ArgumentError(ValueError):
    """Julia's ArgumentError."""


class DimensionMismatch(ValueError):
    """Julia's DimensionMismatch."""


_lib = None


def library_path() -> str:
    return _LIB_PATH


def load_library():
    """Load liblhl_hip.so (built by lhl.jl_amd/build_lib.py).  Fails loudly."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise LHLError(f"{_LIB_PATH} is missing: build it with lhl.jl_amd/build_lib.py")
    # PyTorch ships its own copy of the HIP runtime (same soname).  GPU tensors
    # handed to this library must live in the runtime the library uses, and a
    # second runtime in one process cannot open the GPU: load torch's first
    # when it is installed, so the library binds to it.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(_LIB_PATH)
    vp, i64, ci, cd = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
    sig = {
        "lhl_version": ([], ci),
        "lhl_create": ([ctypes.POINTER(vp), i64, ci, ci], ci),
        "lhl_destroy": ([vp], ci),
        "lhl_last_error": ([vp], ctypes.c_char_p),
        "lhl_reduce": ([vp, vp, i64, i64, ci, ci, ci], ci),
        "lhl_shift_create": ([vp, ctypes.POINTER(vp), ci], ci),
        "lhl_shift_destroy": ([vp], ci),
        "lhl_shift": ([vp, vp, cd, cd, cd, cd, ctypes.POINTER(i64)], ci),
        "lhl_ldiv": ([vp, vp, vp, i64, i64, i64, ci, ci, ci], ci),
        "lhl_refine": ([vp, vp, vp, vp, i64, vp, i64, i64, i64, ci, ci, ci, ci, cd, ctypes.POINTER(ci)], ci),
        "lhl_shift_solve_batched": ([vp, i64, vp, vp, ci, vp, vp, i64, vp, i64, ci, ci, cd, ci, vp, vp], ci),
        "lhl_applyZ": ([vp, vp, i64, i64, i64, ci, ci, ci], ci),
        "lhl_applyZinv": ([vp, vp, i64, i64, i64, ci, ci, ci], ci),
        "lhl_get_info": ([vp, ctypes.POINTER(i64), ctypes.POINTER(ci), ctypes.POINTER(ci), ctypes.POINTER(ci)], ci),
        "lhl_get_ipiv": ([vp, vp], ci),
        "lhl_get_perm": ([vp, vp, vp], ci),
        "lhl_get_scale": ([vp, vp, vp], ci),
        "lhl_get_factors": ([vp, vp], ci),
        "lhl_get_ht": ([vp, vp], ci),
        "lhl_get_shift_state": ([vp, vp, vp, vp, vp, vp, vp, vp, ctypes.POINTER(i64)], ci),
        "lhl_get_shift_info": ([vp, vp, ctypes.POINTER(i64), vp, vp, ctypes.POINTER(ci)], ci),
        "lhl_tune_set": ([vp, ctypes.c_char_p, i64], ci),
        "lhl_tune_get": ([vp, ctypes.c_char_p, ctypes.POINTER(i64)], ci),
        "lhl_applyZ_noalloc": ([vp, ci, vp, i64, i64, i64, ci, ci], ci),
        "lhl_reduce_group": ([i64, vp, vp, vp, vp, ci, ci, ci], ci),
        "lhl_plan_create": ([vp, ctypes.POINTER(vp), i64, ci, vp, vp, i64, vp, i64, ci, cd, ci], ci),
        "lhl_plan_run": ([vp, vp, vp, vp, vp], ci),
        "lhl_plan_stats": ([vp, ctypes.POINTER(ci), ctypes.POINTER(i64), ctypes.POINTER(i64), ctypes.POINTER(i64)], ci),
        "lhl_plan_destroy": ([vp], ci),
        "lhl_sync": ([vp], ci),
        "lhl_get_stream": ([vp, ctypes.POINTER(vp)], ci),
        "lhl_profile_enable": ([vp, ci], ci),
        "lhl_profile_reset": ([vp], ci),
        "lhl_profile_read": ([vp, ctypes.c_char_p, ctypes.POINTER(i64), ctypes.POINTER(cd)], ci),
        "lhl_profile_read_work": ([vp, ctypes.c_char_p, ctypes.POINTER(i64), ctypes.POINTER(cd),
                                   ctypes.POINTER(i64), ctypes.POINTER(i64)], ci),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _check(h, code, result=None):
    if code == LHL_OK:
        return
    msg = _lib.lhl_last_error(h)
    msg = msg.decode() if msg else ""
    if code == LHL_NOT_CONVERGED:
        raise LHLNotConverged(msg, result)
    if code == LHL_ERR_ARG:
        raise ArgumentError(msg)
    if code == LHL_ERR_DIM:
        raise DimensionMismatch(msg)
    raise LHLError(f"liblhl_hip error {code}: {msg}")


def _real_dtype(dt):
    dt = np.dtype(dt)
    if dt in (np.float64, np.complex128):
        return np.dtype(np.float64)
    if dt in (np.float32, np.complex64):
        return np.dtype(np.float32)
    raise ArgumentError(f"element type {dt} is not supported (Float64/Float32 and their complex forms)")


def _code(dt):
    return LHL_F64 if _real_dtype(dt) == np.float64 else LHL_F32


def _full_code(dt):
    """The element-type code with complexity (LHL_C64 / LHL_C32 for complex)."""
    dt = np.dtype(dt)
    if np.issubdtype(dt, np.complexfloating):
        return LHL_C64 if _real_dtype(dt) == np.float64 else LHL_C32
    return _code(dt)


def _is_torch(x):
    return hasattr(x, "data_ptr") and hasattr(x, "is_cuda")


class _Vec:
    """A vector/matrix argument: pointer, dtype, columns, leading dimension."""

    def __init__(self, x, name="x"):
        self.obj = x
        if _is_torch(x):
            import torch
            if not x.is_cuda:
                raise ArgumentError(f"{name}: a torch tensor must be on the GPU")
            self.on_device = 1
            if x.dim() == 1:
                if not x.is_contiguous():
                    raise ArgumentError(f"{name}: tensor must be contiguous")
                self.n, self.ncol, self.ld = x.shape[0], 1, max(x.shape[0], 1)
            else:
                if x.stride(0) != 1:
                    raise ArgumentError(f"{name}: a matrix of right-hand sides must be column-major")
                self.n, self.ncol, self.ld = x.shape[0], x.shape[1], max(x.stride(1), x.shape[0], 1)
            tmap = {torch.float64: np.float64, torch.float32: np.float32,
                    torch.complex128: np.complex128, torch.complex64: np.complex64}
            if x.dtype not in tmap:
                raise ArgumentError(f"{name}: unsupported tensor dtype {x.dtype}")
            self.dtype = np.dtype(tmap[x.dtype])
            self.ptr = x.data_ptr()
        else:
            if not isinstance(x, np.ndarray):
                raise ArgumentError(f"{name}: expected a numpy array or a GPU torch tensor")
            self.on_device = 0
            self.dtype = x.dtype
            _real_dtype(self.dtype)
            if x.ndim == 1:
                if not x.flags.c_contiguous:
                    raise ArgumentError(f"{name}: array must be contiguous")
                self.n, self.ncol, self.ld = x.shape[0], 1, max(x.shape[0], 1)
            elif x.ndim == 2:
                isz = x.itemsize
                if x.flags.f_contiguous:
                    self.n, self.ncol, self.ld = x.shape[0], x.shape[1], max(x.shape[0], 1)
                elif (x.strides[0] == isz and x.strides[1] > 0 and x.strides[1] % isz == 0
                      and x.strides[1] >= x.shape[0] * isz):
                    # a column-major view with padded columns (e.g. big[:n, :m] of
                    # an order='F' array): its pointer and leading dimension
                    self.n, self.ncol, self.ld = x.shape[0], x.shape[1], x.strides[1] // isz
                else:
                    raise ArgumentError(f"{name}: a matrix of right-hand sides must be column-major (order='F')")
            else:
                raise ArgumentError(f"{name}: expected a vector or a matrix")
            self.ptr = x.ctypes.data
        self.complex = int(np.issubdtype(self.dtype, np.complexfloating))
        self.code = _code(self.dtype)


class LHLShift:
    """LHLShift{TG}: the LU of one sigma*I + tau*H (src:84-116).

    LHLShift(ws, dtype) holds an extra shift on ws's reduction; LHLShift(n, dtype)
    is an unbound state that attaches to the first workspace it is used with.
    """

    def __init__(self, ws_or_n, dtype=np.complex128):
        self.dtype = np.dtype(dtype)
        _real_dtype(self.dtype)
        self._ptr = None
        self._ws = None
        self._own = False
        self.sigma = self.dtype.type(0)
        self.tau = self.dtype.type(0)
        self.info = 0
        self._loaded = False
        if isinstance(ws_or_n, LHLWorkspace):
            self._n = ws_or_n.n
            self._bind(ws_or_n)
        else:
            self._n = int(ws_or_n)

    @property
    def is_complex(self):
        return bool(np.issubdtype(self.dtype, np.complexfloating))

    @property
    def n(self):
        return self._n

    def _bind(self, ws):
        if self._ws is ws:
            return
        _check_shift_type(self, ws)
        if self._ptr is not None and not self._own:
            load_library().lhl_shift_destroy(self._ptr)
        p = ctypes.c_void_p()
        _check(ws._h, load_library().lhl_shift_create(ws._h, ctypes.byref(p), int(self.is_complex)))
        self._ptr = p
        self._ws = ws
        self._loaded = False

    def __del__(self):
        try:
            if self._ptr is not None and not self._own and self._ws is not None and self._ws._h is not None:
                _lib.lhl_shift_destroy(self._ptr)
        except Exception:
            pass

    # the reference's shift fields (src:84-96)
    def _state(self):
        ws = self._ws
        n = self._n
        if ws is None or not self._loaded:
            raise LHLError("no shift loaded")
        rdt = _real_dtype(ws.dtype)
        cd = np.dtype(np.complex128 if rdt == np.float64 else np.complex64) if self.is_complex else rdt
        U = np.zeros((n, n), dtype=cd, order="F")
        l = np.zeros(max(n - 1, 1), dtype=cd)
        sw = np.zeros(max(n, 1), dtype=np.uint8)
        rd = np.zeros(max(n, 1), dtype=cd)
        s2 = (ctypes.c_double * 2)()
        t2 = (ctypes.c_double * 2)()
        info = ctypes.c_int64()
        _check(ws._h, _lib.lhl_get_shift_state(ws._h, None if self._own else self._ptr, U.ctypes.data,
                                               l.ctypes.data, sw.ctypes.data, rd.ctypes.data, s2, t2,
                                               ctypes.byref(info)))
        return U, l[: max(n - 1, 0)], sw[:n].astype(bool), rd[:n]

    @property
    def U(self):
        return self._state()[0]

    @property
    def Gt(self):
        """U transposed with the multipliers on the first superdiagonal, the
        reference's Gt layout (src:70-73)."""
        U, l, _, _ = self._state()
        G = U.T.copy(order="F")
        for k in range(len(l)):
            G[k, k + 1] = l[k]
        return G

    @property
    def multipliers(self):
        return self._state()[1]

    @property
    def swap(self):
        return self._state()[2]

    @property
    def rdiag(self):
        return self._state()[3]


def _check_shift_type(sh, ws):
    T = ws.dtype
    TG = sh.dtype
    ok = TG == T or (np.issubdtype(TG, np.complexfloating) and _real_dtype(TG) == T)
    if not ok:
        raise ArgumentError(f"an LHLShift{{{TG}}} cannot serve a workspace of element type {T}")


class LHLWorkspace:
    """LHLWorkspace{T}(n; shift = T): the reduction of one J (src:118-191).
    T is Float64, Float32 or their complex forms; a complex T reduces a complex
    J and its shifts are T."""

    def __init__(self, n, dtype=np.float64, shift=None):
        L = load_library()
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.float64, np.float32, np.complex128, np.complex64):
            raise ArgumentError(f"workspace element type must be Float64, Float32 or their complex forms, "
                                f"got {self.dtype}")
        TG = np.dtype(shift) if shift is not None else self.dtype
        rt = _real_dtype(self.dtype)
        cplx = np.dtype(np.complex128 if rt == np.float64 else np.complex64)
        if TG != self.dtype and TG != cplx:
            raise ArgumentError(f"shift type must be {self.dtype} or {cplx}, got {TG}")
        self.shift_dtype = TG
        h = ctypes.c_void_p()
        code = L.lhl_create(ctypes.byref(h), int(n), _full_code(self.dtype), int(TG == cplx))
        if code != LHL_OK:
            raise LHLError(f"lhl_create failed with code {code}")
        self._h = h
        self.shift = LHLShift.__new__(LHLShift)
        sh = self.shift
        sh.dtype = TG
        sh._ptr = None
        sh._ws = self
        sh._own = True
        sh.sigma = TG.type(0)
        sh.tau = TG.type(0)
        sh.info = 0
        sh._loaded = False
        sh._n = int(n)
        self._n = int(n)
        self._user_knobs = set()      # knobs the caller set (the autotuner leaves them)
        self._tuned_knobs = set()     # knobs the autotuner set last

    @property
    def reduced(self) -> bool:
        """Whether the workspace holds a valid reduction (src:158-160): false
        from the start of a reduction until it completes, so a failed lhl()
        leaves it false."""
        r = ctypes.c_int()
        _check(self._h, _lib.lhl_get_info(self._h, None, None, None, ctypes.byref(r)))
        return bool(r.value)

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None and _lib is not None:
                _lib.lhl_destroy(self._h)
                self._h = None
        except Exception:
            pass

    @property
    def n(self):
        return self._n

    # reduction fields
    @property
    def factors(self):
        F = np.zeros((self.n, self.n), dtype=self.dtype, order="F")
        _check(self._h, _lib.lhl_get_factors(self._h, F.ctypes.data))
        return F

    @property
    def Ht(self):
        F = np.zeros((self.n, self.n), dtype=self.dtype, order="F")
        _check(self._h, _lib.lhl_get_ht(self._h, F.ctypes.data))
        return F

    @property
    def ipiv(self):
        v = np.zeros(max(self.n - 2, 1), dtype=np.int32)
        _check(self._h, _lib.lhl_get_ipiv(self._h, v.ctypes.data))
        return v[: max(self.n - 2, 0)].astype(np.int64) + 1

    @property
    def perm(self):
        p = np.zeros(max(self.n, 1), dtype=np.int32)
        _check(self._h, _lib.lhl_get_perm(self._h, p.ctypes.data, None))
        return p[: self.n].astype(np.int64) + 1

    @property
    def iperm(self):
        p = np.zeros(max(self.n, 1), dtype=np.int32)
        _check(self._h, _lib.lhl_get_perm(self._h, None, p.ctypes.data))
        return p[: self.n].astype(np.int64) + 1

    @property
    def scale(self):
        s = np.zeros(max(self.n, 1))
        _check(self._h, _lib.lhl_get_scale(self._h, s.ctypes.data, None))
        return s[: self.n].astype(_real_dtype(self.dtype))   # Vector{real(T)}

    @property
    def iscale(self):
        s = np.zeros(max(self.n, 1))
        _check(self._h, _lib.lhl_get_scale(self._h, None, s.ctypes.data))
        return s[: self.n].astype(_real_dtype(self.dtype))

    # shift fields forwarded from ws.shift (src:182-191)
    @property
    def info(self):
        return self.shift.info

    @property
    def sigma(self):
        return self.shift.sigma

    @property
    def tau(self):
        return self.shift.tau

    @property
    def Gt(self):
        return self.shift.Gt

    @property
    def rdiag(self):
        return self.shift.rdiag

    @property
    def swap(self):
        return self.shift.swap


def _as_dense_J(J):
    if _is_torch(J):
        v = _Vec(J, "J")
        if v.n != v.ncol:
            raise DimensionMismatch(f"matrix is not square: dimensions are {tuple(J.shape)}")
        return v
    J = np.asarray(J)
    if J.ndim != 2 or J.shape[0] != J.shape[1]:
        raise DimensionMismatch(f"matrix is not square: dimensions are {J.shape}")
    if J.dtype not in (np.float64, np.float32, np.complex128, np.complex64):
        J = J.astype(np.complex128 if np.iscomplexobj(J) else np.float64)
    if not (J.flags.f_contiguous or (J.strides[0] == J.itemsize and J.strides[1] >= J.shape[0] * J.itemsize
                                     and J.strides[1] % J.itemsize == 0)):
        J = np.asfortranarray(J)
    return _Vec(J, "J")


def lhl_reduce_(ws: LHLWorkspace, J, balance: bool, thread=True) -> LHLWorkspace:
    """lhl_reduce!(ws, J, balance) (src:355): reduce J into ws; J is not modified.
    `thread` is accepted and ignored (the result does not depend on it, as the
    reference's does not depend on the thread count, src:712-716).  The
    autotuner's settings for this n and element type are applied first."""
    v = _as_dense_J(J)
    from . import autotune
    autotune.apply(ws, v.n)
    _check(ws._h, _lib.lhl_reduce(ws._h, ctypes.c_void_p(v.ptr), v.n, v.ld, _full_code(v.dtype), int(bool(balance)),
                                  v.on_device))
    ws._n = v.n
    ws.shift._n = v.n
    ws.shift._loaded = False
    return ws


def lhl_(ws: LHLWorkspace, J, balance: bool = True, thread=True) -> LHLWorkspace:
    """lhl!(ws, J; balance) (src:3192)."""
    return lhl_reduce_(ws, J, balance, thread)


def lhl(J, balance: bool = True, shift=None, thread=True, dtype=None) -> LHLWorkspace:
    """lhl(J; balance, shift) (src:3179).  dtype overrides the reduction's element
    type (Float32 on a Float64 J gives the FP32 factors of the precision form)."""
    v = _as_dense_J(J)
    T = np.dtype(dtype) if dtype is not None else np.dtype(v.dtype)
    ws = LHLWorkspace(v.n, T, shift=shift)
    return lhl_reduce_(ws, J, balance, thread)


def lhl_group_(wss, Js, balance: bool = True):
    """lhl!(ws, J) for many workspaces at once (lhl_reduce_group): matrices
    of one size class share one launch per reduction step.  Every J must have
    the same element type and location; the results equal lhl_'s."""
    if len(wss) != len(Js):
        raise DimensionMismatch("one J per workspace")
    if not wss:
        return wss
    vs = [_as_dense_J(J) for J in Js]
    code0, dev0 = _full_code(vs[0].dtype), vs[0].on_device
    for v in vs:
        if _full_code(v.dtype) != code0 or v.on_device != dev0:
            raise ArgumentError("every J of a group must have one element type and location")
    from . import autotune
    for ws, v in zip(wss, vs):
        autotune.apply(ws, v.n)
    M = len(wss)
    hs = (ctypes.c_void_p * M)(*[ws._h for ws in wss])
    ps = (ctypes.c_void_p * M)(*[v.ptr for v in vs])
    ns = np.array([v.n for v in vs], dtype=np.int64)
    lds = np.array([v.ld for v in vs], dtype=np.int64)
    code = _lib.lhl_reduce_group(M, ctypes.cast(hs, ctypes.c_void_p), ctypes.cast(ps, ctypes.c_void_p),
                                 ns.ctypes.data, lds.ctypes.data, code0, int(bool(balance)), dev0)
    if code != LHL_OK:
        # the message sits on the handle the error concerns
        for ws in wss:
            msg = _lib.lhl_last_error(ws._h)
            if msg:
                _check(ws._h, code)
        _check(wss[0]._h, code)
    for ws, v in zip(wss, vs):
        ws._n = v.n
        ws.shift._n = v.n
        ws.shift._loaded = False
    del vs
    return wss


def lhl_group(Js, balance: bool = True, dtype=None):
    """lhl(J) for many matrices at once: a list of workspaces (see lhl_group_)."""
    wss = []
    for J in Js:
        v = _as_dense_J(J)
        T = np.dtype(dtype) if dtype is not None else np.dtype(v.dtype)
        wss.append(LHLWorkspace(v.n, T))
    return lhl_group_(wss, Js, balance)


def _split_shift_args(args, fname):
    if len(args) == 3 and isinstance(args[0], LHLWorkspace):
        ws, s, t = args
        return ws.shift, ws, s, t
    if len(args) == 4 and isinstance(args[0], LHLShift) and isinstance(args[1], LHLWorkspace):
        return args
    raise TypeError(f"{fname}: expected (ws, sigma, tau) or (sh, ws, sigma, tau)")


def lhl_shift_(*args):
    """lhl_shift!(ws, sigma, tau) -> ws; lhl_shift!(sh, ws, sigma, tau) -> sh (src:2456-2489)."""
    sh, ws, s, t = _split_shift_args(args, "lhl_shift_")
    _check_shift_type(sh, ws)
    s = complex(s)
    t = complex(t)
    if not sh.is_complex and (s.imag != 0 or t.imag != 0):
        raise ArgumentError(f"complex shift ({s}, {t}) on an LHLShift{{{sh.dtype}}}; build the workspace "
                            f"with a complex shift type or use a complex LHLShift")
    if not sh._own:
        sh._bind(ws)
    info = ctypes.c_int64()
    _check(ws._h, _lib.lhl_shift(ws._h, None if sh._own else sh._ptr, s.real, s.imag, t.real, t.imag,
                                 ctypes.byref(info)))
    sh._n = ws.n
    sh.info = int(info.value)
    sh.sigma = sh.dtype.type(s if sh.is_complex else s.real)
    sh.tau = sh.dtype.type(t if sh.is_complex else t.real)
    sh._loaded = True
    return args[0]


def _split_ldiv_args(args, fname):
    if len(args) == 1 and isinstance(args[0], LHLWorkspace):
        return args[0].shift, args[0]
    if len(args) == 2 and isinstance(args[0], LHLShift) and isinstance(args[1], LHLWorkspace):
        return args
    raise TypeError(f"{fname}: expected (x, ws) or (x, sh, ws)")


def _sh_ptr(sh):
    return None if sh._own else sh._ptr


def _prep_x(x, sh, ws, name="x"):
    v = _Vec(x, name)
    if v.n != ws.n:
        raise DimensionMismatch(f"{name} has length {v.n}, the workspace is {ws.n}x{ws.n}")
    if sh.n != ws.n:
        raise DimensionMismatch(f"the LHLShift is {sh.n}x{sh.n}, the workspace {ws.n}x{ws.n}")
    if sh.is_complex and not v.complex:
        raise ArgumentError("a real x cannot hold the solution of a complex shift; use a complex vector")
    if not sh._own and sh._ws is not ws:
        raise ArgumentError("the LHLShift is bound to another workspace")
    return v


def lhl_ldiv_(x, *args):
    """lhl_ldiv!(x, ws) / lhl_ldiv!(x, sh, ws) (src:3052-3123): x <- W^-1 x in place."""
    sh, ws = _split_ldiv_args(args, "lhl_ldiv_")
    v = _prep_x(x, sh, ws)
    _check(ws._h, _lib.lhl_ldiv(ws._h, _sh_ptr(sh), ctypes.c_void_p(v.ptr), v.n, v.ncol, v.ld, v.code, v.complex,
                                v.on_device))
    return x


def lhl_refine_(x, A, b, *args, tol: float = 0.0):
    """lhl_refine!(x, A, b, ws, steps) / lhl_refine!(x, A, b, sh, ws, steps)
    (src:3144-3157).  A is the dense system matrix, or None for sigma*I + tau*J
    with the workspace's resident J.  tol > 0 stops early once every column's
    ||b - A x||_inf <= tol ||b||_inf."""
    if len(args) == 2:
        ws, steps = args
        sh = ws.shift
    elif len(args) == 3:
        sh, ws, steps = args
    else:
        raise TypeError("lhl_refine_: expected (x, A, b, ws, steps) or (x, A, b, sh, ws, steps)")
    if steps <= 0:
        return x
    v = _prep_x(x, sh, ws)
    bv = _Vec(b, "b")
    if bv.n != v.n or bv.ncol != v.ncol or bv.dtype != v.dtype or bv.on_device != v.on_device:
        raise DimensionMismatch("b must match x in shape, element type and location")
    if bv.ld != v.ld:
        raise DimensionMismatch("b and x must share a leading dimension")
    Wp, ldw = None, 0
    keep = None
    if A is not None:
        if _is_torch(A):
            # the library reads A with x's element type and location: convert
            # (an FP64 A with an FP32 x is cast, not reinterpreted)
            import torch
            tdt = {np.dtype(np.float64): torch.float64, np.dtype(np.float32): torch.float32,
                   np.dtype(np.complex128): torch.complex128, np.dtype(np.complex64): torch.complex64}[v.dtype]
            if not v.on_device:
                raise ArgumentError("A is a GPU tensor but x is in host memory; pass A as a numpy array")
            keep = A if A.dtype == tdt else A.to(tdt)
            if keep.dim() != 2 or keep.stride(0) != 1:
                keep = keep.t().contiguous().t()
            av = _Vec(keep, "A")
        else:
            Aa = np.asfortranarray(np.asarray(A), dtype=v.dtype)
            if v.on_device:
                import torch
                keep = torch.from_numpy(Aa).to("cuda")
                keep = keep.t().contiguous().t()
                av = _Vec(keep, "A")
            else:
                keep = Aa
                av = _Vec(Aa, "A")
        if av.n != v.n or av.ncol != v.n:
            raise DimensionMismatch("A must be n x n")
        Wp, ldw = ctypes.c_void_p(av.ptr), av.ld
    iters = ctypes.c_int()
    code = _lib.lhl_refine(ws._h, _sh_ptr(sh), ctypes.c_void_p(v.ptr), Wp, ldw, ctypes.c_void_p(bv.ptr),
                           v.n, v.ncol, v.ld, v.code, v.complex, v.on_device, int(steps), float(tol),
                           ctypes.byref(iters))
    del keep
    _check(ws._h, code, x)
    return x


def applyZ_(x, ws: LHLWorkspace):
    """applyZ!(x, ws) (src:2069): x <- Z x."""
    v = _Vec(x)
    if v.n != ws.n:
        raise DimensionMismatch(f"x has length {v.n}, the workspace is {ws.n}x{ws.n}")
    _check(ws._h, _lib.lhl_applyZ(ws._h, ctypes.c_void_p(v.ptr), v.n, v.ncol, v.ld, v.code, v.complex, v.on_device))
    return x


def applyZinv_(x, ws: LHLWorkspace):
    """applyZinv!(x, ws) (src:2104): x <- Z^-1 x."""
    v = _Vec(x)
    if v.n != ws.n:
        raise DimensionMismatch(f"x has length {v.n}, the workspace is {ws.n}x{ws.n}")
    _check(ws._h, _lib.lhl_applyZinv(ws._h, ctypes.c_void_p(v.ptr), v.n, v.ncol, v.ld, v.code, v.complex,
                                     v.on_device))
    return x


def shift_solve_batched(ws: LHLWorkspace, sigmas, taus, B, rhs_offsets, X=None, precision="working",
                        tol: float = 1e-12, max_iter: int = 10, strict: bool = True):
    """Many shifts and right-hand sides in one call: for shift s, columns
    rhs_offsets[s]:rhs_offsets[s+1] of B (n x m, FP64 or complex128, column-major)
    are solved against sigma_s I + tau_s J.  precision="refine64" refines with the
    FP64 J until ||r||_inf <= tol ||b||_inf.  Returns (X, info, iters).  When a
    shift still has a column above tol after max_iter rounds, strict=True raises
    LHLNotConverged (its .result holds the return value); strict=False returns."""
    sig = np.asarray(sigmas, dtype=np.complex128).reshape(-1)
    tau = np.asarray(taus, dtype=np.complex128).reshape(-1)
    if sig.shape != tau.shape:
        raise DimensionMismatch("sigmas and taus differ in length")
    S = sig.shape[0]
    cplx = bool(np.any(sig.imag != 0) or np.any(tau.imag != 0)) or bool(np.issubdtype(ws.dtype, np.complexfloating))
    if _is_torch(B):
        cplx = cplx or bool(B.is_complex())
    else:
        cplx = cplx or bool(np.iscomplexobj(B))
    offs = np.asarray(rhs_offsets, dtype=np.int64)
    if offs.shape != (S + 1,):
        raise DimensionMismatch("rhs_offsets must have one entry more than the shifts")
    if X is None:
        if _is_torch(B):
            X = B.clone()
            X = X.t().contiguous().t() if X.dim() == 2 else X
        else:
            X = np.zeros_like(np.asfortranarray(B, dtype=np.complex128 if cplx else np.float64))
    if not _is_torch(B):
        B = np.asfortranarray(B, dtype=np.complex128 if cplx else np.float64)
    bv, xv = _Vec(B, "B"), _Vec(X, "X")
    if bv.code != LHL_F64 or xv.code != LHL_F64 or bv.complex != int(cplx) or xv.complex != int(cplx):
        raise ArgumentError("B and X must be Float64 (complex128 for complex shifts)")
    if bv.n != ws.n or xv.n != ws.n:
        raise DimensionMismatch("B and X must have n rows")
    if bv.on_device != xv.on_device:
        raise ArgumentError("B and X must both be in host memory or both on the GPU")
    if S > 0 and (offs[0] < 0 or np.any(np.diff(offs) < 0)):
        raise DimensionMismatch("rhs_offsets must be non-negative and non-decreasing")
    if S > 0 and (offs[-1] > bv.ncol or offs[-1] > xv.ncol):
        raise DimensionMismatch(f"rhs_offsets reach column {int(offs[-1])}, beyond B ({bv.ncol}) or X ({xv.ncol})")
    s2 = np.empty(2 * S)
    t2 = np.empty(2 * S)
    s2[0::2], s2[1::2] = sig.real, sig.imag
    t2[0::2], t2[1::2] = tau.real, tau.imag
    info = np.zeros(S, dtype=np.int64)
    iters = np.zeros(S, dtype=np.int32)
    prec = LHL_PREC_REFINE64 if precision == "refine64" else LHL_PREC_WORKING
    code = _lib.lhl_shift_solve_batched(
        ws._h, S, s2.ctypes.data, t2.ctypes.data, int(cplx), offs.ctypes.data, ctypes.c_void_p(bv.ptr), bv.ld,
        ctypes.c_void_p(xv.ptr), xv.ld, bv.on_device, prec, float(tol), int(max_iter), info.ctypes.data,
        iters.ctypes.data)
    if code == LHL_NOT_CONVERGED and not strict:
        return X, info, iters
    _check(ws._h, code, (X, info, iters))
    return X, info, iters


class BatchPlan:
    """A batched shift + solve + refinement step of fixed shapes, replayed for
    new shifts as one HIP graph (lhl_plan_create / lhl_plan_run).  B and X are
    GPU tensors (float64, or complex128 for complex shifts; column-major) that
    the plan reads and writes at each run; rhs_offsets as in
    shift_solve_batched.  The refinement runs on the device (converged shifts
    drop out of later rounds); the first run learns the round count, the
    second captures the graph, later runs replay it."""

    def __init__(self, ws: LHLWorkspace, rhs_offsets, B, X, shifts_complex=False, precision="refine64",
                 tol: float = 1e-12, max_iter: int = 30):
        offs = np.ascontiguousarray(np.asarray(rhs_offsets, dtype=np.int64))
        if offs.ndim != 1 or offs.shape[0] < 2:
            raise DimensionMismatch("rhs_offsets needs at least two entries")
        if not (_is_torch(B) and _is_torch(X)):
            raise ArgumentError("a plan's B and X are GPU tensors")
        bv, xv = _Vec(B, "B"), _Vec(X, "X")
        cplx = int(bool(shifts_complex))
        if bv.code != LHL_F64 or xv.code != LHL_F64 or bv.complex != cplx or xv.complex != cplx:
            raise ArgumentError("B and X must be float64 (complex128 for complex shifts)")
        if bv.n != ws.n or xv.n != ws.n:
            raise DimensionMismatch("B and X must have n rows")
        if offs[0] < 0 or np.any(np.diff(offs) < 0):
            raise DimensionMismatch("rhs_offsets must be non-negative and non-decreasing")
        if offs[-1] > bv.ncol or offs[-1] > xv.ncol:
            raise DimensionMismatch(f"rhs_offsets reach column {int(offs[-1])}, beyond B or X")
        self.ws, self.B, self.X = ws, B, X
        self.nshift = offs.shape[0] - 1
        self.cplx = cplx
        self._offs = offs
        prec = LHL_PREC_REFINE64 if precision == "refine64" else LHL_PREC_WORKING
        p = ctypes.c_void_p()
        _check(ws._h, _lib.lhl_plan_create(ws._h, ctypes.byref(p), self.nshift, cplx, offs.ctypes.data,
                                           ctypes.c_void_p(bv.ptr), bv.ld, ctypes.c_void_p(xv.ptr), xv.ld, prec,
                                           float(tol), int(max_iter)))
        self._p = p

    def run(self, sigmas, taus, strict: bool = True):
        """X for these shifts; returns (info, iters) as numpy arrays."""
        sig = np.asarray(sigmas, dtype=np.complex128).reshape(-1)
        tau = np.asarray(taus, dtype=np.complex128).reshape(-1)
        if sig.shape != (self.nshift,) or tau.shape != (self.nshift,):
            raise DimensionMismatch(f"the plan takes {self.nshift} shifts")
        s2 = np.empty(2 * self.nshift)
        t2 = np.empty(2 * self.nshift)
        s2[0::2], s2[1::2] = sig.real, sig.imag
        t2[0::2], t2[1::2] = tau.real, tau.imag
        info = np.zeros(self.nshift, dtype=np.int64)
        iters = np.zeros(self.nshift, dtype=np.int32)
        code = _lib.lhl_plan_run(self._p, s2.ctypes.data, t2.ctypes.data, info.ctypes.data, iters.ctypes.data)
        if code == LHL_NOT_CONVERGED and not strict:
            return info, iters
        _check(self.ws._h, code, (info, iters))
        return info, iters

    def stats(self):
        """(rounds in the current graph or -1, replays, captures, rounds run beyond a graph)."""
        r, a, b, c = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _check(self.ws._h, _lib.lhl_plan_stats(self._p, ctypes.byref(r), ctypes.byref(a), ctypes.byref(b),
                                               ctypes.byref(c)))
        return int(r.value), int(a.value), int(b.value), int(c.value)

    def close(self):
        if getattr(self, "_p", None) is not None and _lib is not None:
            _lib.lhl_plan_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def profile_read_work(ws: LHLWorkspace, kernel: str):
    """(launches, total ms, shift units, column units) of one kernel family:
    the summed work the launches were given."""
    c, t = ctypes.c_int64(), ctypes.c_double()
    su, cu = ctypes.c_int64(), ctypes.c_int64()
    _check(ws._h, _lib.lhl_profile_read_work(ws._h, kernel.encode(), ctypes.byref(c), ctypes.byref(t),
                                             ctypes.byref(su), ctypes.byref(cu)))
    return int(c.value), float(t.value), int(su.value), int(cu.value)


def tune_set(ws: LHLWorkspace, knob: str, value: int):
    """Set one tuning knob of the workspace (-1 restores its default).  A knob
    set here is the caller's: the autotuner's cached choices no longer replace
    it (set -1 to hand it back)."""
    _check(ws._h, _lib.lhl_tune_set(ws._h, knob.encode(), int(value)))
    if int(value) == -1:
        ws._user_knobs.discard(knob)
    else:
        ws._user_knobs.add(knob)


def tune_get(ws: LHLWorkspace, knob: str) -> int:
    v = ctypes.c_int64()
    _check(ws._h, _lib.lhl_tune_get(ws._h, knob.encode(), ctypes.byref(v)))
    return int(v.value)


def sync(ws: LHLWorkspace):
    """Wait for the workspace's stream."""
    _check(ws._h, _lib.lhl_sync(ws._h))


def profile_enable(ws: LHLWorkspace, on: bool = True):
    """Bracket every kernel launch of ws with HIP events on its stream."""
    _check(ws._h, _lib.lhl_profile_enable(ws._h, int(bool(on))))


def profile_reset(ws: LHLWorkspace):
    _check(ws._h, _lib.lhl_profile_reset(ws._h))


def profile_read(ws: LHLWorkspace, kernel: str):
    """(launches, total milliseconds) of one kernel family since the last reset."""
    c = ctypes.c_int64()
    t = ctypes.c_double()
    _check(ws._h, _lib.lhl_profile_read(ws._h, kernel.encode(), ctypes.byref(c), ctypes.byref(t)))
    return int(c.value), float(t.value)
