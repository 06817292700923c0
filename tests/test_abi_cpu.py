"""The C-ABI library loads and exports every symbol include/lhl_hip.h declares
(no GPU work), and the host mirror's argument checks raise the reference's
error classes before any device call."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "lhl_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(lhl_[A-Za-z_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(lhl_hip):
    lib = ctypes.CDLL(lhl_hip.library_path())
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s


def test_julia_bindings_call_declared_symbols(lhl_hip):
    """Every symbol the Julia module and its LinearSolve adapter ccall is
    declared in the header and exported by the built library."""
    lib = ctypes.CDLL(lhl_hip.library_path())
    decl = set(declared_symbols())
    jdir = os.path.join(ROOT, "lhl.jl_amd", "julia")
    called = set()
    for f in os.listdir(jdir):
        if f.endswith(".jl"):
            txt = open(os.path.join(jdir, f)).read()
            called |= set(re.findall(r"ccall\(\(:(lhl_[A-Za-z_]+)", txt))
            called |= set(re.findall(r"\(:\w+!?, :(lhl_[A-Za-z_]+)\)", txt))
    assert {"lhl_create", "lhl_reduce", "lhl_shift", "lhl_ldiv", "lhl_refine", "lhl_applyZ",
            "lhl_shift_solve_batched", "lhl_get_shift_info"} <= called
    for s in sorted(called):
        assert s in decl, s
        assert hasattr(lib, s), s


def test_version_without_gpu(lhl_hip):
    assert lhl_hip.load_library().lhl_version() == 1


def test_workspace_type_errors(lhl_hip):
    # shift type must be T or Complex{T} (src:167-169)
    with pytest.raises(lhl_hip.ArgumentError):
        lhl_hip.LHLWorkspace.__init__(object.__new__(lhl_hip.LHLWorkspace), 3, np.float64, shift=np.float32)
    with pytest.raises(lhl_hip.ArgumentError):
        lhl_hip.LHLWorkspace.__init__(object.__new__(lhl_hip.LHLWorkspace), 3, np.float64, shift=np.complex64)


def test_non_square_is_dimension_mismatch(lhl_hip):
    with pytest.raises(lhl_hip.DimensionMismatch):
        lhl_hip.lhl(np.zeros((3, 4)))


def test_shift_dtype_mismatch_is_argument_error(lhl_hip):
    class FakeWS:
        pass
    sh = lhl_hip.LHLShift(3, np.complex128)
    ws = object.__new__(lhl_hip.LHLWorkspace)
    ws.dtype = np.dtype(np.float32)
    with pytest.raises(lhl_hip.ArgumentError):
        lhl_hip.lhl_hip_check = lhl_hip._check_shift_type(sh, ws)


def test_no_cpu_fallback_in_product_package():
    """The product package must not import the oracle."""
    pkg = os.path.join(ROOT, "lhl.jl_amd")
    for dp, _, fs in os.walk(pkg):
        for f in fs:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                txt = open(os.path.join(dp, f)).read()
                assert "oracle" not in txt.replace("the oracle", ""), f
