"""Grouped reductions (lhl_reduce_group): many independent matrices of uneven
n, one launch per reduction step for each size class.  Every workspace equals
the one lhl() gives for its matrix alone, bit for bit (pivots, permutation,
scales, factors), in both precisions, with matrices of several size classes,
the smallest orders, host and device J, and a blocked-size matrix mixed in;
and the grouped workspaces solve as the lone ones do."""
import numpy as np
import pytest

from inputs import randn, stiff, tie_matrix

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def same(a, b):
    return (np.array_equal(a.ipiv, b.ipiv) and np.array_equal(a.perm, b.perm) and np.array_equal(a.scale, b.scale)
            and np.array_equal(a.factors, b.factors))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_group_equals_lone(lhl_hip, monkeypatch, dtype):
    monkeypatch.setenv("LHL_TUNE", "0")
    rng = np.random.default_rng(11)
    ns = [3, 4, 17, 64, 129, 300, 700, 1500, 2047, 2049, 2600, 900, 130, 5]
    Js = []
    for i, n in enumerate(ns):
        if i % 3 == 0:
            Js.append(stiff(300 + i, n))
        elif i % 3 == 1:
            Js.append(randn(300 + i, n))
        else:
            Js.append(tie_matrix(300 + i, n))
    say(f"case: group of {len(ns)} {np.dtype(dtype).name}")
    wss = lhl_hip.lhl_group(Js, dtype=dtype)
    for i, (J, ws) in enumerate(zip(Js, wss)):
        lone = lhl_hip.lhl(J, dtype=dtype)
        assert same(ws, lone), (i, ns[i])
        assert ws.reduced
    # the same group again in place, from device J (the handles keep their n)
    import torch
    Jd = [torch.from_numpy(np.asfortranarray(J)).cuda() for J in Js[::-1]]
    wss2 = [lhl_hip.LHLWorkspace(J.shape[0], dtype) for J in Jd]
    lhl_hip.lhl_group_(wss2, Jd)
    for J, ws in zip(Js[::-1], wss2):
        assert same(ws, lhl_hip.lhl(J, dtype=dtype))
    # solves on grouped workspaces
    for i in (2, 6, 9):
        n = ns[i]
        b = rng.standard_normal(n)
        for ws in (wss[i], lhl_hip.lhl(Js[i], dtype=dtype)):
            lhl_hip.lhl_shift_(ws, 1.0, -0.01)
        x1 = lhl_hip.lhl_ldiv_(b.astype(dtype), wss[i])
        lone = lhl_hip.lhl(Js[i], dtype=dtype)
        lhl_hip.lhl_shift_(lone, 1.0, -0.01)
        x2 = lhl_hip.lhl_ldiv_(b.astype(dtype), lone)
        assert np.array_equal(x1, x2)


def test_group_with_blocked_member_and_errors(lhl_hip, monkeypatch):
    monkeypatch.setenv("LHL_TUNE", "0")
    monkeypatch.setenv("LHL_BLOCKED_MIN", "600")       # the 800 matrix takes the blocked path alone
    Js = [randn(400, 300), stiff(401, 800), randn(402, 1000)]
    wss = lhl_hip.lhl_group(Js)
    for J, ws in zip(Js, wss):
        assert same(ws, lhl_hip.lhl(J))
    with pytest.raises(lhl_hip.ArgumentError):
        lhl_hip.lhl_group_([wss[0], wss[0]], [Js[0], Js[0]])       # one handle twice
    with pytest.raises(lhl_hip.DimensionMismatch):
        lhl_hip.lhl_group_(wss, Js[:2])
