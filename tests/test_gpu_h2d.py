"""A host J's upload to the device: staged in chunks of whole columns through
two pinned halves (the knob h2d_chunk_bytes), the reduction of a host J is
bitwise the reduction of the same J already on the device, for chunk sizes
that split the columns unevenly, a padded leading dimension, one chunk, and
the runtime's own copy (0).  The reduced flag (src:158-160) reads the handle:
false before any reduction, true after one, and a call rejected before the
reduction starts leaves it as it was."""
import numpy as np
import pytest

from inputs import randn, stiff

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def state(ws):
    return ws.ipiv.copy(), ws.perm.copy(), ws.scale.copy(), ws.factors.copy()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_host_upload_matches_device_J(lhl_hip, dtype):
    import torch
    for n, pad in ((300, 0), (1500, 7)):
        J = stiff(81, n)
        big = np.zeros((n + pad, n), order="F")
        big[:n, :] = J
        Jh = big[:n, :]                       # a padded column-major view
        assert pad == 0 or not Jh.flags.f_contiguous
        say(f"case: device J n={n} {np.dtype(dtype).name}")
        ws = lhl_hip.LHLWorkspace(n, dtype)
        lhl_hip.lhl_(ws, torch.from_numpy(np.asfortranarray(J)).cuda())
        ref = state(ws)
        for chunk in (-1, 0, 1 << 16, 3 * 8 * n + 5, 1 << 30):
            say(f"case: host J n={n} pad={pad} chunk={chunk}")
            lhl_hip.tune_set(ws, "h2d_chunk_bytes", chunk)
            lhl_hip.lhl_(ws, Jh)
            got = state(ws)
            for a, b in zip(got, ref):
                assert np.array_equal(a, b), (n, chunk)
        # the resident FP64 J is the host J (the refinement reads it)
        lhl_hip.tune_set(ws, "h2d_chunk_bytes", 1 << 16)
        lhl_hip.lhl_(ws, Jh)
        lhl_hip.lhl_shift_(ws, 1.0, -0.01)
        b = np.random.default_rng(2).standard_normal(n)
        x = lhl_hip.lhl_ldiv_(b.astype(dtype), ws)
        x64 = x.astype(np.float64)
        lhl_hip.lhl_refine_(x64, None, b, ws, 5)
        W = np.eye(n) - 0.01 * J
        assert np.abs(b - W @ x64).max() <= 1e-12 * np.abs(b).max()


def test_reduced_flag(lhl_hip):
    n = 64
    ws = lhl_hip.LHLWorkspace(n, np.float64)
    assert ws.reduced is False
    lhl_hip.lhl_(ws, randn(82, n))
    assert ws.reduced is True
    with pytest.raises(lhl_hip.DimensionMismatch):
        lhl_hip.lhl_(ws, np.zeros((n, n + 1)))    # rejected before the reduction starts
    assert ws.reduced is True
    ws2 = lhl_hip.LHLWorkspace(n, np.float64)
    with pytest.raises(lhl_hip.LHLError):
        lhl_hip.lhl_shift_(ws2, 1.0, -0.1)        # no reduction held
    assert ws2.reduced is False
