"""Tuning knobs change speed, not results: the shift LU stays bit-identical to
the oracle with the wide bulk forced at every size, the reduction's pivots stay
the oracle's under every panel setting set through the knob API, and the
solves stay within the reference's tolerances for every L-sweep depth."""
import numpy as np
import pytest

from inputs import randn, stiff, tie_matrix

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_knob_api_roundtrip(lhl_hip):
    ws = lhl_hip.LHLWorkspace(10, np.float64)
    for k in lhl_hip.TUNING_KNOBS:
        assert lhl_hip.tune_get(ws, k) == -1
    lhl_hip.tune_set(ws, "reduce_nb", 8)
    assert lhl_hip.tune_get(ws, "reduce_nb") == 8
    lhl_hip.tune_set(ws, "reduce_switch", 0)
    assert lhl_hip.tune_get(ws, "reduce_switch") == 0
    lhl_hip.tune_set(ws, "reduce_switch", -1)
    assert lhl_hip.tune_get(ws, "reduce_switch") == -1
    with pytest.raises(lhl_hip.ArgumentError):
        lhl_hip.tune_set(ws, "reduce_nb", 12)
    with pytest.raises(lhl_hip.ArgumentError):
        lhl_hip.tune_set(ws, "no_such_knob", 1)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_wide_shift_bulk_bitwise(lhl_hip, oracle, monkeypatch, dtype):
    monkeypatch.setenv("LHL_TUNE", "0")
    for n in (64, 300, 1000, 2500):
        say(f"case: wide shift bulk {np.dtype(dtype).name} n={n}")
        J = randn(n + 1, n, dtype=dtype)
        ws = lhl_hip.LHLWorkspace(n, dtype)
        lhl_hip.tune_set(ws, "shift_wide_min", 1)
        lhl_hip.lhl_(ws, J)
        Ht = ws.Ht
        for s, t in ((1, -0.1), (0.5, -3)):
            lhl_hip.lhl_shift_(ws, s, t)
            ref = oracle.shift(Ht, s, t)
            assert ws.info == ref.info
            assert np.array_equal(ws.shift.swap, ref.swap.astype(bool))
            assert np.array_equal(ws.shift.U, ref.U())
            assert np.array_equal(ws.shift.multipliers, ref.mult())
            assert np.array_equal(ws.shift.rdiag, ref.rdiag)


# the split shift LU: G workgroups per shift, windows of 63 (FP32) / 64 (FP64)
# columns dealt round-robin; sizes on both sides of window boundaries and of
# the owned-column shapes, one shift at a time and batched
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_split_shift_bitwise(lhl_hip, oracle, monkeypatch, dtype):
    monkeypatch.setenv("LHL_TUNE", "0")
    kp = 64 if dtype == np.float64 else 63
    for n in (6 * kp, 6 * kp + 1, 10 * kp - 1, 10 * kp, 10 * kp + 1, 2500, 4096, 6000):
        J = randn(n + 3, n, dtype=dtype)
        ws = lhl_hip.LHLWorkspace(n, dtype)
        lhl_hip.lhl_(ws, J)
        Ht = ws.Ht
        for split in (1, 2, 4, 8):
            say(f"case: split shift {np.dtype(dtype).name} n={n} split={split}")
            lhl_hip.tune_set(ws, "shift_split", split)
            for s, t in ((1, -0.1), (0.5, -3)):
                lhl_hip.lhl_shift_(ws, s, t)
                ref = oracle.shift(Ht, s, t)
                assert ws.info == ref.info
                assert np.array_equal(ws.shift.swap, ref.swap.astype(bool))
                assert np.array_equal(ws.shift.U, ref.U())
                assert np.array_equal(ws.shift.multipliers, ref.mult())
                assert np.array_equal(ws.shift.rdiag, ref.rdiag)


def test_split_shift_batched_equals_unsplit(lhl_hip, monkeypatch):
    monkeypatch.setenv("LHL_TUNE", "0")
    n, S, R = 2048, 24, 3
    say(f"case: split shift batched n={n} S={S}")
    J = stiff(5, n)
    gam = np.logspace(-4, -1, S)
    offs = np.arange(0, S * R + 1, R, dtype=np.int64)
    B = np.asfortranarray(np.random.default_rng(9).standard_normal((n, S * R)))
    out = {}
    for split in (1, 4):
        ws = lhl_hip.LHLWorkspace(n, np.float32)
        lhl_hip.tune_set(ws, "shift_split", split)
        lhl_hip.lhl_(ws, J)
        out[split] = lhl_hip.shift_solve_batched(ws, np.ones(S), -gam, B, offs, precision="refine64", tol=1e-12,
                                                 max_iter=10)
    assert np.array_equal(out[1][0], out[4][0])
    assert np.array_equal(np.asarray(out[1][1]), np.asarray(out[4][1]))


# the row-split U back substitution (G workgroups per 16 columns of a shift)
# against the one-workgroup-per-4-columns kernel and dense FP64 solves: a
# partial last block, fewer and more than 16 columns per shift, G = 2, 4, 8,
# and a batch of more groups than fit on the device at once (several launches)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_utrsm_split_solves(lhl_hip, monkeypatch, dtype):
    monkeypatch.setenv("LHL_TUNE", "0")
    rtol = 1e-11 if dtype == np.float64 else 5e-3
    for n, S, R in ((200, 3, 16), (700, 2, 5), (1000, 2, 17), (2049, 2, 33), (4096, 2, 16), (512, 300, 16)):
        J = np.asfortranarray(randn(n + 7, n) / np.sqrt(n))
        gam = np.linspace(0.05, 0.5, S)
        offs = np.arange(0, S * R + 1, R, dtype=np.int64)
        B = np.asfortranarray(np.random.default_rng(n).standard_normal((n, S * R)))
        ws = lhl_hip.LHLWorkspace(n, dtype)
        lhl_hip.lhl_(ws, J)
        X = {}
        for split in (1, 2, -1):   # 2: tiles where G = 2 gives each thread at most two blocks
            say(f"case: utrsm split {np.dtype(dtype).name} n={n} S={S} R={R} split={split}")
            lhl_hip.tune_set(ws, "utrsm_split", split)
            X[split], info, _ = lhl_hip.shift_solve_batched(ws, np.ones(S), -gam, B, offs)
            assert np.all(np.asarray(info) == 0)
        for s in sorted({0, S // 2, S - 1}):
            c = slice(offs[s], offs[s + 1])
            ref = np.linalg.solve(np.eye(n) - gam[s] * J, B[:, c])
            den = np.abs(ref).max()
            err = {split: np.abs(X[split][:, c] - ref).max() / den for split in (1, 2, -1)}
            # FP64: all at the dense solve; FP32: the split kernel no further
            # from it than the other (the FP32 factors set the error)
            assert err[1] <= rtol, (n, S, R, s, err)
            for split in (2, -1):
                assert err[split] <= (rtol if dtype == np.float64 else 3 * err[1] + 1e-6), (n, S, R, s, err)


# complex shifts on a real J (U in two planes of tiles and the row-split U
# solve; FP64 up to n = 4096, packed columns beyond, as at 4160): against the
# one-workgroup-per-4-columns kernel and dense complex FP64 solves
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_utrsm_split_complex_shifts(lhl_hip, monkeypatch, dtype):
    monkeypatch.setenv("LHL_TUNE", "0")
    rtol = 5e-3 if dtype == np.float32 else 1e-11
    for n, S, R in ((300, 3, 8), (1000, 2, 9), (4096, 2, 8), (4160, 1, 3)):
        J = np.asfortranarray(randn(n + 11, n) / np.sqrt(n))
        gam = np.linspace(0.05, 0.5, S) * (0.6 + 0.3j)
        offs = np.arange(0, S * R + 1, R, dtype=np.int64)
        rng = np.random.default_rng(n + 1)
        B = np.asfortranarray(rng.standard_normal((n, S * R)) + 1j * rng.standard_normal((n, S * R)))
        ws = lhl_hip.LHLWorkspace(n, dtype)
        lhl_hip.lhl_(ws, J)
        X = {}
        for split in (1, -1):
            say(f"case: utrsm split complex shifts {np.dtype(dtype).name} n={n} S={S} R={R} split={split}")
            lhl_hip.tune_set(ws, "utrsm_split", split)
            X[split], info, _ = lhl_hip.shift_solve_batched(ws, np.ones(S), -gam, B, offs)
            assert np.all(np.asarray(info) == 0)
        for s in sorted({0, S - 1}):
            c = slice(offs[s], offs[s + 1])
            ref = np.linalg.solve(np.eye(n) - gam[s] * J, B[:, c])
            den = np.abs(ref).max()
            err = {split: np.abs(X[split][:, c] - ref).max() / den for split in (1, -1)}
            assert err[1] <= rtol, (n, S, R, s, err)
            assert err[-1] <= (rtol if dtype == np.float64 else 3 * err[1] + 1e-6), (n, S, R, s, err)


def test_reduce_knobs_keep_pivots(lhl_hip, oracle, monkeypatch):
    monkeypatch.setenv("LHL_TUNE", "0")
    n = 600
    for J in (randn(61, n), tie_matrix(62, 200, -1, 1)):
        m = J.shape[0]
        red = oracle.reduce(J, blocked=0)
        for knobs in ({"reduce_blocked_min": 3, "reduce_switch": 0}, {"reduce_blocked_min": 3, "reduce_nb": 8,
                                                                     "reduce_rmin": 4, "reduce_switch": 500},
                      {"reduce_blocked_min": 3, "reduce_rmin": 64}, {"reduce_blocked_min": 1 << 30}):
            say(f"case: reduce knobs n={m} {knobs}")
            ws = lhl_hip.LHLWorkspace(m, np.float64)
            for k, v in knobs.items():
                lhl_hip.tune_set(ws, k, v)
            lhl_hip.lhl_(ws, J)
            assert np.array_equal(ws.ipiv - 1, red.ipiv), knobs
            rel = np.linalg.norm(ws.factors - red.A) / np.linalg.norm(red.A)
            assert rel <= 1e-12 * m, (knobs, rel)


def test_lsweep_depth_solves(lhl_hip, monkeypatch):
    monkeypatch.setenv("LHL_TUNE", "0")
    n, S, R = 1100, 3, 2
    J = stiff(1100, n)
    gam = np.logspace(-4, -1, S)
    offs = np.arange(0, S * R + 1, R, dtype=np.int64)
    B = np.asfortranarray(np.random.default_rng(5).standard_normal((n, S * R)))
    ws = lhl_hip.lhl(J)
    X0 = None
    for depth in (0, 1, 2, 3):
        say(f"case: lsweep depth {depth}")
        lhl_hip.tune_set(ws, "lsweep_depth", depth)
        X, info, _ = lhl_hip.shift_solve_batched(ws, np.ones(S), -gam, B, offs)
        assert np.all(info == 0)
        for s in range(S):
            W = np.eye(n) - gam[s] * J
            c = slice(offs[s], offs[s + 1])
            r = np.abs(B[:, c] - W @ X[:, c]).max() / np.abs(B[:, c]).max()
            assert r <= 1e-10, (depth, s, r)
        if X0 is None:
            X0 = X
        assert np.linalg.norm(X - X0) <= 1e-11 * np.linalg.norm(X0)
