"""The autotuner: picks the library's tuning knobs per (element type, n) by
timing them on the device, and applies the stored choice to every reduction.

The reference fixes these as constants (src/LHLFactorization.jl:703-706
_lhl_block_min / _lhl_panel_width, 1451 _LHL_GEMV_GROUP, 2494
_lhl_shift_fused_min).  Here they are per-handle settings (lhl_tune_set):

  reduce_blocked_min, reduce_nb, reduce_rmin, reduce_switch   lhl(): whether and how the
      delayed-update panels run (width, rows per step workgroup, the step where the
      row-resident loop takes over)
  lsweep_depth, shift_wide_min, shift_split, utrsm_split      the batched solve: halvings of
      the L sweeps' triangular products, the n from which the shift LU's bulk is wide, the
      workgroups per shift LU, and the row-split U solve's workgroups per 16 columns

Search: coordinate descent from the built-in defaults, one knob at a time over
its candidate values, keeping a change only when it is faster by more than
`gain` (noise guard), for `rounds` passes.  Every timing is the minimum over
`reps` runs of the whole call (device-resident inputs, synchronized).  The
result goes into a JSON cache keyed "f64:4096"; `apply` uses the entry with the
nearest n of the same element type (within a factor of 1.5), so a cache tuned
at a few sizes covers the sizes between them.

Every knob leaves the results unchanged beyond rounding: the pivots of a
reduction are the same for every panel setting (the GPU tests check this), so
tuning never trades accuracy for time.

Command line (on a GPU box):  python -m lhl_hip.autotune --n 2048 4096 --dtype f64 f32
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_CACHE = os.path.join(_HERE, "tuned.json")

REDUCE_CANDIDATES = {
    "reduce_blocked_min": [-1, 512, 1 << 30],     # default (FP64 from 1024, FP32 never), from 512, never
    "reduce_switch": [0, 200, 300, 400, 500, 600, 800],
    "reduce_nb": [8, 16],
    "reduce_rmin": [8, 16, 32],
}
SOLVE_CANDIDATES = {
    "lsweep_depth": [0, 1, 2, 3],
    "shift_wide_min": [-1, 1],                     # default threshold, or wide at every size
    "shift_split": [-1, 2, 8],                     # workgroups per shift LU (default 4)
    "utrsm_split": [-1, 1, 4],                     # row-split U solve (default up to 8; 1: packed columns)
}

_cache = None


def cache_path() -> str:
    return os.environ.get("LHL_TUNE_CACHE", DEFAULT_CACHE)


def load_cache() -> dict:
    global _cache
    if _cache is None:
        p = cache_path()
        try:
            with open(p) as f:
                _cache = json.load(f)
        except (OSError, ValueError):
            _cache = {}
    return _cache


def _key(dtype, n) -> str:
    return f"{'f64' if np.dtype(dtype) == np.float64 else 'f32'}:{int(n)}"


def lookup(dtype, n) -> dict:
    """The cached settings for the nearest tuned n of this element type."""
    if os.environ.get("LHL_TUNE", "1") == "0":
        return {}
    c = load_cache()
    tag = "f64" if np.dtype(dtype) == np.float64 else "f32"
    best, bd = None, None
    for k, v in c.items():
        t, _, m = k.partition(":")
        if t != tag or not m.isdigit():
            continue
        m = int(m)
        if max(m, n) > 1.5 * min(m, n):
            continue
        d = abs(np.log(m / max(n, 1)))
        if bd is None or d < bd:
            best, bd = v, d
    return dict(best.get("knobs", {})) if best else {}


def apply(ws, n) -> dict:
    """Set the cached knobs of (ws.dtype, n) on the workspace, and return the
    knobs this tuner set for an earlier n to their defaults.  Knobs the caller
    set with tune_set are left alone, and LHL_TUNE=0 leaves every knob."""
    from . import _check, _lib
    if os.environ.get("LHL_TUNE", "1") == "0":
        return {}
    if np.issubdtype(ws.dtype, np.complexfloating):   # the knobs tune the real paths only
        return {}
    knobs = {k: v for k, v in lookup(ws.dtype, n).items() if k not in ws._user_knobs}

    def put(k, v):
        _check(ws._h, _lib.lhl_tune_set(ws._h, k.encode(), int(v)))
    for k in ws._tuned_knobs - set(knobs) - ws._user_knobs:
        put(k, -1)
    for k, v in knobs.items():
        put(k, v)
    ws._tuned_knobs = set(knobs)
    return knobs


def _stiff(seed, n):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((n, n))
    u = np.linspace(0.0, 6.0, n)
    return np.asfortranarray(G / np.sqrt(n) - np.diag(10.0 ** u))


def _time(fn, reps):
    import torch
    best = float("inf")
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best


def _descend(ws, candidates, run, reps, gain, rounds, log):
    from . import tune_set
    cur = {k: -1 for k in candidates}
    for k, v in cur.items():
        tune_set(ws, k, v)
    run()
    base = _time(run, reps)
    for _ in range(rounds):
        changed = False
        for k, vals in candidates.items():
            for v in vals:
                if v == cur[k]:
                    continue
                tune_set(ws, k, v)
                t = _time(run, reps)
                log(f"  {k}={v}: {1e3 * t:.2f} ms (best {1e3 * base:.2f})")
                if t < base * (1 - gain):
                    base, cur[k], changed = t, v, True
                else:
                    tune_set(ws, k, cur[k])
        if not changed:
            break
    return cur, base


def tune(n, dtype=np.float64, shifts=64, rhs=16, reps=3, gain=0.01, rounds=2, log=print):
    """Tune the knobs for (dtype, n) on the current GPU; returns the entry."""
    import torch
    from . import LHLWorkspace, lhl_, shift_solve_batched, tune_set, load_library
    load_library()
    dev = torch.device("cuda:0")
    J = torch.from_numpy(_stiff(4096, n)).to(dev)
    ws = LHLWorkspace(n, dtype)
    os.environ["LHL_TUNE"] = "0"   # time the knobs as set here (tune_set marks them the caller's anyway)
    try:
        def red():
            lhl_(ws, J)
        log(f"autotune {_key(dtype, n)}: lhl()")
        kred, tred = _descend(ws, REDUCE_CANDIDATES, red, reps, gain, rounds, log)
        for k, v in kred.items():
            tune_set(ws, k, v)
        lhl_(ws, J)
        S, R = shifts, rhs
        gam = np.logspace(-4, -1, S)
        offs = np.arange(0, S * R + 1, R, dtype=np.int64)
        B = torch.from_numpy(np.asfortranarray(np.random.default_rng(7).standard_normal((n, S * R)))).to(dev)
        X = torch.zeros_like(B)

        def solve():
            shift_solve_batched(ws, np.ones(S), -gam, B, offs, X=X, precision="refine64", tol=1e-12, max_iter=30)
        log(f"autotune {_key(dtype, n)}: batched solve {S}x{R}")
        ksol, tsol = _descend(ws, SOLVE_CANDIDATES, solve, reps, gain, rounds, log)
    finally:
        os.environ.pop("LHL_TUNE", None)
    knobs = {k: v for k, v in {**kred, **ksol}.items() if v != -1}
    return {"knobs": knobs, "lhl_ms": 1e3 * tred, "solve_ms": 1e3 * tsol,
            "workload": f"stiff J n={n}; batched {S} shifts x {R} RHS refined to 1e-12"}


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="tune liblhl_hip's knobs on this GPU")
    ap.add_argument("--n", type=int, nargs="+", default=[4096])
    ap.add_argument("--dtype", nargs="+", default=["f64", "f32"])
    ap.add_argument("--out", default=cache_path())
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args(argv)
    out = {}
    if os.path.exists(args.out):
        with open(args.out) as f:
            out = json.load(f)
    for dt in args.dtype:
        for n in args.n:
            d = np.float64 if dt == "f64" else np.float32
            e = tune(n, d, reps=args.reps, log=lambda *a: print(*a, flush=True))
            out[_key(d, n)] = e
            print(f"{_key(d, n)}: {e['knobs']} lhl {e['lhl_ms']:.2f} ms, solve {e['solve_ms']:.2f} ms", flush=True)
            with open(args.out, "w") as f:
                json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
