#!/usr/bin/env python3
"""Benchmark of the LHL hot path on one MI355X.

Workload (the batched precision form of the shifted-solve family): n = 4096,
J = -diag(10^u) + G/sqrt(n) (u linearly spaced in [0, 6], G ~ N(0,1), seeded),
64 real shifts (1, -gamma) with gamma log-spaced in [1e-4, 1e-1], 16 N(0,1)
right-hand sides per shift.  One step = one batched pass: the 64 shifted
Hessenberg LUs and the 1024 solves, refined with FP64 residuals until every
column has ||b - W x||_inf <= 1e-12 ||b||_inf.

  value               FP32-factor + FP64-refinement solves/s (whole batch / time)
  fp64_solves_per_s   the same batch with FP64 factors (refined to the same bar)
  lhl_ms              the reduction lhl(J) with J resident in HBM (FP64 and FP32)

J and the right-hand sides are resident in HBM before the timed region; the
timed region brackets K steps with a device synchronization on both sides.
The roofline object describes the kernel with the largest share of the step,
timed live with HIP events on the library's stream; the CPU baseline is the
C restatement of the reference algorithm (oracle/) on this host's cores, on a
bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "lhl.jl_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "solves/s @n=4096 (FP32+refine vs FP64, rel resid≤1e-12) + lhl() time"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def stiff(seed, n):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((n, n))
    u = np.linspace(0.0, 6.0, n)
    return np.asfortranarray(G / np.sqrt(n) - np.diag(10.0 ** u))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic(kernel, n, S, ncol, es):
    """Algorithmic bytes (or flops) one launch of a kernel family moves for the
    whole batch: (amount, unit, bound)."""
    tri = n * (n + 1) / 2
    if kernel == "shift":
        # Ht's upper Hessenberg part read once, S packed U's written
        return (tri + n) * es + S * tri * es, "bytes", "hbm"
    if kernel == "hess":
        # the forward interchange chain: every column read and written once,
        # each shift's multipliers and swap flags read once
        return 2 * ncol * n * es + S * 2 * n * es, "bytes", "hbm"
    if kernel in ("trsm", "utrsm", "trmm"):
        # the multipliers' triangle and the columns once each: n^2 ncol FMAs
        return 2.0 * (n * n / 2) * ncol, "flops", "valu"
    if kernel == "residual":
        # J x for every column in FP64: 2 n^2 flops per column
        return 2.0 * n * n * ncol, "flops", "mfma"
    if kernel == "reduce_trailing":
        return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--shifts", type=int, default=64)
    ap.add_argument("--rhs", type=int, default=16)
    ap.add_argument("--tol", type=float, default=1e-12)
    ap.add_argument("--cpu-baseline", type=int, default=1, help="0 skips the CPU leg")
    ap.add_argument("--lhl-reps", type=int, default=2)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (a multi-GPU run would shard the shift batch "
              "by shifts after one broadcast of the reduction); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    import lhl_hip

    lhl_hip.load_library()
    dev = torch.device("cuda:0")
    n, S, R = args.n, args.shifts, args.rhs
    ncol = S * R
    log(f"bench: n={n} shifts={S} rhs/shift={R}")
    J = stiff(4096, n)
    gam = np.logspace(-4, -1, S)
    sig = np.ones(S)
    tau = -gam
    offs = np.arange(0, ncol + 1, R, dtype=np.int64)
    rng = np.random.default_rng(7)
    Bh = np.asfortranarray(rng.standard_normal((n, ncol)))
    Jd = torch.from_numpy(J).to(dev)                  # column-major strides preserved
    Bd = torch.from_numpy(Bh).to(dev)
    Xd = torch.zeros_like(Bd)
    assert Jd.stride(0) == 1 and Bd.stride(0) == 1 and Xd.stride(0) == 1
    torch.cuda.synchronize()

    # ---- lhl(): FP64 and FP32 factors from the resident FP64 J
    ws64 = lhl_hip.LHLWorkspace(n, np.float64)
    ws32 = lhl_hip.LHLWorkspace(n, np.float32)
    lhl_ms = {}
    for name, ws in (("fp64", ws64), ("fp32", ws32)):
        times = []
        for r in range(args.lhl_reps + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            lhl_hip.lhl_(ws, Jd)
            lhl_hip.sync(ws)
            t1 = time.perf_counter()
            if r > 0 or args.lhl_reps == 0:
                times.append(t1 - t0)
        lhl_ms[name] = 1e3 * min(times)
        log(f"bench: lhl {name} {lhl_ms[name]:.2f} ms")

    def step(ws):
        _, info, iters = lhl_hip.shift_solve_batched(ws, sig, tau, Bd, offs, X=Xd, precision="refine64",
                                                     tol=args.tol, max_iter=30)
        return info, iters

    results = {}
    for name, ws in (("fp64", ws64), ("fp32", ws32)):
        for _ in range(args.warmup):
            step(ws)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            info, iters = step(ws)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / max(args.steps, 1)
        # verify the accuracy bar on the device in FP64: r = b - (x - gamma J x)
        JX = Jd @ Xd
        g = torch.from_numpy(np.repeat(gam, R)).to(dev)
        Rd = Bd - (Xd - JX * g[None, :])
        rel = (Rd.abs().amax(0) / Bd.abs().amax(0)).max().item()
        results[name] = dict(ms=1e3 * dt, sps=ncol / dt, iters=iters.tolist(), info=int(np.max(np.abs(info))),
                             rel=rel)
        log(f"bench: {name} step {1e3 * dt:.2f} ms, {ncol / dt:.0f} solves/s, max rel resid {rel:.2e}, "
            f"refine iters {int(iters.min())}-{int(iters.max())}")

    # ---- live per-kernel timing of one FP32+refine step (outside the timed region)
    lhl_hip.profile_reset(ws32)
    lhl_hip.profile_enable(ws32, True)
    step(ws32)
    fams = ["shift", "gather", "trsm", "hess", "utrsm", "trmm", "residual", "convert"]
    prof = {k: lhl_hip.profile_read(ws32, k) for k in fams}
    lhl_hip.profile_enable(ws32, False)
    log("bench: kernels " + ", ".join(f"{k} {c}x {t:.2f} ms" for k, (c, t) in prof.items()))
    dom = max(("shift", "trsm", "hess", "utrsm", "trmm", "residual"), key=lambda k: prof[k][1])
    c, t = prof[dom]
    amt, unit, bound = algorithmic(dom, n, S, ncol, 4)
    avg_s = (t / c) / 1e3 if c else float("nan")
    if bound == "valu":
        achieved = amt / (t / 1e3) / 1e12      # all launches of the family in one pass
        roof = {"bound": "mfma", "achieved": achieved, "peak": 157.3, "unit": "TFLOP/s",
                "frac": achieved / 157.3, "traffic": None, "kernel": dom, "launches": c, "avg_ms": t / c,
                "algorithmic_flops": amt}
    elif bound == "hbm":
        # one solve launch per refinement round covers the whole batch
        achieved = amt / avg_s / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dom,
                "launches": c, "avg_ms": t / c, "algorithmic_bytes": amt}
    else:
        achieved = amt / avg_s / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None, "kernel": dom,
                "launches": c, "avg_ms": t / c, "algorithmic_flops": amt}

    # ---- CPU baseline: the C restatement of the reference algorithm (oracle/)
    cpu = None
    if args.cpu_baseline:
        from oracle import oracle as O

        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        os.environ.setdefault("OMP_NUM_THREADS", str(cores))
        log(f"bench: CPU baseline (oracle, {cores} threads for the reduction)")
        t0 = time.perf_counter()
        red = O.reduce(J)
        t_red = time.perf_counter() - t0
        ns, nr = 2, 2
        t0 = time.perf_counter()
        t_shift = t_solve = 0.0
        for s in range(ns):
            a = time.perf_counter()
            sh = O.shift(red.Ht, 1.0, float(tau[s * (S // ns)]))
            b_ = time.perf_counter()
            t_shift += b_ - a
            for r in range(nr):
                a = time.perf_counter()
                W = None
                x = O.ldiv(red, sh, Bh[:, s * R + r])
                t_solve += time.perf_counter() - a
        per_shift = t_shift / ns
        per_solve = t_solve / (ns * nr)
        # one FP64 refinement round per solve doubles the solve count; the residual is a dense matvec
        a = time.perf_counter()
        _ = J @ Bh[:, 0]
        t_mv = time.perf_counter() - a
        batch_s = S * per_shift + ncol * (2 * per_solve + t_mv)
        cpu = {"value": ncol / batch_s, "unit": "solves/s", "cores": cores, "kind": "port",
               "sample": (f"oracle (C restatement of the reference, not the Julia reference): lhl(J) at n={n} "
                          f"once ({t_red:.2f} s, OpenMP {cores} threads); {ns} shifts x {nr} solves timed serially "
                          f"({1e3 * per_shift:.1f} ms/shift, {1e3 * per_solve:.1f} ms/solve) scaled to the "
                          f"{S}x{R} batch with one refinement round per solve"),
               "lhl_s": t_red}
        log(f"bench: CPU lhl {t_red:.2f} s, {ncol / batch_s:.2f} solves/s")

    fp32, fp64 = results["fp32"], results["fp64"]
    out = {
        "metric": METRIC,
        "value": fp32["sps"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fp32["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 factors/solves + f64 residual/refinement",
        "data": "synthetic (seeded stiff J and N(0,1) right-hand sides)",
        "config": {"workload": (f"n={n} stiff J=-diag(10^u)+G/sqrt(n), {S} real shifts (1,-gamma) gamma "
                                f"log-spaced [1e-4,1e-1] x {R} N(0,1) RHS; refined with FP64 residuals to "
                                f"rel resid <= {args.tol:g}"),
                   "n": n, "shifts": S, "rhs_per_shift": R, "parallelism": "1 GPU"},
        "fp64_solves_per_s": fp64["sps"],
        "fp64_ms_per_step": fp64["ms"],
        "fp32_refine_over_fp64": fp32["sps"] / fp64["sps"],
        "lhl_ms": lhl_ms,
        "max_rel_resid": {"fp32_refine": fp32["rel"], "fp64": fp64["rel"]},
        "refine_rounds": {"fp32": [min(fp32["iters"]), max(fp32["iters"])],
                          "fp64": [min(fp64["iters"]), max(fp64["iters"])]},
        "kernel_ms_per_step_fp32": {k: v[1] for k, v in prof.items()},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["lhl_speedup_vs_cpu"] = cpu["lhl_s"] * 1e3 / lhl_ms["fp64"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
