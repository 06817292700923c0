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
The roofline object describes the hand-written kernel with the largest share
of the step, timed live with HIP events on the library's stream (the rocBLAS
GEMM phases, trsm, trmm and gemm_jx, are listed beside it: each is several
library kernels, and the first pass's forward L sweep runs on a second stream
beside the shift LUs, so its event span includes the time it waits for CUs); the CPU baseline is the
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
FP32_PEAK_TFLOPS = 157.3


def stiff(seed, n):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((n, n))
    u = np.linspace(0.0, 6.0, n)
    return np.asfortranarray(G / np.sqrt(n) - np.diag(10.0 ** u))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


KERNELS = ["shift", "uinv", "ut_pack", "gather", "trsm", "fwd_chain", "utrsm", "trmm", "scatter", "gemm_jx",
           "resid_combine", "colnorms", "convert"]
# phases made of rocBLAS GEMMs (several library kernels each)
LIBRARY_PHASES = {"trsm", "trmm", "gemm_jx"}
PEAK = {"hbm": (HBM_PEAK_GBS, "GB/s"), "mfma32": (FP32_PEAK_TFLOPS, "TFLOP/s"),
        "mfma64": (FP64_PEAK_TFLOPS, "TFLOP/s")}


def config2(lhl_hip, torch, dev):
    """The plain FP64 drop-in on the stiff family, n in {2000, 4096}: lhl(J)
    with J resident in HBM, then for each of 8 real shifts (1, -gamma), gamma
    log-spaced in [1e-4, 1e-1], one call each of lhl_shift!, lhl_ldiv! on one
    N(0,1) right-hand side and one lhl_refine! step with the resident J, the
    vectors in host memory as a Julia caller passes them.  Every call is timed
    on its own (each returns with its result in host memory)."""
    res = {"workload": "stiff J, FP64, 8 real shifts x 1 RHS, host vectors; per-call times (median over shifts)"}
    for n in (2000, 4096):
        Jh = stiff(2000 + n, n)
        J = torch.from_numpy(Jh).to(dev)
        b = np.random.default_rng(n).standard_normal(n)
        ws = lhl_hip.LHLWorkspace(n, np.float64)
        lhl_hip.lhl_(ws, J)
        lhl_hip.sync(ws)
        t0 = time.perf_counter()
        lhl_hip.lhl_(ws, J)
        lhl_hip.sync(ws)
        t_lhl = time.perf_counter() - t0
        ts, tl, tr, rel = [], [], [], 0.0
        for g in np.logspace(-4, -1, 8):
            a = time.perf_counter()
            lhl_hip.lhl_shift_(ws, 1.0, -g)
            ts.append(time.perf_counter() - a)
            x = b.copy()
            a = time.perf_counter()
            lhl_hip.lhl_ldiv_(x, ws)
            tl.append(time.perf_counter() - a)
            a = time.perf_counter()
            lhl_hip.lhl_refine_(x, None, b, ws, 1)
            tr.append(time.perf_counter() - a)
            rel = max(rel, float(np.abs(b - (x - g * (Jh @ x))).max() / np.abs(b).max()))
        res[f"n{n}"] = {"lhl_ms": 1e3 * t_lhl, "shift_ms": 1e3 * float(np.median(ts)),
                        "ldiv_ms": 1e3 * float(np.median(tl)), "refine_step_ms": 1e3 * float(np.median(tr)),
                        "max_rel_resid": rel}
        log(f"bench: config-2 n={n}: lhl {1e3 * t_lhl:.1f} ms, per shift: lhl_shift! {1e3 * np.median(ts):.2f} ms, "
            f"lhl_ldiv! {1e3 * np.median(tl):.2f} ms, lhl_refine! step {1e3 * np.median(tr):.2f} ms, "
            f"max rel resid {rel:.2e}")
        del ws
    return res


def config5(lhl_hip, torch, dev, tol, threads):
    """The uneven batch: 64 independent stiff J (the family above) with n drawn
    uniformly from [128, 3000], each with 4-32 real shifts (1, -gamma), gamma
    log-spaced in [1e-4, 1e-1], and 1-8 N(0,1) right-hand sides per shift
    (seeded).  Grouped: one lhl_reduce_group call (one launch per reduction
    step for each size class), then each matrix's batched solve refined with
    FP64 residuals to `tol`, the solves handed out largest first to `threads`
    host threads (each matrix's handle has its own stream).  For comparison,
    the same work with one lhl() per matrix on the host threads."""
    from concurrent.futures import ThreadPoolExecutor

    rng = np.random.default_rng(505)
    M = 64
    ns = rng.integers(128, 3001, M)
    nsh = rng.integers(4, 33, M)
    nrhs = rng.integers(1, 9, M)
    mats = []
    for m in range(M):
        n, S, R = int(ns[m]), int(nsh[m]), int(nrhs[m])
        J = torch.from_numpy(stiff(5000 + m, n)).to(dev)
        B = torch.from_numpy(np.asfortranarray(rng.standard_normal((n, S * R)))).to(dev)
        assert J.stride(0) == 1 and B.stride(0) == 1
        mats.append(dict(n=n, S=S, R=R, J=J, B=B, gam=np.logspace(-4, -1, S),
                         offs=np.arange(0, S * R + 1, R, dtype=np.int64)))
    order = sorted(range(M), key=lambda m: -(mats[m]["n"] ** 3 + mats[m]["S"] * mats[m]["n"] ** 2))
    ncol = int(sum(d["S"] * d["R"] for d in mats))
    res = {"workload": (f"{M} independent stiff J, n uniform in [128,3000] (sum n = {int(ns.sum())}, max "
                        f"{int(ns.max())}), 4-32 real shifts x 1-8 RHS each ({ncol} solves in all); grouped "
                        f"reductions (lhl_reduce_group) + one batched solve per matrix refined to rel resid <= "
                        f"{tol:g}")}
    Js = [d["J"] for d in mats]
    for name, dt in (("fp64", np.float64), ("fp32", np.float32)):
        wss = [lhl_hip.LHLWorkspace(d["n"], dt) for d in mats]
        Xs = [torch.zeros_like(d["B"]) for d in mats]
        torch.cuda.synchronize()
        its = [None] * M

        def solve(m):
            d = mats[m]
            _, info, it = lhl_hip.shift_solve_batched(wss[m], np.ones(d["S"]), -d["gam"], d["B"], d["offs"],
                                                      X=Xs[m], precision="refine64", tol=tol, max_iter=30)
            assert np.all(info == 0)
            its[m] = it

        def lone(m):
            lhl_hip.lhl_(wss[m], mats[m]["J"])
            solve(m)

        def job(grouped, nt):
            if grouped:
                lhl_hip.lhl_group_(wss, Js)
            f = solve if grouped else lone
            if nt == 1:
                for m in order:
                    f(m)
            else:
                with ThreadPoolExecutor(nt) as ex:
                    list(ex.map(f, order))
            for ws in wss:
                lhl_hip.sync(ws)

        job(True, threads)   # warm-up: the handles' buffers, rocBLAS, code objects
        job(False, threads)
        torch.cuda.synchronize()
        t = {}
        for key, grouped, nt in (("grouped", True, threads), ("grouped_1", True, 1), ("lone", False, threads)):
            t0 = time.perf_counter()
            job(grouped, nt)
            torch.cuda.synchronize()
            t[key] = time.perf_counter() - t0
        # the reductions alone, grouped
        t0 = time.perf_counter()
        lhl_hip.lhl_group_(wss, Js)
        torch.cuda.synchronize()
        t_red = time.perf_counter() - t0
        job(True, threads)       # leave the grouped results in Xs
        torch.cuda.synchronize()
        rel = 0.0
        for m, d in enumerate(mats):
            g = torch.from_numpy(np.repeat(d["gam"], d["R"])).to(dev)
            Rd = d["B"] - (Xs[m] - (d["J"] @ Xs[m]) * g[None, :])
            rel = max(rel, (Rd.abs().amax(0) / d["B"].abs().amax(0)).max().item())
        itv = np.concatenate(its)
        res[name] = {"ms": 1e3 * t["grouped"], "solves_per_s": ncol / t["grouped"], "host_threads": threads,
                     "grouped_reductions_ms": 1e3 * t_red, "ms_one_thread": 1e3 * t["grouped_1"],
                     "per_matrix_lhl_ms": 1e3 * t["lone"], "max_rel_resid": rel,
                     "refine_rounds": [int(itv.min()), int(itv.max())]}
        log(f"bench: config-5 {name}: {1e3 * t['grouped']:.1f} ms grouped ({1e3 * t_red:.1f} ms of it the 64 "
            f"reductions; {1e3 * t['grouped_1']:.1f} ms with one host thread), {1e3 * t['lone']:.1f} ms with one "
            f"lhl() per matrix; {ncol / t['grouped']:.0f} solves/s, max rel resid {rel:.2e}")
        del wss, Xs
    res["fp32_refine_over_fp64"] = res["fp32"]["solves_per_s"] / res["fp64"]["solves_per_s"]
    return res


def lhl_roofline(n, es, ms):
    """lhl() against the delayed-update (blocked) algorithm's bound: a
    read-only GEMV stream of s n^3/3 bytes at the HBM peak plus n^3 flops of
    panel GEMMs at the MFMA peak.  `frac` = that bound / the measured time, so
    it is the fraction of the algorithm's roofline reached whatever the
    kernels' own traffic.  For reference, the unblocked step loop's own bytes
    (each step reads the n x (n-k) trailing columns and writes back the
    (n-k)^2 block: 5/6 s n^3) are reported beside it."""
    unb = es * 5.0 / 6.0 * n ** 3
    blk_b = es * n ** 3 / 3.0
    blk_s = blk_b / (HBM_PEAK_GBS * 1e9) + n ** 3 / ((FP64_PEAK_TFLOPS if es == 8 else FP32_PEAK_TFLOPS) * 1e12)
    return {"bound": "hbm", "algorithmic_bytes": blk_b, "achieved": blk_b / (ms * 1e-3) / 1e9,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": 1e3 * blk_s / ms, "bound_ms": 1e3 * blk_s,
            "unblocked_bytes": unb, "unblocked_frac": unb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}


def algorithmic(kernel, n, S, ncol, es):
    """Algorithmic work of ONE launch of a kernel over the whole batch (S
    shifts, ncol columns, factor element size es): (bytes, flops, bound).
    bytes counts each operand once (what an ideal kernel moves to or from
    HBM); flops counts the multiply-adds as 2."""
    tri = n * (n + 1) / 2
    col = n * ncol * es                      # one pass over the batch's columns
    nb = (n + 63) // 64
    if kernel == "shift":
        # H's upper Hessenberg part read, each shift's packed U, multipliers,
        # reciprocal pivots and swap flags written; n^2/2 multiply-adds per shift
        return (tri + n) * es + S * (tri + 3 * n) * es, S * n * n, "hbm"
    if kernel == "uinv":
        return S * nb * 64 * 64 * es * 2, S * nb * 2 * 64 ** 3 / 3, "hbm"
    if kernel == "ut_pack":
        # every shift's packed U read and written again in packed columns
        return 2 * S * tri * es, 0, "hbm"
    if kernel == "gather":
        return col + ncol * n * 8 + n * 12, 0, "hbm"
    if kernel == "trsm":
        # L' (strict lower triangle) and its diagonal-block inverses read once,
        # the columns read and written: n^2 flops per column
        return n * n / 2 * es + nb * 64 * 64 * es + 2 * col, n * n * ncol, "mfma32" if es == 4 else "mfma64"
    if kernel == "fwd_chain":
        return 2 * col + S * 2 * n * es, 4 * n * ncol, "hbm"
    if kernel == "utrsm":
        # every shift's packed U and diagonal-block inverses read once
        return S * tri * es + S * nb * 64 * 64 * es + 2 * col, n * n * ncol, "hbm"
    if kernel == "trmm":
        # (I + L') y': L' read once, the columns read and written
        return n * n / 2 * es + 2 * col, n * n * ncol, "mfma32" if es == 4 else "mfma64"
    if kernel == "scatter":
        # the swept columns read, the FP64 solution read-modify-written
        return col + 2 * n * ncol * 8, 0, "hbm"
    if kernel == "gemm_jx":
        # J x in FP64 for every column: J (n^2) read once, x read, J x written
        return n * n * 8 + 2 * n * ncol * 8, 2.0 * n * n * ncol, "mfma64"
    if kernel == "resid_combine":
        return 4 * n * ncol * 8, 3 * n * ncol, "hbm"
    if kernel == "colnorms":
        return 2 * n * ncol * 8, 2 * n * ncol, "hbm"
    if kernel == "convert":
        return 2 * col, 0, "hbm"
    return 0, 0, "hbm"


# HBM bytes per full-width launch at n = 4096, 64 shifts x 16 RHS, FP32
# factors: two rocprofv3 counter passes over `bench.py --pmc-workload 3`
# (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch, averaged over the launches),
# each scaled by what the same counter reads in the same run for k_rows_load,
# which reads the FP64 J (134.2 MB) once and writes its row-major copy once:
# FETCH_SIZE counts 0.501 and WRITE_SIZE 1.000 of those bytes.
MEASURED_TRAFFIC_N4096_S64X16_F32 = {
    "shift": 2.725e9, "uinv": 0.118e9, "gather": 0.192e9,
    "fwd_chain": 0.120e9, "utrsm": 2.461e9, "scatter": 0.116e9,
}


def roofline_of(kernel, launches, total_ms, n, S, ncol, es, su=None, cu=None):
    """Roofline figures of one kernel from its live per-launch average.  The
    algorithmic work is priced from what the launches processed: su shifts and
    cu internal columns summed over the launches (every formula above is
    affine in S and ncol, so the average launch's work times the launch count
    is the summed work); without them every launch counts the whole batch."""
    if su is not None and cu is not None and launches:
        S_l, c_l = su / launches, cu / launches
        byts, flops, bound = algorithmic(kernel, n, S_l if su else S, c_l if cu else ncol, es)
    else:
        byts, flops, bound = algorithmic(kernel, n, S, ncol, es)
    avg_s = total_ms / launches / 1e3
    peak, unit = PEAK[bound]
    if bound == "hbm":
        achieved = byts / avg_s / 1e9
    else:
        achieved = flops / avg_s / 1e12
    return {"bound": "hbm" if bound == "hbm" else "mfma", "achieved": achieved, "peak": peak, "unit": unit,
            "frac": achieved / peak,
            "traffic": (MEASURED_TRAFFIC_N4096_S64X16_F32.get(kernel)
                        if (n, S, ncol, es) == (4096, 64, 1024, 4) else None),
            "traffic_basis": "HBM bytes of one full-width launch (64 shifts x 1024 columns)",
            "kernel": kernel, "launches": launches,
            "avg_ms": total_ms / launches, "algorithmic_bytes": byts, "algorithmic_flops": flops,
            "shift_units": su, "col_units": cu}


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
    ap.add_argument("--config4", type=int, default=1,
                    help="1 also times the n=4096 sweep of 256 complex (i*omega, -1) + 256 real shifts x 8 RHS")
    ap.add_argument("--config2", type=int, default=0,
                    help="1 also times the plain FP64 drop-in calls (lhl, lhl_shift!, lhl_ldiv!, lhl_refine!) "
                         "at n = 2000 and 4096")
    ap.add_argument("--config5", type=int, default=8,
                    help="T > 0 also times the uneven batch of 64 independent matrices on T host threads")
    ap.add_argument("--calib-copy", type=int, default=0,
                    help="MiB of a plain device copy run at the end (the HBM-counter calibration for --pmc runs)")
    ap.add_argument("--pmc-workload", type=int, default=0,
                    help="K > 0: only lhl() once and K + 1 plain batched solves per precision, then exit (the "
                         "short program the rocprofv3 counter passes run)")
    ap.add_argument("--plan", type=int, default=1, help="0 skips the graph-captured (plan) leg of the main step")
    ap.add_argument("--tune", action="append", default=[], metavar="KNOB=VALUE",
                    help="set a tuning knob on the main workload's workspaces (repeatable)")
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
    for kv in args.tune:
        k, v = kv.split("=")
        for w_ in (ws64, ws32):
            lhl_hip.tune_set(w_, k, int(v))
    if args.pmc_workload:
        # every kernel of a full-width solve pass, a few times, and nothing else
        # (k_rows_load in it reads the FP64 J once and writes the row-major
        # copy once: the counters' calibration in the same run)
        for name, ws in (("fp64", ws64), ("fp32", ws32)):
            lhl_hip.lhl_(ws, Jd)
            Xd = torch.zeros_like(Bd)
            for _ in range(args.pmc_workload + 1):
                lhl_hip.shift_solve_batched(ws, sig, tau, Bd, offs, X=Xd)
            lhl_hip.sync(ws)
            log(f"bench: pmc workload {name} done ({args.pmc_workload + 1} solve calls)")
        return
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
    # one more lhl() per precision with every launch timed (outside the timed runs)
    lhl_kernels = {}
    for name, ws in (("fp64", ws64), ("fp32", ws32)):
        lhl_hip.profile_reset(ws)
        lhl_hip.profile_enable(ws, True)
        lhl_hip.lhl_(ws, Jd)
        lhl_hip.sync(ws)
        lhl_hip.profile_enable(ws, False)
        lhl_kernels[name] = {}
        for k in ("balance", "reduce_copy", "reduce_step", "reduce_panel", "reduce_gemm", "reduce_top", "reduce_pivot",
                  "reduce_swap", "reduce_trailing", "reduce_finish"):
            c, t = lhl_hip.profile_read(ws, k)
            if c:
                lhl_kernels[name][k] = {"launches": c, "total_ms": t, "avg_us": 1e3 * t / c}
        log(f"bench: lhl {name} kernels " + ", ".join(f"{k} {v['launches']}x {v['avg_us']:.1f} us"
                                                    for k, v in lhl_kernels[name].items()))

    def step(ws):
        _, info, iters = lhl_hip.shift_solve_batched(ws, sig, tau, Bd, offs, X=Xd, precision="refine64",
                                                     tol=args.tol, max_iter=30)
        return info, iters

    results = {}
    prof = {}
    for name, ws in (("fp64", ws64), ("fp32", ws32)):
        for _ in range(args.warmup):
            step(ws)
        torch.cuda.synchronize()
        # the timed steps run without per-launch events
        t0 = time.perf_counter()
        for _ in range(args.steps):
            info, iters = step(ws)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / max(args.steps, 1)
        # the kernel breakdown: the same steps again, every launch bracketed by
        # HIP events on the library's stream (the stream the kernels run on)
        lhl_hip.profile_reset(ws)
        lhl_hip.profile_enable(ws, True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step(ws)
        torch.cuda.synchronize()
        dt_prof = (time.perf_counter() - t0) / max(args.steps, 1)
        lhl_hip.profile_enable(ws, False)
        prof[name] = {k: lhl_hip.profile_read_work(ws, k) for k in KERNELS}
        # verify the accuracy bar on the device in FP64: r = b - (x - gamma J x)
        JX = Jd @ Xd
        g = torch.from_numpy(np.repeat(gam, R)).to(dev)
        Rd = Bd - (Xd - JX * g[None, :])
        rel = (Rd.abs().amax(0) / Bd.abs().amax(0)).max().item()
        results[name] = dict(ms=1e3 * dt, sps=ncol / dt, iters=iters.tolist(), info=int(np.max(np.abs(info))),
                             rel=rel, ms_profiled=1e3 * dt_prof)
        log(f"bench: {name} step {1e3 * dt:.2f} ms, {ncol / dt:.0f} solves/s, max rel resid {rel:.2e}, "
            f"refine iters {int(iters.min())}-{int(iters.max())}")
        if name == "fp64":
            X64 = Xd.clone()
            # the plain FP64 path: one solve pass, no refinement
            for _ in range(args.warmup):
                lhl_hip.shift_solve_batched(ws, sig, tau, Bd, offs, X=Xd)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                lhl_hip.shift_solve_batched(ws, sig, tau, Bd, offs, X=Xd)
            torch.cuda.synchronize()
            dtp = (time.perf_counter() - t0) / max(args.steps, 1)
            Rd = Bd - (Xd - (Jd @ Xd) * g[None, :])
            relp = (Rd.abs().amax(0) / Bd.abs().amax(0)).max().item()
            results["fp64_plain"] = dict(ms=1e3 * dtp, sps=ncol / dtp, rel=relp)
            log(f"bench: fp64 plain (no refinement) step {1e3 * dtp:.2f} ms, {ncol / dtp:.0f} solves/s, "
                f"max rel resid {relp:.2e}")
        else:
            # each shift's error against the FP64 solution (the FP64 path is
            # bit-checked against the oracle at sizes the oracle finishes):
            # max over its RHS of ||x32 - x64||_inf / ||x64||_inf
            e = ((Xd - X64).abs().amax(0) / X64.abs().amax(0)).reshape(S, R).amax(1).cpu().numpy()
            results[name]["err_vs_fp64"] = [float(f"{v:.3g}") for v in e]
            log(f"bench: fp32+refine error vs fp64 per shift: max {e.max():.2e}, median {np.median(e):.2e}")

    # ---- the same step as one graph launch (a plan): device-side refinement
    # control, replayed for the same shifts (learning run and capture untimed)
    plan_res = {}
    if args.plan:
        g = torch.from_numpy(np.repeat(gam, R)).to(dev)
        for name, ws in (("fp64", ws64), ("fp32", ws32)):
            Xp = torch.zeros_like(Bd)
            plan = lhl_hip.BatchPlan(ws, offs, Bd, Xp, precision="refine64", tol=args.tol, max_iter=30)
            for _ in range(max(args.warmup, 0) + 2):
                plan.run(sig, tau)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                info, iters = plan.run(sig, tau)
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / max(args.steps, 1)
            Rd = Bd - (Xp - (Jd @ Xp) * g[None, :])
            rel = (Rd.abs().amax(0) / Bd.abs().amax(0)).max().item()
            gr, replays, captures, extra = plan.stats()
            plan_res[name] = {"ms": 1e3 * dt, "solves_per_s": ncol / dt, "max_rel_resid": rel,
                              "refine_rounds": [int(iters.min()), int(iters.max())], "graph_rounds": gr,
                              "replays": replays, "captures": captures, "rounds_beyond_graph": extra}
            log(f"bench: plan {name} step {1e3 * dt:.2f} ms, {ncol / dt:.0f} solves/s, max rel resid {rel:.2e}, "
                f"graph rounds {gr}, replays {replays}")
            plan.close()
            del Xp
        plan_res["fp32_refine_over_fp64"] = plan_res["fp32"]["solves_per_s"] / plan_res["fp64"]["solves_per_s"]

    # ---- optional: the complex + real shift sweep on one real reduction
    sweep4 = None
    if args.config4:
        S4, R4 = 256, 8
        rng4 = np.random.default_rng(404)
        A4 = np.asfortranarray(rng4.standard_normal((n, n)) / np.sqrt(n) - 2.0 * np.eye(n))
        A4d = torch.from_numpy(A4).to(dev)
        om = np.logspace(-2, 2, S4)
        gam4 = np.logspace(-4, -1, S4)
        offs4 = np.arange(0, S4 * R4 + 1, R4, dtype=np.int64)
        Bc = torch.from_numpy(np.asfortranarray(rng4.standard_normal((n, S4 * R4)) +
                                                1j * rng4.standard_normal((n, S4 * R4)))).to(dev)
        Br = torch.from_numpy(np.asfortranarray(rng4.standard_normal((n, S4 * R4)))).to(dev)
        Xc, Xr = torch.zeros_like(Bc), torch.zeros_like(Br)
        sweep4 = {"workload": f"n={n} A=G/sqrt(n)-2I, {S4} complex shifts (i*omega,-1) omega log-spaced [1e-2,1e2] "
                              f"+ {S4} real shifts (1,-gamma) gamma log-spaced [1e-4,1e-1], {R4} RHS each"}
        for name, dt in (("fp64", np.float64), ("fp32", np.float32)):
            ws4 = lhl_hip.LHLWorkspace(n, dt)
            lhl_hip.lhl_(ws4, A4d)
            prec = "refine64"   # both precisions refined to the same tolerance, as in the main step

            # the sweep as two graph launches: plans for the complex and the
            # real shifts (learning run and capture untimed)
            pc = lhl_hip.BatchPlan(ws4, offs4, Bc, Xc, shifts_complex=True, precision=prec, tol=args.tol,
                                   max_iter=30)
            pr = lhl_hip.BatchPlan(ws4, offs4, Br, Xr, precision=prec, tol=args.tol, max_iter=30)

            def sweep_plan():
                _, itc = pc.run(1j * om, -np.ones(S4))
                _, itr = pr.run(np.ones(S4), -gam4)
                return np.concatenate([itc, itr])
            for _ in range(3):
                sweep_plan()
            torch.cuda.synchronize()
            reps4 = 3
            t0 = time.perf_counter()
            for _ in range(reps4):
                its = sweep_plan()
            torch.cuda.synchronize()
            dt4 = (time.perf_counter() - t0) / reps4
            plan_stats = {"complex": pc.stats(), "real": pr.stats()}
            pc.close()
            pr.close()
            torch.cuda.synchronize()
            # accuracy on the device: r = b - (sigma x + tau A x)
            sig = torch.from_numpy(np.repeat(1j * om, R4)).to(dev)
            Rc = Bc - (Xc * sig[None, :] - A4d.to(torch.complex128) @ Xc)
            relc = (Rc.abs().amax(0) / Bc.abs().amax(0)).max().item()
            g4 = torch.from_numpy(np.repeat(gam4, R4)).to(dev)
            Rr = Br - (Xr - (A4d @ Xr) * g4[None, :])
            relr = (Rr.abs().amax(0) / Br.abs().amax(0)).max().item()
            sweep4[name] = {"solves_per_s": 2 * S4 * R4 / dt4, "ms": 1e3 * dt4, "max_rel_resid": max(relc, relr),
                            "refine_rounds": [int(its.min()), int(its.max())], "path": "plans (HIP graphs)",
                            "plan_stats_graph_rounds_replays_captures_extra": plan_stats}
            log(f"bench: config-4 sweep {name}: {1e3 * dt4:.1f} ms as two graph launches, "
                f"{2 * S4 * R4 / dt4:.0f} solves/s, max rel resid {max(relc, relr):.2e}")
            del ws4

    # ---- optional: the plain drop-in calls, one shift at a time
    sweep2 = config2(lhl_hip, torch, dev) if args.config2 else None

    # ---- optional: the uneven batch of independent matrices
    sweep5 = config5(lhl_hip, torch, dev, args.tol, args.config5) if args.config5 else None

    # ---- roofline of every kernel of the FP32+refine step; the dominant one leads
    steps = max(args.steps, 1)
    kernels = {}
    for prec, es in (("fp32", 4), ("fp64", 8)):
        kernels[prec] = {}
        for k, (c, t, su, cu) in prof[prec].items():
            if c:
                r = roofline_of(k, c, t, n, S, ncol, es, su, cu)
                r["ms_per_step"] = t / steps
                kernels[prec][k] = r
    log("bench: fp32 kernels " + ", ".join(f"{k} {v['launches'] // steps}x {v['ms_per_step']:.2f} ms"
                                          for k, v in kernels["fp32"].items()))
    own = [k for k in kernels["fp32"] if k not in LIBRARY_PHASES]
    dom = max(own, key=lambda k: kernels["fp32"][k]["ms_per_step"])
    roof = dict(kernels["fp32"][dom])

    # ---- CPU baseline: the C restatement of the reference algorithm (oracle/),
    # on a bounded sample: lhl() of an n/2 matrix of the same family scaled by
    # n^3 (the reduction is 5/3 n^3 flops), and shifts/solves at full n timed on
    # the GPU's own factors (their cost does not depend on the values).
    cpu = None
    if args.cpu_baseline:
        from oracle import oracle as O

        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        os.environ.setdefault("OMP_NUM_THREADS", str(cores))
        ns_red = max(n // 2, 2)
        log(f"bench: CPU baseline (oracle, {cores} threads for the reduction, lhl sample n={ns_red})")
        Js = stiff(4096, ns_red)
        t0 = time.perf_counter()
        O.reduce(Js)
        t_red_s = time.perf_counter() - t0
        t_red = t_red_s * (n / ns_red) ** 3
        i32 = lambda v: np.ascontiguousarray(np.asarray(v) - 1, dtype=np.int32)
        red = O.Reduction(np.asfortranarray(ws64.factors, dtype=np.float64), i32(ws64.ipiv), i32(ws64.perm),
                          i32(ws64.iperm), np.ascontiguousarray(ws64.scale, dtype=np.float64),
                          np.ascontiguousarray(ws64.iscale, dtype=np.float64),
                          np.asfortranarray(ws64.Ht, dtype=np.float64))
        # four shifts spread over gamma, one right-hand side each: the oracle's
        # FP64 shift LU and solve (timed), and each GPU solution's error
        # against it (max |x - x_oracle| / max |x_oracle|)
        pick = sorted({0, S // 3, (2 * S) // 3, S - 1})
        t_shift = t_solve = 0.0
        X32h, X64h = Xd.cpu().numpy(), X64.cpu().numpy()
        err_oracle = {"shifts": pick, "fp32_refine": [], "fp64": []}
        for s_ in pick:
            a = time.perf_counter()
            sh = O.shift(red.Ht, 1.0, float(tau[s_]))
            t_shift += time.perf_counter() - a
            a = time.perf_counter()
            xo = O.ldiv(red, sh, Bh[:, s_ * R])
            t_solve += time.perf_counter() - a
            den = np.abs(xo).max()
            err_oracle["fp32_refine"].append(float(f"{np.abs(X32h[:, s_ * R] - xo).max() / den:.3g}"))
            err_oracle["fp64"].append(float(f"{np.abs(X64h[:, s_ * R] - xo).max() / den:.3g}"))
        per_shift = t_shift / len(pick)
        per_solve = t_solve / len(pick)
        log(f"bench: error vs the oracle's FP64 solves (shifts {pick}): fp32+refine {err_oracle['fp32_refine']}, "
            f"fp64 {err_oracle['fp64']}")
        # one FP64 refinement round per solve doubles the solve count; the residual is a dense matvec
        a = time.perf_counter()
        _ = J @ Bh[:, 0]
        t_mv = time.perf_counter() - a
        batch_s = S * per_shift + ncol * (2 * per_solve + t_mv)
        # the reference's own published single-thread reduction time (112 ms at
        # n=1024, README.md:208-216) scaled by n^3, for scale
        ref_1t = 0.112 * (n / 1024) ** 3
        cpu = {"value": ncol / batch_s, "unit": "solves/s", "cores": cores, "kind": "port",
               "sample": (f"oracle (C restatement of the reference, not the Julia reference): lhl() of a stiff "
                          f"n={ns_red} J ({t_red_s:.2f} s, OpenMP {cores} threads) scaled by (n/{ns_red})^3 "
                          f"to {t_red:.1f} s; {len(pick)} shifts x 1 solve at n={n} timed serially "
                          f"({1e3 * per_shift:.1f} ms/shift, {1e3 * per_solve:.1f} ms/solve) and scaled to the "
                          f"{S}x{R} batch with one refinement round per solve"),
               "lhl_s": t_red,
               "reference_published_lhl_1thread_s_scaled": ref_1t,
               "err_vs_oracle_fp64_solve": err_oracle}
        log(f"bench: CPU lhl {t_red:.2f} s (sample {t_red_s:.2f} s), {ncol / batch_s:.2f} solves/s")

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
        "fp64_plain": {"solves_per_s": results["fp64_plain"]["sps"], "ms_per_step": results["fp64_plain"]["ms"],
                       "max_rel_resid": results["fp64_plain"]["rel"],
                       "note": "one FP64 solve pass, no refinement (the residual bar is not enforced here)"},
        "fp32_refine_over_fp64_plain": fp32["sps"] / results["fp64_plain"]["sps"],
        "ms_per_step_with_launch_events": {"fp32": fp32["ms_profiled"], "fp64": fp64["ms_profiled"]},
        "lhl_ms": lhl_ms,
        "lhl_kernel_ms": lhl_kernels,
        "lhl_roofline": {p: lhl_roofline(n, 8 if p == "fp64" else 4, lhl_ms[p]) for p in lhl_ms},
        "max_rel_resid": {"fp32_refine": fp32["rel"], "fp64": fp64["rel"]},
        "fp32_refine_err_vs_fp64_per_shift": fp32["err_vs_fp64"],
        "refine_rounds": {"fp32": [min(fp32["iters"]), max(fp32["iters"])],
                          "fp64": [min(fp64["iters"]), max(fp64["iters"])]},
        "kernel_ms_per_step": {p: {k: v["ms_per_step"] for k, v in kernels[p].items()} for p in kernels},
        "roofline": roof,
        "rooflines_fp32": {k: {f: v[f] for f in ("bound", "achieved", "unit", "frac", "avg_ms", "traffic",
                                                 "shift_units", "col_units")}
                           for k, v in kernels["fp32"].items()},
        "cpu_baseline": cpu,
    }
    if plan_res:
        out["plan"] = plan_res
    if sweep4:
        out["config4_sweep"] = sweep4
    if sweep2:
        out["config2_dropin"] = sweep2
    if sweep5:
        out["config5_uneven"] = sweep5
    if cpu:
        # The comparators of lhl() that mean something: the reference's own
        # published single-thread time scaled by n^3, and that time divided by
        # its published 16-thread speedup at n = 2000 (8.4x, README.md:126-130).
        # (The C port's own time is in cpu_baseline; it is a plain restatement,
        # about 17x slower than the reference's tuned code at n = 1024.)
        ref1 = cpu["reference_published_lhl_1thread_s_scaled"] * 1e3
        out["lhl_speedup_vs_reference_published_1thread"] = ref1 / lhl_ms["fp64"]
        out["lhl_speedup_vs_reference_16thread_estimate"] = ref1 / 8.4 / lhl_ms["fp64"]
    print(json.dumps(out), flush=True)
    if args.calib_copy > 0:
        # a plain copy of known size: FETCH_SIZE / WRITE_SIZE of this dispatch
        # against its exact byte count calibrates the counters of the kernels above
        nb = args.calib_copy * 1024 * 1024
        src = torch.ones(nb // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"bench: calibration copy of {args.calib_copy} MiB ({nb / 1024:.0f} KiB read, {nb / 1024:.0f} KiB written)")


if __name__ == "__main__":
    main()
