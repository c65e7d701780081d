"""Benchmark of the linear-solve hot path on one MI355X.

Workload (default): the seeded 60x220x85 black-oil-like Jacobian (1,122,000 cells, BSR 3x3 FP64, five wells with
BHP rows Schur-eliminated), solved to relative residual 1e-8.  One step = one Newton-iteration solve with every
input already resident in HBM: new matrix values handed over on the device (D2D), numeric preconditioner set-up
(CPR quasi-IMPES weights, pressure matrix, Galerkin refresh of the resident AMG hierarchy, coarsest inverse,
BILU0 factorisation) and the Krylov solve.  The symbolic set-up (levels, slices, aggregation) happens once per
pattern, before the timed region, as in the Newton loop.

Printed: one JSON line.  value = solves/s of the FP64 CPR-BiCGSTAB drop-in (variant `cpr_mc`: quasi-IMPES
weights, aggregation AMG V(1,1) with damped-Jacobi smoothing, multicolour BILU0 second stage); beside it every other
variant (natural-order BILU0 second stage, plain FP64 ILU0-BiCGSTAB, mixed precision; --extra adds the exact
Gauss-Seidel AMG, which reproduces the reference CPU arithmetic and is slow), the roofline of the dominant kernel and
the CPU baseline (the reference's own solver code compiled from its sources, OpenMP, timed on a bounded sample on
this host).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class DeviceInputs:
    """Matrix values, right-hand side and solution kept in HBM."""

    def __init__(self, rs, ctx, s):
        self.rs, self.ctx = rs, ctx
        L = rs.lib()
        self.ptrs = []

        def up(a):
            a = np.ascontiguousarray(a)
            p = ctypes.c_void_p()
            rs._check(L.rsh_device_alloc(ctx.h, a.nbytes, ctypes.byref(p)), "alloc")
            rs._check(L.rsh_memcpy_h2d(ctx.h, p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes), "h2d")
            self.ptrs.append(p)
            return p

        self.val = up(np.asarray(s.val, np.float64).reshape(-1))
        self.b = up(np.asarray(s.rhs, np.float64))
        self.x = up(np.zeros(s.n * s.nb))
        self.N = s.n * s.nb

    def x_host(self):
        out = np.empty(self.N)
        self.rs._check(self.rs.lib().rsh_memcpy_d2h(self.ctx.h, out.ctypes.data_as(ctypes.c_void_p), self.x, out.nbytes), "d2h")
        return out

    def free(self):
        for p in self.ptrs:
            self.rs.lib().rsh_device_free(self.ctx.h, p)
        self.ptrs = []


def step(rs, sysd, dev, opts, tol, max_it):
    L = rs.lib()
    rs._check(L.rsh_system_set_values_device(sysd.h, dev.val), "set_values_device")
    rs._check(L.rsh_system_setup(sysd.h, ctypes.byref(opts)), "setup")
    its = ctypes.c_int(0)
    rel = ctypes.c_double(0)
    rc = L.rsh_solve_device(sysd.h, dev.b, dev.x, tol, max_it, ctypes.byref(its), ctypes.byref(rel))
    rs._check(rc, "solve")
    return int(its.value), float(rel.value)


def time_config(rs, sysd, dev, opts, tol, max_it, steps, warmup):
    for _ in range(warmup):
        step(rs, sysd, dev, opts, tol, max_it)
    sysd.ctx.synchronize()
    t0 = time.perf_counter()
    its = rel = None
    for _ in range(steps):
        its, rel = step(rs, sysd, dev, opts, tol, max_it)
    sysd.ctx.synchronize()
    dt = (time.perf_counter() - t0) / steps
    info = sysd.info()
    return dt, its, rel, info


def cpu_baseline(s, tol, lean=False):
    """The reference's BlockCPR + BiCGSTAB (BlockCPR::setup + bicgstab, bsr.cpp:993-1084, linsolve.cpp:159-216)
    compiled from its sources with its release flags (OpenMP, -O3 -mavx2 -mfma), on this host's cores, solving the
    same system to the same tolerance.  value: the Newton loop's steady state, the work a GPU step does: the CPR set-up
    with the AMG aggregates of an earlier set-up reused (AMG::refresh, amg.cpp:257-271: weights, pressure matrix,
    Galerkin products, coarsest factorisation, BILU0) + the solve.  fresh: a complete set-up (aggregation included) +
    the solve, as the first Newton iteration of a pattern.  About 40 s of CPU work in all.  lean: the fresh solve with
    the BILU0 stage factorised lean (BlockCPR::setLean, the engine's default from 500k cells)."""
    import oracle
    ref = oracle.reference(fast=True)
    kind = "reference"
    if ref is None:
        ref = oracle.oracle()
        kind = "port"
    threads = ref.set_threads(0)
    _, its, rel, conv, setup_s, solve_s = ref.krylov_cpr_steady(s, s.rhs, tol, 3000)
    dt = setup_s + solve_s
    t0 = time.perf_counter()
    _, its_f, rel_f, conv_f = ref.krylov(s, s.rhs, "cpr", "bicgstab", tol, 3000)
    dt_f = time.perf_counter() - t0
    out = {"value": 1.0 / dt, "unit": "solves/s", "cores": int(threads), "kind": kind,
           "sample": f"one CPR-BiCGSTAB step of the {s.n}-cell system to {tol:g} in the Newton loop's steady state (CPR set-up "
                     f"with the AMG aggregates reused {setup_s:.2f} s + solve {solve_s:.1f} s, {its} iterations, rel.res {rel:.2e}, "
                     f"converged {conv})",
           "iterations": its,
           "fresh": {"value": 1.0 / dt_f, "unit": "solves/s", "iterations": its_f, "rel_res": rel_f, "converged": bool(conv_f),
                     "sample": f"complete set-up (aggregation included) + solve: {dt_f:.1f} s"}}
    if lean:
        t0 = time.perf_counter()
        _, its, rel, conv = ref.krylov(s, s.rhs, "cpr_lean", "bicgstab", tol, 3000)
        dt = time.perf_counter() - t0
        out["lean"] = {"value": 1.0 / dt, "unit": "solves/s", "iterations": its, "rel_res": rel, "converged": bool(conv),
                       "sample": f"the fresh solve with the lean BILU0 stage: {dt:.1f} s"}
    return out


def error_vs_reference(s, solutions):
    """Each variant's solution against the reference's own FP64 CPR-BiCGSTAB solution of the same system to 1e-10
    (compiled from its sources, oracle/_ref; the oracle's restatement where it was not built): relative 2-norm error.
    The checker's solve is not timed."""
    import oracle
    ref = oracle.reference(fast=True) or oracle.oracle()
    ref.set_threads(0)
    t0 = time.perf_counter()
    xr, its, rel, conv = ref.krylov(s, s.rhs, "cpr", "bicgstab", 1e-10, 3000)
    if not conv:
        raise RuntimeError("the reference solve did not reach 1e-10")
    nr = float(np.linalg.norm(xr))
    return {"sample": f"{its} iterations to rel.res {rel:.1e} in {time.perf_counter() - t0:.1f} s",
            "errors": {k: float(np.linalg.norm(x - xr) / nr) for k, x in solutions.items()}}


def _upload(rs, ctx, a, ptrs):
    a = np.ascontiguousarray(a)
    p = ctypes.c_void_p()
    rs._check(rs.lib().rsh_device_alloc(ctx.h, max(a.nbytes, 8), ctypes.byref(p)), "alloc")
    rs._check(rs.lib().rsh_memcpy_h2d(ctx.h, p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes), "h2d")
    ptrs.append(p)
    return p


def batched_leg(rs, gen, args):
    """256 independent systems (block sizes 2/3/4, 432..45k cells) in one batch: one step = new values (D2D) +
    numeric setup + the batched solve of every system to 1e-8, inputs resident in HBM; against the single-system
    path looped over the same systems (same preconditioner, resident inputs)."""
    L = rs.lib()
    t = time.perf_counter()
    systems = gen.batch_workload(args.batch_count)
    cells = int(sum(s.n for s in systems))
    log(f"batched workload: {len(systems)} systems, {cells} cells, generated in {time.perf_counter() - t:.1f}s")
    ctx = rs.default_context()
    ptrs = []
    val = _upload(rs, ctx, np.concatenate([np.asarray(s.val, np.float64).reshape(-1) for s in systems]), ptrs)
    rhs = _upload(rs, ctx, np.concatenate([s.rhs for s in systems]), ptrs)
    N = int(sum(s.n * s.nb for s in systems))
    xd = _upload(rs, ctx, np.zeros(N), ptrs)
    m = len(systems)
    tol = np.full(m, args.tol)
    its = np.zeros(m, np.int32)
    rel = np.zeros(m)
    st = np.zeros(m, np.int32)
    t = time.perf_counter()
    B = rs.Batch(systems, ctx)
    log(f"batch created in {time.perf_counter() - t:.1f}s")
    out = {"workload": f"{m} systems, grid dims log-uniform in [10x10x3, 64x64x16], block size 2/3/4 in ratio 2:5:1, "
                       f"sigma ln k in [0.5, 2.5], 1-2 wells each; {cells} cells in total", "systems": m, "cells": cells}

    def bstep(o):
        rs._check(L.rsh_batch_set_values_device(B.h, val), "batch set_values")
        rs._check(L.rsh_batch_setup(B.h, ctypes.byref(o)), "batch setup")
        rc = L.rsh_batch_solve_device(B.h, rhs, xd, rs._p(tol, ctypes.c_double), args.max_it, rs._p(its, ctypes.c_int),
                                      rs._p(rel, ctypes.c_double), rs._p(st, ctypes.c_int))
        rs._check(rc, "batch solve")

    xs = {}
    legs = (("batched_mixed", "mcbilu0", rs.MIXED), ("batched_store32", "mcbilu0", rs.MIXED_STORE), ("batched_fp64", "mcbilu0", rs.FP64),
            ("batched_cpr_fp64", "cpr_mc", rs.FP64), ("batched_cpr_store32", "cpr_mc", rs.MIXED_STORE))
    for name, pc, prec in legs:
        o = rs.default_options(precond=pc, precision=prec, smoother=rs.SMOOTH_JACOBI, smooth_steps=2, jacobi_omega=1.0)
        t = time.perf_counter()
        bstep(o)                      # symbolic set-up and tuning (once per pattern) outside the timed region
        log(f"{name}: first step (symbolic + tuning) {time.perf_counter() - t:.1f}s")
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.batch_steps):
            bstep(o)
        ctx.synchronize()
        dt = (time.perf_counter() - t0) / args.batch_steps
        lat = B.latencies()
        info = B.info()
        x = np.empty(N)
        rs._check(L.rsh_memcpy_d2h(ctx.h, x.ctypes.data_as(ctypes.c_void_p), xd, x.nbytes), "d2h")
        xs[name] = x
        out[name] = {"solves_per_s": m / dt, "ms_per_batch": dt * 1e3, "setup_ms": info["last_setup_ms"], "solve_ms": info["last_solve_ms"],
                     "latency_p50_ms": float(np.percentile(lat, 50)), "latency_p99_ms": float(np.percentile(lat, 99)),
                     "iterations_mean": float(its.mean()), "iterations_max": int(its.max()), "sweeps": info["last_sweeps"],
                     "outer": info["last_outer"], "max_rel_res": float(rel.max()), "all_converged": bool((st == 0).all()),
                     "chunk": info["chunk"], "fuse_max": [g.get("fuse_max") for g in info["groups"]],
                     "run_rows": [g.get("run_rows") for g in info["groups"]],
                     "spmv_lanes": [g.get("spmv_lanes") for g in info["groups"]],
                     "spmv_lanes_ms": [g.get("spmv_lanes_ms") for g in info["groups"]],
                     "amg": [[g.get("amg_levels"), g.get("amg_coarsest")] for g in info["groups"]] if pc.startswith("cpr") else None}
        log(f"{name}: {m / dt:.1f} solves/s ({dt * 1e3:.1f} ms per batch: setup {info['last_setup_ms']:.1f} + solve {info['last_solve_ms']:.1f}), "
            f"latency p50 {out[name]['latency_p50_ms']:.1f} ms p99 {out[name]['latency_p99_ms']:.1f} ms, its mean {its.mean():.0f} max {its.max()}, "
            f"SpMV lanes per row {out[name]['spmv_lanes']} (tuner ms for 1/2/4: {out[name]['spmv_lanes_ms']})")
    B.close()

    def loop_leg(name, o, ref):
        """the single-system path looped over the same systems (resident inputs, same options)"""
        sysd, devs = [], []
        t = time.perf_counter()
        for s in systems:
            d = rs.System(s, ctx)
            dv = DeviceInputs(rs, ctx, s)
            step(rs, d, dv, o, args.tol, args.max_it)
            sysd.append(d)
            devs.append(dv)
        log(f"{name}: symbolic set-up of {m} systems {time.perf_counter() - t:.1f}s")
        ctx.synchronize()
        t0 = time.perf_counter()
        lits = []
        for d, dv in zip(sysd, devs):
            k, _ = step(rs, d, dv, o, args.tol, args.max_it)
            lits.append(k)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        err = 0.0
        off = 0
        for s, dv in zip(systems, devs):
            xl = dv.x_host()
            n = s.n * s.nb
            err = max(err, float(np.linalg.norm(xs[ref][off:off + n] - xl) / max(np.linalg.norm(xl), 1e-300)))
            off += n
        out[name] = {"solves_per_s": m / dt, "ms_total": dt * 1e3, "iterations_mean": float(np.mean(lits)),
                     f"max_rel_diff_{ref}_vs_loop": err}
        log(f"{name}: {m / dt:.1f} solves/s ({dt * 1e3:.0f} ms for {m} systems), {ref} vs loop max rel diff {err:.2e}")
        for d, dv in zip(sysd, devs):
            dv.free()
            d.close()

    loop_leg("loop_fp64", rs.default_options(precond="mcbilu0", precision=rs.FP64), "batched_fp64")
    loop_leg("loop_cpr_fp64", rs.default_options(precond="cpr_mc", precision=rs.FP64, smoother=rs.SMOOTH_JACOBI, smooth_steps=2,
                                                 jacobi_omega=1.0), "batched_cpr_fp64")
    out["speedup_mixed_vs_loop"] = out["batched_mixed"]["solves_per_s"] / out["loop_fp64"]["solves_per_s"]
    out["speedup_store32_vs_loop"] = out["batched_store32"]["solves_per_s"] / out["loop_fp64"]["solves_per_s"]
    out["speedup_fp64_vs_loop"] = out["batched_fp64"]["solves_per_s"] / out["loop_fp64"]["solves_per_s"]
    out["speedup_cpr_fp64_vs_loop_cpr"] = out["batched_cpr_fp64"]["solves_per_s"] / out["loop_cpr_fp64"]["solves_per_s"]
    out["speedup_cpr_store32_vs_loop_cpr"] = out["batched_cpr_store32"]["solves_per_s"] / out["loop_cpr_fp64"]["solves_per_s"]
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_sample(systems, 16, "batched")
    for p in ptrs:
        L.rsh_device_free(ctx.h, p)
    return out


def serving_leg(rs, gen, args):
    """The serving loop under Poisson arrivals: 64 resident models (5k-500k cells, log-uniform), each request one
    Newton iteration of a random model (every block row scaled by +-5 % on the device from the model's base values,
    new numeric setup with the resident AMG aggregates refreshed, CPR-BiCGSTAB to 1e-8); latency = submit -> done."""
    t = time.perf_counter()
    models = gen.serving_models(args.serve_models)
    cells = int(sum(m.n for m in models))
    log(f"serving workload: {len(models)} models, {cells} cells, generated in {time.perf_counter() - t:.1f}s")
    lambdas = [float(x) for x in args.serve_lambdas.split(",") if x]
    out = {"workload": f"{len(models)} resident models, 5k-500k cells log-uniform ({cells} cells), nb 3, CPR(Jacobi AMG)+MC-BILU0 "
                       f"BiCGSTAB to 1e-8, block rows scaled by +-5 % per request", "models": len(models), "cells": cells,
           "slots": args.serve_slots, "requests_per_rate": args.serve_requests, "hbm_budget_GB": 200}
    forms = (("fp64", rs.FP64, False), ("mixed", rs.MIXED, False), ("store32", rs.MIXED_STORE, False), ("fp64_batched", rs.FP64, True),
             ("store32_batched", rs.MIXED_STORE, True))
    only = [f for f in args.serve_forms.split(",") if f] or ["fp64", "mixed", "store32", "fp64_batched"]
    for name, prec, batched in forms:
        if only and name not in only:
            continue
        # batched: the requests queued for one group of models (at most serve_group_cells cells) share each round
        srv = rs.Server(slots=args.serve_slots, hbm_budget=200e9, batched=batched, group_cells=args.serve_group_cells, group_systems=16)
        ids = [srv.add_model(m, precond="cpr_mc", smoother=rs.SMOOTH_JACOBI, smooth_steps=2, jacobi_omega=1.0, precision=prec)
               for m in models]
        t = time.perf_counter()
        warm = [srv.submit(i, m.rhs, perturb=0.05, seed=1, tol=args.tol, max_it=args.max_it, want_x=False) for i, m in zip(ids, models)]
        for tk in warm:
            srv.wait(tk)
        log(f"serving {name}: models loaded (symbolic set-up + first iteration) in {time.perf_counter() - t:.1f}s; {srv.stats()}")
        newton = [2] * len(models)
        rows = []
        rng = np.random.default_rng(20240606)    # every form serves the same arrival times and models
        for lam in lambdas:
            gaps = rng.exponential(1.0 / lam, args.serve_requests)
            picks = rng.integers(0, len(models), args.serve_requests)
            tickets, sent = [], []
            t0 = time.perf_counter()
            due = t0
            for g, k in zip(gaps, picks):
                due += g
                d = due - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                seed = 1000 * int(k) + newton[k]
                tickets.append(srv.submit(ids[k], models[k].rhs, perturb=0.05, seed=seed, tol=args.tol, max_it=args.max_it, want_x=False))
                sent.append((int(k), seed))
                newton[k] = 2 + (newton[k] - 1) % 11      # 4-12 Newton iterations per step: cycle through the seeds
            res = [srv.wait(tk) for tk in tickets]
            t1 = time.perf_counter()
            lat = np.array([r["latency_ms"] for r in res])
            ok = sum(r["status"] == 0 for r in res)
            row = {"lambda": lam, "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                   "mean_service_ms": float(np.mean([r["service_ms"] for r in res])), "throughput_req_s": len(res) / (t1 - t0),
                   "converged": ok, "requests": len(res), "iterations_mean": float(np.mean([r["iterations"] for r in res]))}
            if ok < len(res):
                row["not_converged"] = [{"model": k, "seed": sd, "status": r["status"], "rel_res": r["rel_res"]}
                                        for (k, sd), r in zip(sent, res) if r["status"] != 0]
            rows.append(row)
            log(f"serving {name} lambda {lam:g}/s: p50 {row['p50_ms']:.1f} ms, p99 {row['p99_ms']:.1f} ms, throughput "
                f"{row['throughput_req_s']:.1f} req/s, {ok}/{len(res)} converged")
        st = srv.stats()
        srv.close()
        out[name] = {"rates": rows, "peak_running": st["peak_running"], "hbm_used_GB": st["hbm_used"] / 1e9, "evictions": st["evictions"]}
        if batched:
            out[name].update({"groups": st["groups"], "group_models": st["group_models"], "rounds": st["rounds"], "group_cells": args.serve_group_cells})
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_sample(models, 8, "serving")
    return out


def cpu_sample(models, count, what):
    """The reference's OpenMP BlockCPR + BiCGSTAB (bsr.cpp:993-1084, linsolve.cpp:159-216) on this host's cores, on a
    bounded, size-stratified sample of the systems of a leg (count of them, evenly spaced by size): one solve = the
    system's set-up + solve to 1e-8, one at a time with every thread on it (seconds of CPU work for the batched
    systems, 10-20 s for the served models)."""
    import oracle
    ref = oracle.reference(fast=True)
    kind = "reference"
    if ref is None:
        ref = oracle.oracle()
        kind = "port"
    threads = ref.set_threads(0)
    order = sorted(range(len(models)), key=lambda i: models[i].n)
    pick = order[len(order) // (2 * count)::max(1, len(order) // count)][:count]
    ms, cells = [], []
    for i in pick:
        m = models[i]
        t0 = time.perf_counter()
        _, its, rel, conv = ref.krylov(m, m.rhs, "cpr", "bicgstab", 1e-8, 3000)
        ms.append((time.perf_counter() - t0) * 1e3)
        cells.append(int(m.n))
    lat = np.array(ms)
    log(f"{what} cpu baseline: {len(pick)} solves, mean {lat.mean():.1f} ms, p50 {np.percentile(lat, 50):.1f} ms, max {lat.max():.1f} ms "
        f"on {threads} threads ({kind})")
    return {"solves": len(pick), "cells": cells, "latency_ms": [round(x, 2) for x in ms], "mean_ms": float(lat.mean()),
            "p50_ms": float(np.percentile(lat, 50)), "max_ms": float(lat.max()), "solves_per_s": float(1e3 / lat.mean()),
            "cores": int(threads), "kind": kind,
            "sample": f"{len(pick)} of the {len(models)} systems, evenly spaced by size, base values: set-up + CPR-BiCGSTAB to 1e-8, "
                      f"one at a time"}


# the CPR second stage of the variants that do not use the multicolour BILU0 (the workload text of a run led by them)
SECOND_STAGE = {"cpr": "natural-order BILU0", "cpr_mixed": "natural-order BILU0", "cpr_exact_gs": "natural-order BILU0",
                "cpr_bj": "block-Jacobi", "cpr_rbt": "BILU0 of red / black 10x10x6 tiles"}


def serving_process(args):
    """The serving leg in a process of its own, as a serving loop runs (a child of this one, which keeps its own
    device allocations meanwhile): after the other legs in one process its saturation rate read about a quarter lower
    (195 against 257 requests/s at 2000/s) for reasons not pinned down.  Its log goes to this stderr, its JSON back."""
    import subprocess
    cmd = [sys.executable, "-u", os.path.abspath(__file__), "--serving-only", "--serve-models", str(args.serve_models), "--serve-slots",
           str(args.serve_slots), "--serve-requests", str(args.serve_requests), "--serve-lambdas", args.serve_lambdas,
           "--serve-group-cells", str(args.serve_group_cells), "--tol", repr(args.tol), "--max-it", str(args.max_it)]
    if args.serve_forms:
        cmd += ["--serve-forms", args.serve_forms]
    if args.no_cpu:
        cmd.append("--no-cpu")
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=None, text=True, timeout=900)
    if p.returncode != 0:
        raise RuntimeError(f"serving process exited with {p.returncode}")
    out = json.loads(p.stdout.strip().splitlines()[-1])
    out["process"] = "own (bench.py --serving-only)"
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=float, default=1.0, help="shrink every grid dimension (quick runs only)")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--max-it", type=int, default=3000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-lean", action="store_true", help="also time the reference's lean-BILU0 CPR on the host (about 20 s more)")
    ap.add_argument("--extra", action="store_true", help="also time the exact Gauss-Seidel CPR (slow)")
    ap.add_argument("--configs", default="", help="comma list of variants to run (default: all but the --extra ones)")
    ap.add_argument("--kernels-only", action="store_true", help="only the per-kernel timing loops (for counter runs)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-systems leg")
    ap.add_argument("--batched-only", action="store_true", help="only the batched-systems leg")
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--dump", default="", help="solve a Jacobian dumped by the engine (RESSIM_DUMP format) instead of the synthetic system")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-loop leg")
    ap.add_argument("--serving-only", action="store_true", help="only the serving-loop leg")
    ap.add_argument("--serve-models", type=int, default=64)
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-requests", type=int, default=128)
    ap.add_argument("--serve-lambdas", default="10,25,50,200,2000")
    ap.add_argument("--serve-forms", default="", help="comma list of serving forms (default: all)")
    ap.add_argument("--serve-group-cells", type=int, default=150000, help="batched serving: cells per model group")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (requested --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    import ressimstudio_amd as rs
    from ressimstudio_amd import gen

    if args.batched_only:
        print(json.dumps(batched_leg(rs, gen, args)))
        return
    if args.serving_only:
        print(json.dumps(serving_leg(rs, gen, args)))
        return
    t = time.perf_counter()
    acc = None
    if args.dump:
        from ressimstudio_amd import dump
        dd = dump.read_dump(args.dump)
        s = dump.to_system(dd)
        acc = dd["acc"]           # the engine wrote accumulation blocks: true-IMPES weights, as the engine used them
        args.scale = 0.0          # not the 1.1M-cell synthetic workload: no traffic file, no batched / serving legs
    else:
        s = gen.large(scale=args.scale)
    log(f"workload: {s.n} cells x {s.nb}, {s.nnzb} blocks ({s.fill_blocks} well fill), generated in {time.perf_counter() - t:.1f}s")
    ctx = rs.default_context()
    sysd = rs.System(s, ctx)
    if acc is not None:
        sysd.set_accumulation(acc)
        log("CPR weights: true-IMPES from the dump's accumulation blocks")
    dev = DeviceInputs(rs, ctx, s)

    J, GS, FP64, MIXED, BI, FG = rs.SMOOTH_JACOBI, rs.SMOOTH_GS, rs.FP64, rs.MIXED, rs.KRYLOV_BICGSTAB, rs.KRYLOV_FGMRES
    all_cfgs = {
        # head: two Jacobi steps per level with weight 1 (DeviceCprAmg's one step with weight 0.6667 is cpr_mc_j1; the
        # weight 1 takes 7 % fewer iterations on average over four seeds of this workload)
        "cpr_mc": rs.default_options(precond="cpr_mc", precision=FP64, krylov=BI, smoother=J, smooth_steps=2, jacobi_omega=1.0),
        "cpr_mc_w067": rs.default_options(precond="cpr_mc", precision=FP64, krylov=BI, smoother=J, smooth_steps=2),
        "cpr_mc_j1": rs.default_options(precond="cpr_mc", precision=FP64, krylov=BI, smoother=J, smooth_steps=1),
        # DeviceCprAmg: Jacobi V-cycle (one step, weight 0.6667) + block-Jacobi second stage (no ILU)
        "cpr_bj": rs.default_options(precond="cpr_bj", precision=FP64, krylov=BI, smoother=J, smooth_steps=1),
        "cpr": rs.default_options(precond="cpr", precision=FP64, krylov=BI, smoother=J, smooth_steps=2, jacobi_omega=1.0),
        # the head's CPR with the BILU0 of red / black tiles of 10 x 10 x 6 cells instead of the cell colouring
        "cpr_rbt": rs.default_options(precond="cpr_mc", precision=FP64, krylov=BI, smoother=J, smooth_steps=2, jacobi_omega=1.0,
                                      ilu_rb_tiles=10 | (10 << 8) | (6 << 16)),
        "ilu0": rs.default_options(precond="bilu0", precision=FP64, krylov=BI),
        "mcilu0": rs.default_options(precond="mcbilu0", precision=FP64, krylov=BI),
        # FP32-stored preconditioner (factors, hierarchy, pressure column) with FP64 vectors and arithmetic, FP64 BiCGSTAB
        "cpr_mc_store32": rs.default_options(precond="cpr_mc", precision=rs.MIXED_STORE, krylov=BI, smoother=J, smooth_steps=2,
                                             jacobi_omega=1.0),
        # the head with the reference's lean BILU0 factorisation (its engine default from 500k cells): float32 factors
        # eliminated against rounded rows, FP64 hierarchy, vectors and arithmetic
        "cpr_mc_lean": rs.default_options(precond="cpr_mc", precision=FP64, krylov=BI, smoother=J, smooth_steps=2, jacobi_omega=1.0,
                                          ilu_lean=1),
        # FP32 preconditioner inside FP64 BiCGSTAB / FGMRES(30); FP64 refinement around FP32-operator inner solves
        "cpr_mc_mixed": rs.default_options(precond="cpr_mc", precision=MIXED, krylov=BI, smoother=J, smooth_steps=2, jacobi_omega=1.0),
        "cpr_mc_mixed_fgmres": rs.default_options(precond="cpr_mc", precision=MIXED, krylov=FG, smoother=J, restart=30, smooth_steps=2,
                                                  jacobi_omega=1.0),
        "cpr_mc_mixed_ir": rs.default_options(precond="cpr_mc", precision=rs.MIXED_IR, krylov=BI, smoother=J, smooth_steps=2, jacobi_omega=1.0),
        "cpr_mixed": rs.default_options(precond="cpr", precision=MIXED, krylov=BI, smoother=J, smooth_steps=2, jacobi_omega=1.0),
        "cpr_exact_gs": rs.default_options(precond="cpr", precision=FP64, krylov=BI, smoother=GS),
    }
    names = [c for c in args.configs.split(",") if c] or [n for n in all_cfgs if n != "cpr_exact_gs" or args.extra]
    cfgs = {n: all_cfgs[n] for n in names}
    if args.kernels_only:
        L = rs.lib()
        rs._check(L.rsh_system_set_values_device(sysd.h, dev.val), "set_values_device")
        rs._check(L.rsh_system_setup(sysd.h, ctypes.byref(cfgs[names[0]])), "setup")
        kinds = [(4, "copy"), (0, "spmv"), (2, "ilu_sweeps")]
        if cfgs[names[0]].smoother == J and cfgs[names[0]].precond in (rs.PREC_CPR, rs.PREC_CPR_MC, rs.PREC_CPR_BJ):
            kinds += [(5, "smoother_l0"), (6, "smoother_l1")]
        for kind, name in kinds:
            ms, by = sysd.time_kernel(kind, 20)
            log(f"kernel {name}: {ms * 1e3:.1f} us" + (f", {by / ms / 1e6:.0f} GB/s" if by > 0 else ""))
        return
    results = {}
    solutions = {}
    x_fp64 = None
    for name, o in cfgs.items():
        # symbolic set-up (once per pattern) outside the timed region
        step(rs, sysd, dev, o, args.tol, args.max_it)
        steps = args.steps if name == names[0] else 1
        dt, its, rel, info = time_config(rs, sysd, dev, o, args.tol, args.max_it, steps, args.warmup if name == names[0] else 0)
        x = dev.x_host()
        solutions[name] = x
        results[name] = {"solves_per_s": 1.0 / dt, "ms_per_solve": dt * 1e3, "iterations": its, "rel_res": rel, "steps": steps,
                         "setup_ms": info["last_setup_ms"], "solve_ms": info["last_solve_ms"], "outer": info["last_outer"]}
        if x_fp64 is None and o.precision == FP64:
            x_fp64 = x
        elif x_fp64 is not None:
            results[name]["err_vs_fp64"] = float(np.linalg.norm(x - x_fp64) / np.linalg.norm(x_fp64))
        log(f"{name}: {1.0 / dt:.3f} solves/s ({dt * 1e3:.1f} ms: setup {info['last_setup_ms']:.1f} + solve {info['last_solve_ms']:.1f}), "
            f"{its} its, rel {rel:.2e}")

    # kernel timing (HIP events on the context's stream) for the roofline of the dominant kernel
    head = names[0]
    sysd.setup(cfgs[head])
    k = {}
    kinds = [(4, "copy"), (0, "spmv"), (2, "ilu_sweeps"), (3, "amg_vcycle"), (1, "cpr_apply")]
    if cfgs[head].smoother == J:
        kinds += [(5, "smoother_l0"), (6, "smoother_l1")]   # one damped-Jacobi step of AMG level 0 / level 1
    for kind, name in kinds:
        try:
            ms, by = sysd.time_kernel(kind, 10 if kind in (1, 2, 3) else 50)
        except rs.RshError:
            if kind in (0, 2, 4):
                raise
            continue   # no AMG (or no Jacobi-smoothed AMG) in this preconditioner
        k[name] = {"ms": ms, "bytes": by, "GBs": (by / (ms * 1e-3) / 1e9) if by > 0 else None}
        log(f"kernel {name}: {ms * 1e3:.1f} us" + (f", {k[name]['GBs']:.0f} GB/s" if by > 0 else ""))
    its_cpr = results[head]["iterations"]
    # the single kernel with the largest share of the solve's GPU time is the BSR SpMV (two per iteration; the
    # kernel trace in profiles/); the BILU0 apply (four to six launches) is reported beside it
    dom = "spmv"
    roof = {"bound": "hbm", "kernel": f"kSpmvSliced<{s.nb},double>: BSR SpMV of the {s.n}-cell system (largest single-kernel share of the solve)",
            "achieved": k[dom]["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": k[dom]["GBs"] / HBM_PEAK_GBS, "traffic": None,
            "algorithmic_bytes": k[dom]["bytes"], "ilu_apply_GBs": k["ilu_sweeps"]["GBs"],
            "ilu_apply_frac": k["ilu_sweeps"]["GBs"] / HBM_PEAK_GBS, "ilu_apply_bytes": k["ilu_sweeps"]["bytes"],
            "copy_GBs": k["copy"]["GBs"]}
    for name in ("smoother_l0", "smoother_l1"):
        if name in k:
            roof[name] = {"us": k[name]["ms"] * 1e3, "algorithmic_bytes": k[name]["bytes"], "GBs": k[name]["GBs"],
                          "frac": k[name]["GBs"] / HBM_PEAK_GBS}
    # the natural-order BILU0 apply (tiled sweeps, the factors the reference computes) beside it, when that variant ran
    if "cpr" in cfgs:
        sysd.setup(cfgs["cpr"])
        ms, by = sysd.time_kernel(2, 10)
        inf = sysd.info()
        roof["ilu_natural_apply"] = {"us": ms * 1e3, "algorithmic_bytes": by, "GBs": by / (ms * 1e-3) / 1e9, "frac": by / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                     "tiled": inf.get("ilu_tiled", 0), "tiles": inf.get("ilu_tiles")}
        log(f"kernel ilu_natural_apply: {ms * 1e3:.1f} us, {by / (ms * 1e-3) / 1e9:.0f} GB/s")
        sysd.setup(cfgs[head])
    # HBM bytes per launch from the FETCH_SIZE / WRITE_SIZE counter runs of this workload (profiles/)
    tfile = os.path.join(ROOT, "profiles", "traffic_cpr_mc_1.1M.json")
    if args.scale == 1.0 and head == "cpr_mc" and os.path.exists(tfile):
        with open(tfile) as f:
            tr = json.load(f)
        roof["traffic"] = tr["spmv"]["traffic_bytes"]
        roof["ilu_apply_traffic"] = tr["mc_bilu0_apply"]["traffic_bytes"]
        for name in ("smoother_l0", "smoother_l1"):
            if name in roof and name in tr:
                roof[name]["traffic"] = tr[name]["traffic_bytes"]
        if "ilu_natural_apply" in roof and "natural_bilu0_apply" in tr:
            roof["ilu_natural_apply"]["traffic"] = tr["natural_bilu0_apply"]["traffic_bytes"]
        roof["traffic_source"] = "profiles/cpr_mc_fp64_1.1M_hbm_traffic.md"

    batched = None
    if not args.no_batched and args.scale == 1.0:
        try:
            batched = batched_leg(rs, gen, args)
        except Exception as e:  # noqa: BLE001
            log(f"batched leg failed: {e}")
            batched = {"error": str(e)}

    serving = None
    if not args.no_serving and args.scale == 1.0:
        try:
            serving = serving_process(args)
        except Exception as e:  # noqa: BLE001
            log(f"serving leg failed: {e}")
            serving = {"error": str(e)}

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(s, args.tol, lean=args.cpu_lean)
            log(f"cpu baseline: {cpu['value']:.4f} solves/s on {cpu['cores']} threads ({cpu['sample']}); fresh set-up "
                f"{cpu['fresh']['value']:.4f} solves/s ({cpu['fresh']['sample']})")
            if "lean" in cpu:
                log(f"cpu baseline, lean BILU0: {cpu['lean']['value']:.4f} solves/s, {cpu['lean']['iterations']} iterations ({cpu['lean']['sample']})")
        except Exception as e:  # noqa: BLE001
            log(f"cpu baseline unavailable: {e}")
        try:
            ref_err = error_vs_reference(s, solutions)
            for name, e in ref_err["errors"].items():
                results[name]["err_vs_reference"] = e
            log(f"error against the reference's FP64 solution ({ref_err['sample']}): " +
                ", ".join(f"{k} {v:.1e}" for k, v in ref_err["errors"].items()))
        except Exception as e:  # noqa: BLE001
            log(f"reference solution unavailable: {e}")

    main_r = results[head]
    # the mixed-precision forms against the plain FP64 GPU ILU0-BiCGSTAB and the FP64 head
    mixed = {}
    for name in ("cpr_mc_store32", "cpr_mc_lean", "cpr_mc_mixed", "cpr_mc_mixed_ir", "cpr_mc_mixed_fgmres"):
        if name in results:
            r = results[name]
            mixed[name] = {"solves_per_s": r["solves_per_s"], "err_vs_fp64": r.get("err_vs_fp64"), "err_vs_reference": r.get("err_vs_reference"),
                           "outer": r.get("outer"),
                           "vs_fp64_head": r["solves_per_s"] / main_r["solves_per_s"]}
            if "ilu0" in results:
                mixed[name]["vs_plain_fp64_ilu0"] = r["solves_per_s"] / results["ilu0"]["solves_per_s"]
    out = {
        "metric": "CPR/ILU0 solves/s @1.1M cells, rel.res 1e-8; HBM GB/s; p50/p99 batched latency",
        "value": main_r["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_r["ms_per_solve"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, ressimstudio_amd/csrc/gen)",
        "config": {"workload": (f"engine Jacobian dump {os.path.basename(args.dump)}, {s.n} cells x nb {s.nb}" if args.dump else
                                f"black-oil Jacobian {s.n} cells 60x220x85 x nb 3, 5 wells") + f"; variant {head}: BiCGSTAB to 1e-8 "
                               f"preconditioned by CPR ({'true' if acc is not None else 'quasi'}-IMPES weights, aggregation AMG V-cycle with two Jacobi steps (weight 1) per level) + "
                               f"{SECOND_STAGE.get(head, 'multicolour BILU0')}; numeric setup + solve per step",
                   "cells": s.n, "block": s.nb, "blocks": s.nnzb, "tol": args.tol, "scale": args.scale},
        "iterations": main_r["iterations"],
        "variants": results,
        "mixed_precision": mixed,
        "kernels": k,
        "roofline": roof,
        "cpu_baseline": cpu,
        "batched": batched,
        "serving": serving,
    }
    print(json.dumps(out))
    dev.free()


if __name__ == "__main__":
    main()
