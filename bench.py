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


def cpu_baseline(s, tol):
    """The reference's BlockCPR + BiCGSTAB (BlockCPR::setup + bicgstab, bsr.cpp:993-1084, linsolve.cpp:159-216)
    compiled from its sources with its release flags (OpenMP, -O3 -mavx2 -mfma), on this host's cores: one complete
    set-up + solve of the same system to the same tolerance (about 20 s of CPU work)."""
    import oracle
    ref = oracle.reference(fast=True)
    kind = "reference"
    if ref is None:
        ref = oracle.oracle()
        kind = "port"
    threads = ref.set_threads(0)
    t0 = time.perf_counter()
    _, its, rel, conv = ref.krylov(s, s.rhs, "cpr", "bicgstab", tol, 3000)
    dt = time.perf_counter() - t0
    return {"value": 1.0 / dt, "unit": "solves/s", "cores": int(threads), "kind": kind,
            "sample": f"one full CPR-BiCGSTAB set-up + solve of the {s.n}-cell system to {tol:g}: {dt:.1f} s, {its} iterations, "
                      f"rel.res {rel:.2e}, converged {conv}",
            "iterations": its}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=float, default=1.0, help="shrink every grid dimension (quick runs only)")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--max-it", type=int, default=3000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--extra", action="store_true", help="also time the exact Gauss-Seidel CPR (slow)")
    ap.add_argument("--configs", default="", help="comma list of variants to run (default: all but the --extra ones)")
    ap.add_argument("--kernels-only", action="store_true", help="only the per-kernel timing loops (for counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (requested --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    import ressimstudio_amd as rs
    from ressimstudio_amd import gen

    t = time.perf_counter()
    s = gen.large(scale=args.scale)
    log(f"workload: {s.n} cells x {s.nb}, {s.nnzb} blocks ({s.fill_blocks} well fill), generated in {time.perf_counter() - t:.1f}s")
    ctx = rs.default_context()
    sysd = rs.System(s, ctx)
    dev = DeviceInputs(rs, ctx, s)

    J, GS, FP64, MIXED, BI, FG = rs.SMOOTH_JACOBI, rs.SMOOTH_GS, rs.FP64, rs.MIXED, rs.KRYLOV_BICGSTAB, rs.KRYLOV_FGMRES
    all_cfgs = {
        "cpr_mc": rs.default_options(precond="cpr_mc", precision=FP64, krylov=BI, smoother=J),
        "cpr": rs.default_options(precond="cpr", precision=FP64, krylov=BI, smoother=J),
        "ilu0": rs.default_options(precond="bilu0", precision=FP64, krylov=BI),
        "mcilu0": rs.default_options(precond="mcbilu0", precision=FP64, krylov=BI),
        # FP32 preconditioner inside FP64 BiCGSTAB / FGMRES(30); FP64 refinement around FP32-operator inner solves
        "cpr_mc_mixed": rs.default_options(precond="cpr_mc", precision=MIXED, krylov=BI, smoother=J),
        "cpr_mc_mixed_fgmres": rs.default_options(precond="cpr_mc", precision=MIXED, krylov=FG, smoother=J, restart=30),
        "cpr_mc_mixed_ir": rs.default_options(precond="cpr_mc", precision=rs.MIXED_IR, krylov=BI, smoother=J),
        "cpr_mixed": rs.default_options(precond="cpr", precision=MIXED, krylov=BI, smoother=J),
        "cpr_exact_gs": rs.default_options(precond="cpr", precision=FP64, krylov=BI, smoother=GS),
    }
    names = [c for c in args.configs.split(",") if c] or [n for n in all_cfgs if n != "cpr_exact_gs" or args.extra]
    cfgs = {n: all_cfgs[n] for n in names}
    if args.kernels_only:
        L = rs.lib()
        rs._check(L.rsh_system_set_values_device(sysd.h, dev.val), "set_values_device")
        rs._check(L.rsh_system_setup(sysd.h, ctypes.byref(cfgs[names[0]])), "setup")
        for kind, name in [(4, "copy"), (0, "spmv"), (2, "ilu_sweeps")]:
            ms, by = sysd.time_kernel(kind, 20)
            log(f"kernel {name}: {ms * 1e3:.1f} us" + (f", {by / ms / 1e6:.0f} GB/s" if by > 0 else ""))
        return
    results = {}
    x_fp64 = None
    for name, o in cfgs.items():
        # symbolic set-up (once per pattern) outside the timed region
        step(rs, sysd, dev, o, args.tol, args.max_it)
        steps = args.steps if name == names[0] else 1
        dt, its, rel, info = time_config(rs, sysd, dev, o, args.tol, args.max_it, steps, args.warmup if name == names[0] else 0)
        x = dev.x_host()
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
    for kind, name in [(4, "copy"), (0, "spmv"), (2, "ilu_sweeps"), (3, "amg_vcycle"), (1, "cpr_apply")]:
        ms, by = sysd.time_kernel(kind, 10 if kind in (1, 2, 3) else 50)
        k[name] = {"ms": ms, "bytes": by, "GBs": (by / (ms * 1e-3) / 1e9) if by > 0 else None}
        log(f"kernel {name}: {ms * 1e3:.1f} us" + (f", {k[name]['GBs']:.0f} GB/s" if by > 0 else ""))
    its_cpr = results[head]["iterations"]
    # the preconditioner dominates (two applies per iteration); within it the BILU0 sweeps are the largest part
    dom = "ilu_sweeps"
    roof = {"bound": "hbm", "kernel": f"BILU0 sweeps of the {head} preconditioner (forward + backward, all launches of one apply)",
            "achieved": k[dom]["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": k[dom]["GBs"] / HBM_PEAK_GBS, "traffic": None,
            "algorithmic_bytes": k[dom]["bytes"], "spmv_GBs": k["spmv"]["GBs"], "spmv_frac": k["spmv"]["GBs"] / HBM_PEAK_GBS,
            "copy_GBs": k["copy"]["GBs"]}
    # HBM bytes per launch from the FETCH_SIZE / WRITE_SIZE counter runs of this workload (profiles/)
    tfile = os.path.join(ROOT, "profiles", "traffic_cpr_mc_1.1M.json")
    if args.scale == 1.0 and head == "cpr_mc" and os.path.exists(tfile):
        with open(tfile) as f:
            tr = json.load(f)
        roof["traffic"] = tr["mc_bilu0_apply"]["traffic_bytes"]
        roof["spmv_traffic"] = tr["spmv"]["traffic_bytes"]
        roof["traffic_source"] = "profiles/cpr_mc_fp64_1.1M_hbm_traffic.md"

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(s, args.tol)
            log(f"cpu baseline: {cpu['value']:.4f} solves/s on {cpu['cores']} threads ({cpu['sample']})")
        except Exception as e:  # noqa: BLE001
            log(f"cpu baseline unavailable: {e}")

    main_r = results[head]
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
        "config": {"workload": f"black-oil Jacobian {s.n} cells 60x220x85 x nb 3, 5 wells; variant {head}: BiCGSTAB to 1e-8 "
                               f"preconditioned by CPR (quasi-IMPES, aggregation AMG V(1,1)) + BILU0; numeric setup + solve per step",
                   "cells": s.n, "block": s.nb, "blocks": s.nnzb, "tol": args.tol, "scale": args.scale},
        "iterations": main_r["iterations"],
        "variants": results,
        "kernels": k,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    dev.free()


if __name__ == "__main__":
    main()
