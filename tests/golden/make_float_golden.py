"""Writes tests/golden/float_solve.json: the reference's float32 device solve (CpuFloatBlockSolver::solve,
engine/src/bsr.cpp:1130-1215) and the FP64 refinement around it (refineWithDevice, bsr.cpp:1373-1396), compiled from
its sources into oracle/_ref/libressim_ref.so by oracle/Makefile, on seeded systems from ressimstudio_amd.gen: sha256
digests of the solutions, iteration counts and residuals, so the oracle's restatement is pinned where the reference
is absent.

Run from the repository root:  python tests/golden/make_float_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from ressimstudio_amd import gen  # noqa: E402


def cases():
    return {"small": gen.small(), "nb2": gen.reservoir(12, 10, 3, 2, model=0, sigma=1.5, wells=2, seed=5),
            "nb4": gen.reservoir(10, 9, 3, 4, model=0, sigma=0.8, wells=1, seed=6),
            "layered": gen.reservoir(12, 30, 14, 3, model=1, sigma=2.5, kvkh=0.1, wells=5, seed=7)}


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, np.float64).tobytes()).hexdigest()


def main():
    ref = oracle.reference()
    if ref is None:
        raise SystemExit("oracle/_ref/libressim_ref.so is not built (needs the reference sources)")
    out = {}
    for name, s in cases().items():
        for kind in ("bilu0f", "mcbilu0", "devcpr"):
            x, its, fin = ref.float_solve(s, s.rhs, kind, 1e-4, 500)
            xr, its_r, rel_r, conv = ref.refine_float(s, s.rhs, kind, 1e-8, 10, 500)
            out[f"{name}/{kind}"] = {"inner": {"its": its, "finite": fin, "x": digest(x)},
                                     "refined": {"its": its_r, "rel": rel_r, "converged": conv, "x": digest(xr)}}
            print(name, kind, its, its_r, rel_r, conv, flush=True)
    with open(os.path.join(HERE, "float_solve.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
