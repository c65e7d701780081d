"""Writes tests/golden/ref_vectors.npz + ref_vectors.json: outputs of the reference's own solver code compiled from
its sources (oracle/_ref/libressim_ref.so, built by oracle/Makefile where the reference is mounted) on seeded
systems from ressimstudio_amd.gen.  The fixtures let the oracle be pinned where the reference is absent.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from ressimstudio_amd import gen  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def cases():
    return {
        "small": gen.small(),
        "chain": gen.chain(),
        "laplace": gen.block_laplace(40, 30, 3, 11),
        "nb2": gen.reservoir(12, 10, 3, 2, model=0, sigma=1.5, wells=2, seed=5),
        "layered": gen.reservoir(12, 30, 14, 3, model=1, sigma=2.5, kvkh=0.1, wells=5, seed=7),
    }


KEEP_X = {("small", "cpr", "bicgstab"), ("chain", "bilu0", "gmres"), ("nb2", "mcbilu0", "bicgstab")}
KEEP_VEC = {"small", "chain", "nb2"}   # full preconditioner outputs kept for these (checksums for the rest)


def main():
    ref = oracle.reference()
    if ref is None:
        raise SystemExit("oracle/_ref/libressim_ref.so is not built (needs the reference sources)")
    meta = {"generator": "ressimstudio_amd.gen", "tol": 1e-8, "cases": {}}
    arrays = {}
    for name, s in cases().items():
        c = {"n": s.n, "nb": s.nb, "nnzb": s.nnzb, "val_sum": float(np.sum(s.val)), "rhs_sum": float(np.sum(s.rhs)), "solves": {}}
        for prec in ["bilu0", "bilu0f", "mcbilu0", "cpr"]:
            for meth in ["bicgstab", "gmres"]:
                x, its, rel, conv = ref.krylov(s, s.rhs, prec, meth, 1e-8, 2000, 40)
                c["solves"][f"{prec}/{meth}"] = {"its": its, "rel": rel, "conv": conv, "x_norm": float(np.linalg.norm(x)),
                                                 "x_sum": float(np.sum(x))}
                if (name, prec, meth) in KEEP_X:
                    arrays[f"{name}__{prec}__{meth}__x"] = x
        r = np.cos(0.7 * np.arange(s.n * s.nb))
        cp = ref.cpr(s)
        vecs = {"bilu0_apply": ref.bilu0(s, 1).apply(r), "mcbilu0_apply": ref.mcbilu0(s).apply(r), "cpr_weights": cp.weights(),
                "cpr_apply": cp.apply(r)}
        c["checksums"] = {k2: [float(np.sum(v)), float(np.linalg.norm(v))] for k2, v in vecs.items()}
        if name in KEEP_VEC:
            for k2, v in vecs.items():
                arrays[f"{name}__{k2}"] = v
        k, colour, perm = ref.mcbilu0(s).colouring()
        c["colours"] = k
        arrays[f"{name}__colour"] = colour.astype(np.int16 if k < 32000 else np.int32)
        c["amg_levels"] = cp.levels()
        nc, agg = cp.aggregates(0) if len(cp.levels()) > 1 else (0, np.zeros(0, np.int32))
        arrays[f"{name}__agg0"] = agg
        meta["cases"][name] = c
        print(name, {k2: v["its"] for k2, v in c["solves"].items()}, "levels", c["amg_levels"], flush=True)
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **arrays)
    with open(os.path.join(HERE, "ref_vectors.json"), "w") as f:
        json.dump(meta, f, indent=1)


if __name__ == "__main__":
    main()
