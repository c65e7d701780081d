"""Writes tests/golden/well_schur.json: the engine's own well Schur reduction (ressim::WellSchur::build / reduceRhs /
recover, engine/src/bsr.cpp:733-990, compiled from its sources into oracle/_ref/libressim_ref.so by oracle/Makefile)
of the Jacobians in tests/golden/well_cases.py, so that the reduction is pinned where the reference is absent.

Run from the repository root:  python tests/golden/make_well_golden.py
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
sys.path.insert(0, HERE)

import oracle  # noqa: E402
import well_cases  # noqa: E402


def main():
    ref = oracle.reference()
    if ref is None:
        raise SystemExit("oracle/_ref/libressim_ref.so is not built (needs the reference sources)")
    out = {}
    for name, J in well_cases.cases().items():
        w = ref.wellschur(J["row_ptr"], J["col"], J["val"], J["na"], J["nv"], J["nw"])
        assert w is not None, name
        rp, ci, v = w.reduced()
        xc = np.cos(0.3 * np.arange(J["na"] * J["nv"]))
        arrs = {"row_ptr": rp, "col": ci, "val": v, "bc": w.reduce_rhs(J["rhs"]), "x": w.recover(J["rhs"], xc)}
        out[name] = {"n": w.n, "nnzb": w.nnzb, "fill": w.fill, "sha256": {k: digest(a) for k, a in arrs.items()}}
        print(name, w.n, w.nnzb, "fill", w.fill, flush=True)
    for name, J in well_cases.refused().items():
        assert ref.wellschur(J["row_ptr"], J["col"], J["val"], J["na"], J["nv"], J["nw"]) is None, name
    with open(os.path.join(HERE, "well_schur.json"), "w") as f:
        json.dump(out, f, indent=1)


def digest(a) -> str:
    """sha256 of an array's bytes (int32 indices, float64 values): equal digests mean bit-identical arrays"""
    a = np.ascontiguousarray(a)
    a = a.astype(np.int32) if a.dtype.kind == "i" else a.astype(np.float64)
    return hashlib.sha256(a.tobytes()).hexdigest()

if __name__ == "__main__":
    main()
