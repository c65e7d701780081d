"""The reference engine's Jacobian dump (``RESSIM_DUMP``, simulator.cpp:3252-3262): the scalar CSR Jacobian of the
first Newton iteration and its right-hand side, as written by the engine, plus the optional accumulation blocks
``<file>.acc`` (simulator.cpp:3313-3315, written when true-IMPES weights are in use).

Layout (little-endian): int32 n, na, nv, nw, nnz; int32 rowPtr[n+1]; int32 col[nnz]; float64 val[nnz]; float64
rhs[n].  Rows are cells in blocks of nv unknowns (pressure first), then the nw well BHP rows; ``to_system`` applies
the well Schur reduction to obtain the BSR system the solver boundary receives.  (The reference's
tools/lsbench.cpp reads a different header order and is stale.)"""
from __future__ import annotations

import os

import numpy as np


def read_dump(path: str) -> dict:
    with open(path, "rb") as f:
        data = f.read()
    hdr = np.frombuffer(data, dtype="<i4", count=5)
    n, na, nv, nw, nnz = (int(v) for v in hdr)
    if n <= 0 or nnz <= 0 or na * nv + nw != n:
        raise ValueError(f"{path}: not a Jacobian dump (n={n}, na={na}, nv={nv}, nw={nw}, nnz={nnz})")
    expect = 20 + 4 * (n + 1) + 4 * nnz + 8 * nnz + 8 * n
    if len(data) != expect:
        raise ValueError(f"{path}: {len(data)} bytes, expected {expect}")
    off = 20
    rp = np.frombuffer(data, dtype="<i4", count=n + 1, offset=off).astype(np.int32)
    off += 4 * (n + 1)
    ci = np.frombuffer(data, dtype="<i4", count=nnz, offset=off).astype(np.int32)
    off += 4 * nnz
    val = np.frombuffer(data, dtype="<f8", count=nnz, offset=off).astype(np.float64)
    off += 8 * nnz
    rhs = np.frombuffer(data, dtype="<f8", count=n, offset=off).astype(np.float64)
    if rp[0] != 0 or rp[-1] != nnz or np.any(np.diff(rp) < 0) or ci.min() < 0 or ci.max() >= n:
        raise ValueError(f"{path}: inconsistent CSR arrays")
    out = {"n": n, "na": na, "nv": nv, "nw": nw, "row_ptr": rp, "col": ci, "val": val, "rhs": rhs, "acc": None}
    acc_path = path + ".acc"
    if os.path.exists(acc_path):
        acc = np.fromfile(acc_path, dtype="<f8")
        if acc.size != na * nv * nv:
            raise ValueError(f"{acc_path}: {acc.size} values, expected {na * nv * nv}")
        out["acc"] = acc.reshape(na, nv, nv)
    return out


def write_dump(path: str, d: dict) -> None:
    n, nnz = int(d["n"]), int(d["row_ptr"][-1])
    with open(path, "wb") as f:
        f.write(np.array([n, d["na"], d["nv"], d["nw"], nnz], "<i4").tobytes())
        f.write(np.asarray(d["row_ptr"], "<i4").tobytes())
        f.write(np.asarray(d["col"], "<i4").tobytes())
        f.write(np.asarray(d["val"], "<f8").tobytes())
        f.write(np.asarray(d["rhs"], "<f8").tobytes())
    if d.get("acc") is not None:
        np.asarray(d["acc"], "<f8").tofile(path + ".acc")


def to_system(d: dict):
    """Well-Schur reduction (WellSchur::build, bsr.cpp:733-840) of a dump into the BSR system of the solver
    boundary; the returned system's ``recover(rhs_full, x_cells)`` gives the full solution with the BHPs."""
    from . import gen
    return gen.from_csr(d["row_ptr"], d["col"], d["val"], d["rhs"], d["na"], d["nv"], d["nw"])
