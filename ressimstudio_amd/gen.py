"""Seeded synthetic reservoir Jacobians (the workload of the linear-solve path).

Thin ctypes binding of ``libressim_gen.so`` (csrc/gen/rs_gen.cpp).  The generator assembles the scalar CSR
Jacobian of a black-oil-like model (cells in blocks of nb unknowns, pressure first, one BHP unknown per well)
and applies the engine's exact well Schur reduction (ressim::WellSchur, engine/src/bsr.cpp:733-990) to obtain
the BSR system handed to the solver boundary (ressim::GpuBlockSolver::solve, bsr.h:143-151).

Named workloads:
  * ``small``   20x20x4, nb 3, iid lognormal k (sigma ln k = 1.0), 2 wells      (1,600 cells + 2 BHP rows)
  * ``large``   60x220x85, nb 3, layered + channelised k (sigma 2.5, kv/kh 0.1), 5 wells  (1,122,000 cells)
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

# accumulation scales: PV c_t / dt for the pressure column (absolute), inverse cell CFL number for the others
ACC_P = 1e-2
ACC_S = 10.0


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libressim_gen.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(path)
        c_ip, c_dp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
        lib.rsg_reservoir.restype = ctypes.c_void_p
        lib.rsg_reservoir.argtypes = [ctypes.c_int] * 6 + [ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_ulonglong,
                                                            ctypes.c_double, ctypes.c_double]
        lib.rsg_reservoir.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_int, ctypes.c_ulonglong, ctypes.c_double, ctypes.c_double]
        lib.rsg_block_laplace.restype = ctypes.c_void_p
        lib.rsg_block_laplace.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint]
        lib.rsg_chain.restype = ctypes.c_void_p
        lib.rsg_chain.argtypes = []
        lib.rsg_from_csr.restype = ctypes.c_void_p
        lib.rsg_from_csr.argtypes = [ctypes.c_int, c_ip, c_ip, c_dp, c_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.rsg_free.argtypes = [ctypes.c_void_p]
        lib.rsg_dims.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong)]
        lib.rsg_get_bsr.argtypes = [ctypes.c_void_p, c_ip, c_ip, c_dp, c_dp]
        lib.rsg_get_full.argtypes = [ctypes.c_void_p, c_ip, c_ip, c_dp, c_dp]
        lib.rsg_get_perm.argtypes = [ctypes.c_void_p, c_dp]
        lib.rsg_reduce_rhs.argtypes = [ctypes.c_void_p, c_dp, c_dp]
        lib.rsg_recover.argtypes = [ctypes.c_void_p, c_dp, c_dp, c_dp]
        lib.rsg_last_error.restype = ctypes.c_char_p
        _LIB = lib
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


@dataclass
class BlockSystem:
    """A Schur-reduced BSR system (the solver's input) plus, when generated from a full Jacobian, the full
    scalar system and the well-Schur data to reduce right-hand sides and recover BHPs."""
    n: int
    nb: int
    row_ptr: np.ndarray
    col: np.ndarray
    val: np.ndarray          # (nnzb, nb, nb) row-major blocks
    rhs: np.ndarray          # n * nb
    full: dict | None = None
    fill_blocks: int = 0
    handle: object = field(default=None, repr=False)

    @property
    def nnzb(self) -> int:
        return int(self.row_ptr[-1])

    def reduce_rhs(self, b_full: np.ndarray) -> np.ndarray:
        out = np.empty(self.n * self.nb)
        b_full = np.ascontiguousarray(b_full, dtype=np.float64)
        if _lib().rsg_reduce_rhs(self.handle, _p(b_full, ctypes.c_double), _p(out, ctypes.c_double)) != 0:
            raise RuntimeError(_lib().rsg_last_error().decode())
        return out

    def recover(self, b_full: np.ndarray, xc: np.ndarray) -> np.ndarray:
        nw = self.full["nw"]
        out = np.empty(self.n * self.nb + nw)
        b_full = np.ascontiguousarray(b_full, dtype=np.float64)
        xc = np.ascontiguousarray(xc, dtype=np.float64)
        if _lib().rsg_recover(self.handle, _p(b_full, ctypes.c_double), _p(xc, ctypes.c_double), _p(out, ctypes.c_double)) != 0:
            raise RuntimeError(_lib().rsg_last_error().decode())
        return out

    def __del__(self):
        if self.handle is not None and _LIB is not None:
            _LIB.rsg_free(self.handle)
            self.handle = None


def _wrap(h) -> BlockSystem:
    lib = _lib()
    if not h:
        raise RuntimeError(lib.rsg_last_error().decode())
    d = (ctypes.c_longlong * 10)()
    lib.rsg_dims(h, d)
    n, nb, nnzb, has_full, nf, nnzf, na, nv, nw, fill = (int(v) for v in d)
    rp = np.empty(n + 1, np.int32)
    ci = np.empty(nnzb, np.int32)
    v = np.empty(nnzb * nb * nb, np.float64)
    rhs = np.empty(n * nb, np.float64)
    lib.rsg_get_bsr(h, _p(rp, ctypes.c_int), _p(ci, ctypes.c_int), _p(v, ctypes.c_double), _p(rhs, ctypes.c_double))
    full = None
    if has_full:
        frp = np.empty(nf + 1, np.int32)
        fci = np.empty(nnzf, np.int32)
        fv = np.empty(nnzf, np.float64)
        fb = np.empty(nf, np.float64)
        lib.rsg_get_full(h, _p(frp, ctypes.c_int), _p(fci, ctypes.c_int), _p(fv, ctypes.c_double), _p(fb, ctypes.c_double))
        full = dict(n=nf, row_ptr=frp, col=fci, val=fv, rhs=fb, na=na, nv=nv, nw=nw)
    return BlockSystem(n=n, nb=nb, row_ptr=rp, col=ci, val=v.reshape(nnzb, nb, nb), rhs=rhs, full=full, fill_blocks=fill,
                       handle=h)


def reservoir(nx: int, ny: int, nz: int, nb: int = 3, *, model: int = 0, sigma: float = 1.0, kvkh: float = 1.0,
              wells: int = 2, seed: int = 1, acc_p: float = ACC_P, acc_s: float = ACC_S) -> BlockSystem:
    """Seeded reservoir Jacobian, Schur-reduced (see csrc/gen/rs_gen.cpp for the exact recipe)."""
    return _wrap(_lib().rsg_reservoir(nx, ny, nz, nb, model, sigma, kvkh, wells, seed, acc_p, acc_s))


def small(seed: int = 20240601) -> BlockSystem:
    """20x20x4 nb 3, ln k ~ N(ln 100, 1.0), two wells with BHP rows (1,600 cells)."""
    return reservoir(20, 20, 4, 3, model=0, sigma=1.0, kvkh=1.0, wells=2, seed=seed)


def large(seed: int = 20240602, scale: float = 1.0) -> BlockSystem:
    """60x220x85 nb 3 (1,122,000 cells), layered/channelised k, sigma 2.5, kv/kh 0.1, five wells.  ``scale`` < 1
    shrinks every grid dimension (same recipe) for quick checks."""
    nx, ny, nz = (max(2, int(round(d * scale))) for d in (60, 220, 85))
    return reservoir(nx, ny, nz, 3, model=1, sigma=2.5, kvkh=0.1, wells=5, seed=seed)


def batch_workload(count: int = 256, seed: int = 20240604) -> list:
    """Independent systems of uneven size and block size: grid dimensions log-uniform between 10x10x3 and 64x64x16
    (each dimension drawn on its own), block size 2, 3 or 4 in ratio 2:5:1, sigma ln k uniform in [0.5, 2.5], one or
    two wells with BHP rows; system i uses generator seed ``seed + i``.  The draws come from numpy's PCG64 seeded
    with ``seed``."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        nx, ny = (int(round(np.exp(rng.uniform(np.log(10), np.log(64))))) for _ in range(2))
        nz = int(round(np.exp(rng.uniform(np.log(3), np.log(16)))))
        nb = int(rng.choice([2, 3, 4], p=[0.25, 0.625, 0.125]))
        sigma = float(rng.uniform(0.5, 2.5))
        out.append(reservoir(nx, ny, nz, nb, model=0, sigma=sigma, wells=1 + i % 2, seed=seed + i))
    return out


def serving_models(count: int = 64, seed: int = 20240605) -> list:
    """Resident ensemble models for the serving loop: cells log-uniform in [5,000, 500,000] (nz uniform in 4..20,
    nx = ny from the cell count), nb 3, lognormal k with sigma ln k uniform in [0.5, 1.5], two to four wells; model i
    uses generator seed ``seed + i``; the draws come from numpy's PCG64 seeded with ``seed``."""
    return [reservoir(*a, **k) for a, k in _serving_params(count, seed)]


def serving_model(i: int, seed: int = 20240605):
    """Model i of ``serving_models(count > i, seed)`` alone."""
    a, k = _serving_params(i + 1, seed)[i]
    return reservoir(*a, **k)


def _serving_params(count: int, seed: int) -> list:
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        cells = float(np.exp(rng.uniform(np.log(5e3), np.log(5e5))))
        nz = int(rng.integers(4, 21))
        nx = max(4, int(round(np.sqrt(cells / nz))))
        sigma = float(rng.uniform(0.5, 1.5))
        wells = int(rng.integers(2, 5))
        out.append(((nx, nx, nz, 3), dict(model=0, sigma=sigma, wells=wells, seed=seed + i)))
    return out


def block_laplace(nx: int, ny: int, nb: int, seed: int) -> BlockSystem:
    """The reference tests' 2-D block Laplacian (run_tests.cpp:3694-3724), rhs = 1."""
    return _wrap(_lib().rsg_block_laplace(nx, ny, nb, seed))


def chain() -> BlockSystem:
    """The reference tests' coupled 1-D chain with two wells (run_tests.cpp:1534-1556), b = J x_t."""
    return _wrap(_lib().rsg_chain())


def from_csr(row_ptr, col, val, b, na: int, nv: int, nw: int) -> BlockSystem:
    """Well-Schur-reduce a caller-assembled scalar CSR Jacobian (cells in blocks of nv, then nw BHP rows)."""
    rp = np.ascontiguousarray(row_ptr, np.int32)
    ci = np.ascontiguousarray(col, np.int32)
    v = np.ascontiguousarray(val, np.float64)
    bb = np.ascontiguousarray(b, np.float64)
    return _wrap(_lib().rsg_from_csr(len(rp) - 1, _p(rp, ctypes.c_int), _p(ci, ctypes.c_int), _p(v, ctypes.c_double),
                                     _p(bb, ctypes.c_double), na, nv, nw))
