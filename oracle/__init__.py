"""ORACLE - test infrastructure only.

ctypes access to the two checker libraries that implement oracle/rs_check.h:
  * ``liboracle.so``            the clean-room CPU restatement of the reference path (oracle.cpp)
  * ``_ref/libressim_ref.so``   the reference's own sources compiled where they lie (built only where the
                                reference is mounted; the built file travels with the tree)
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this package, and only as the checker.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libressim_ref.so")
REF_FAST_LIB = os.path.join(HERE, "_ref", "libressim_ref_fast.so")

PREC = {"bilu0": 1, "bilu0_f32factors": 2, "bilu0f": 3, "mcbilu0": 4, "cpr": 5, "bilu0_lean": 6, "cpr_lean": 7}
METHOD = {"bicgstab": 0, "gmres": 1}

_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def _bsr_args(sys_or_parts):
    s = sys_or_parts
    rp = np.ascontiguousarray(s.row_ptr, np.int32)
    ci = np.ascontiguousarray(s.col, np.int32)
    v = np.ascontiguousarray(s.val, np.float64).reshape(-1)
    return s.n, s.nb, rp, ci, v


class Checker:
    """One checker library (oracle restatement or compiled reference)."""

    def __init__(self, path: str):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.path = path
        lib = ctypes.CDLL(path)
        bsr = [ctypes.c_int, ctypes.c_int, _ip, _ip, _dp]
        lib.chk_spmv.argtypes = bsr + [_dp, _dp]
        lib.chk_bilu0_new.restype = _vp
        lib.chk_bilu0_new.argtypes = bsr + [ctypes.c_int]
        lib.chk_bilu0_apply.argtypes = [_vp, _dp, _dp]
        lib.chk_bilu0_regularised.argtypes = [_vp]
        lib.chk_bilu0_free.argtypes = [_vp]
        lib.chk_mcbilu0_new.restype = _vp
        lib.chk_mcbilu0_new.argtypes = bsr
        lib.chk_mcbilu0_apply.argtypes = [_vp, _dp, _dp]
        lib.chk_mcbilu0_colours.argtypes = [_vp]
        lib.chk_mcbilu0_colouring.argtypes = [_vp, _ip, _ip]
        lib.chk_mcbilu0_free.argtypes = [_vp]
        lib.chk_cpr_new.restype = _vp
        lib.chk_cpr_new.argtypes = bsr
        lib.chk_cpr_new_acc.restype = _vp
        lib.chk_cpr_new_acc.argtypes = bsr + [_dp]
        lib.chk_cpr_fallbacks.argtypes = [_vp]
        lib.chk_cpr_refresh.argtypes = [_vp, _dp]
        lib.chk_cpr_apply.argtypes = [_vp, _dp, _dp]
        lib.chk_cpr_weights.argtypes = [_vp, _dp]
        lib.chk_cpr_nlevels.argtypes = [_vp]
        lib.chk_cpr_level.argtypes = [_vp, ctypes.c_int, _ip, _ip]
        lib.chk_cpr_level_matrix.argtypes = [_vp, ctypes.c_int, _ip, _ip, _dp]
        lib.chk_cpr_aggregates.argtypes = [_vp, ctypes.c_int, _ip]
        lib.chk_cpr_amg_apply.argtypes = [_vp, _dp, _dp]
        lib.chk_cpr_free.argtypes = [_vp]
        lib.chk_devcpr_new.restype = _vp
        lib.chk_devcpr_new.argtypes = bsr
        lib.chk_devcpr_apply.argtypes = [_vp, _dp, _dp]
        lib.chk_devcpr_free.argtypes = [_vp]
        lib.chk_krylov.argtypes = bsr + [_dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, _ip, _dp]
        lib.chk_krylov_cpr_steady.argtypes = bsr + [_dp, _dp, ctypes.c_double, ctypes.c_int, _ip, _dp, _dp]
        if hasattr(lib, "chk_bilu0_factors"):
            lib.chk_bilu0_factors.argtypes = [_vp, _dp]
        if hasattr(lib, "chk_set_threads"):
            lib.chk_set_threads.argtypes = [ctypes.c_int]
        if hasattr(lib, "chk_float_solve"):
            lib.chk_float_solve.argtypes = bsr + [_dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _ip]
            lib.chk_refine_float.argtypes = bsr + [_dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, _ip, _dp]
        if hasattr(lib, "chk_wellschur_new"):
            lib.chk_wellschur_new.restype = _vp
            lib.chk_wellschur_new.argtypes = [ctypes.c_int, _ip, _ip, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip]
            lib.chk_wellschur_reduced.argtypes = [_vp, _ip, _ip, _dp]
            lib.chk_wellschur_reduce_rhs.argtypes = [_vp, _dp, _dp]
            lib.chk_wellschur_recover.argtypes = [_vp, _dp, _dp, _dp]
            lib.chk_wellschur_free.argtypes = [_vp]
        self.lib = lib

    # --- kernels
    def spmv(self, s, x):
        n, nb, rp, ci, v = _bsr_args(s)
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty(n * nb)
        self.lib.chk_spmv(n, nb, _p(rp, ctypes.c_int), _p(ci, ctypes.c_int), _p(v, ctypes.c_double), _p(x, ctypes.c_double),
                          _p(y, ctypes.c_double))
        return y

    def krylov(self, s, b, precond="bilu0", method="bicgstab", tol=1e-8, maxit=1000, restart=40):
        n, nb, rp, ci, v = _bsr_args(s)
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(n * nb)
        its = ctypes.c_int(0)
        rel = ctypes.c_double(0)
        conv = self.lib.chk_krylov(n, nb, _p(rp, ctypes.c_int), _p(ci, ctypes.c_int), _p(v, ctypes.c_double), _p(b, ctypes.c_double),
                                   _p(x, ctypes.c_double), PREC[precond], METHOD[method], restart, tol, maxit, ctypes.byref(its),
                                   ctypes.byref(rel))
        if conv < 0:
            raise RuntimeError("checker krylov failed")
        return x, int(its.value), float(rel.value), bool(conv)

    def krylov_cpr_steady(self, s, b, tol=1e-8, maxit=1000):
        """CPR-BiCGSTAB as in the Newton loop after the first iteration: the CPR set-up with the AMG aggregates
        reused, then the solve; returns x, iterations, rel.res, converged, set-up seconds, solve seconds."""
        n, nb, rp, ci, v = _bsr_args(s)
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(n * nb)
        its = ctypes.c_int(0)
        rel = ctypes.c_double(0)
        secs = np.zeros(2)
        conv = self.lib.chk_krylov_cpr_steady(n, nb, _p(rp, ctypes.c_int), _p(ci, ctypes.c_int), _p(v, ctypes.c_double),
                                              _p(b, ctypes.c_double), _p(x, ctypes.c_double), tol, maxit, ctypes.byref(its),
                                              ctypes.byref(rel), _p(secs, ctypes.c_double))
        return x, int(its.value), float(rel.value), bool(conv), float(secs[0]), float(secs[1])

    FLOAT_KIND = {"bilu0f": 0, "mcbilu0": 1, "devcpr": 2}

    def float_solve(self, s, b, kind="bilu0f", tol=1e-4, maxit=1000):
        """The float32 device solve (CpuFloatBlockSolver::solve, bsr.cpp:1130-1215): (x, iterations, finite)."""
        n, nb, rp, ci, v = _bsr_args(s)
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(n * nb)
        its = ctypes.c_int(0)
        rc = self.lib.chk_float_solve(n, nb, _p(rp, ctypes.c_int), _p(ci, ctypes.c_int), _p(v, ctypes.c_double), _p(b, ctypes.c_double),
                                      _p(x, ctypes.c_double), self.FLOAT_KIND[kind], tol, maxit, ctypes.byref(its))
        if rc < 0:
            raise RuntimeError("float solve set-up failed")
        return x, int(its.value), bool(rc)

    def refine_float(self, s, b, kind="bilu0f", tol=1e-8, max_outer=10, inner_it=500):
        """FP64 refinement around the float device solve (refineWithDevice, bsr.cpp:1373-1396):
        (x, total inner iterations, relative residual, converged)."""
        n, nb, rp, ci, v = _bsr_args(s)
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(n * nb)
        its = ctypes.c_int(0)
        rel = ctypes.c_double(0)
        rc = self.lib.chk_refine_float(n, nb, _p(rp, ctypes.c_int), _p(ci, ctypes.c_int), _p(v, ctypes.c_double), _p(b, ctypes.c_double),
                                       _p(x, ctypes.c_double), self.FLOAT_KIND[kind], tol, max_outer, inner_it, ctypes.byref(its), ctypes.byref(rel))
        if rc < 0:
            raise RuntimeError("float solve set-up failed")
        return x, int(its.value), float(rel.value), bool(rc)

    def set_threads(self, n: int) -> int:
        return int(self.lib.chk_set_threads(n)) if hasattr(self.lib, "chk_set_threads") else 1

    def bilu0(self, s, mode=1):
        return _Bilu0(self, s, mode)

    def mcbilu0(self, s):
        return _McBilu0(self, s)

    def cpr(self, s, acc=None):
        """BlockCPR (bsr.cpp:993-1084); acc: n accumulation blocks for true-IMPES weights"""
        return _Cpr(self, s, acc)

    def wellschur(self, row_ptr, col, val, na, nv, nw):
        """WellSchur (bsr.cpp:733-990) of a scalar Jacobian (compiled reference only); None when it refuses."""
        if not hasattr(self.lib, "chk_wellschur_new"):
            raise NotImplementedError("this checker has no well Schur reduction")
        w = _WellSchur(self, row_ptr, col, val, na, nv, nw)
        return w if w.h else None

    def devcpr(self, s):
        """DeviceCprAmg (bsr.cpp:1235-1371): float32 CPR with damped-Jacobi AMG and block-Jacobi second stage."""
        return _DevCpr(self, s)


class _Handle:
    _free = None

    def __del__(self):
        if getattr(self, "h", None):
            getattr(self.c.lib, self._free)(self.h)
            self.h = None


class _Bilu0(_Handle):
    _free = "chk_bilu0_free"

    def __init__(self, c, s, mode):
        self.c, self.N, self.nb, self.nnzb = c, s.n * s.nb, s.nb, int(s.row_ptr[-1])
        n, nb, rp, ci, v = _bsr_args(s)
        self.h = c.lib.chk_bilu0_new(n, nb, _p(rp, ctypes.c_int), _p(ci, ctypes.c_int), _p(v, ctypes.c_double), mode)
        if not self.h:
            raise RuntimeError("bilu0 factorisation unavailable in this mode")

    def apply(self, r):
        r = np.ascontiguousarray(r, np.float64)
        z = np.empty(self.N)
        self.c.lib.chk_bilu0_apply(self.h, _p(r, ctypes.c_double), _p(z, ctypes.c_double))
        return z

    @property
    def regularised(self):
        return int(self.c.lib.chk_bilu0_regularised(self.h))

    def factors(self):
        if not hasattr(self.c.lib, "chk_bilu0_factors"):
            raise NotImplementedError("this checker does not expose its factors")
        lu = np.empty(self.nnzb * self.nb * self.nb)
        self.c.lib.chk_bilu0_factors(self.h, _p(lu, ctypes.c_double))
        return lu


class _McBilu0(_Handle):
    _free = "chk_mcbilu0_free"

    def __init__(self, c, s):
        self.c, self.N, self.n = c, s.n * s.nb, s.n
        n, nb, rp, ci, v = _bsr_args(s)
        self.h = c.lib.chk_mcbilu0_new(n, nb, _p(rp, ctypes.c_int), _p(ci, ctypes.c_int), _p(v, ctypes.c_double))

    def apply(self, r):
        r = np.ascontiguousarray(r, np.float64)
        z = np.empty(self.N)
        self.c.lib.chk_mcbilu0_apply(self.h, _p(r, ctypes.c_double), _p(z, ctypes.c_double))
        return z

    def colouring(self):
        colour = np.empty(self.n, np.int32)
        perm = np.empty(self.n, np.int32)
        k = self.c.lib.chk_mcbilu0_colouring(self.h, _p(colour, ctypes.c_int), _p(perm, ctypes.c_int))
        return int(k), colour, perm


class _DevCpr(_Handle):
    _free = "chk_devcpr_free"

    def __init__(self, c, s):
        self.c, self.N = c, s.n * s.nb
        n, nb, rp, ci, v = _bsr_args(s)
        self.h = c.lib.chk_devcpr_new(n, nb, _p(rp, ctypes.c_int), _p(ci, ctypes.c_int), _p(v, ctypes.c_double))
        if not self.h:
            raise RuntimeError("DeviceCprAmg build failed")

    def apply(self, r):
        r = np.ascontiguousarray(r, np.float64)
        z = np.empty(self.N)
        self.c.lib.chk_devcpr_apply(self.h, _p(r, ctypes.c_double), _p(z, ctypes.c_double))
        return z


class _WellSchur(_Handle):
    _free = "chk_wellschur_free"

    def __init__(self, c, row_ptr, col, val, na, nv, nw):
        self.c = c
        self._rp = np.ascontiguousarray(row_ptr, np.int32)
        self._ci = np.ascontiguousarray(col, np.int32)
        self._v = np.ascontiguousarray(val, np.float64)
        self.n_full, self.nc, self.nv = len(self._rp) - 1, na * nv, nv
        sizes = np.zeros(3, np.int32)
        self.h = c.lib.chk_wellschur_new(self.n_full, _p(self._rp, ctypes.c_int), _p(self._ci, ctypes.c_int), _p(self._v, ctypes.c_double),
                                         na, nv, nw, _p(sizes, ctypes.c_int))
        self.n, self.nnzb, self.fill = int(sizes[0]), int(sizes[1]), int(sizes[2])

    def reduced(self):
        rp = np.empty(self.n + 1, np.int32)
        ci = np.empty(self.nnzb, np.int32)
        v = np.empty(self.nnzb * self.nv * self.nv)
        self.c.lib.chk_wellschur_reduced(self.h, _p(rp, ctypes.c_int), _p(ci, ctypes.c_int), _p(v, ctypes.c_double))
        return rp, ci, v

    def reduce_rhs(self, b):
        b = np.ascontiguousarray(b, np.float64)
        bc = np.empty(self.nc)
        self.c.lib.chk_wellschur_reduce_rhs(self.h, _p(b, ctypes.c_double), _p(bc, ctypes.c_double))
        return bc

    def recover(self, b, xc):
        b = np.ascontiguousarray(b, np.float64)
        xc = np.ascontiguousarray(xc, np.float64)
        x = np.empty(self.n_full)
        self.c.lib.chk_wellschur_recover(self.h, _p(b, ctypes.c_double), _p(xc, ctypes.c_double), _p(x, ctypes.c_double))
        return x


class _Cpr(_Handle):
    _free = "chk_cpr_free"

    def __init__(self, c, s, acc=None):
        self.c, self.N, self.n, self.nb = c, s.n * s.nb, s.n, s.nb
        n, nb, rp, ci, v = _bsr_args(s)
        if acc is None:
            self.h = c.lib.chk_cpr_new(n, nb, _p(rp, ctypes.c_int), _p(ci, ctypes.c_int), _p(v, ctypes.c_double))
        else:
            self._acc = np.ascontiguousarray(acc, np.float64).reshape(-1)
            self.h = c.lib.chk_cpr_new_acc(n, nb, _p(rp, ctypes.c_int), _p(ci, ctypes.c_int), _p(v, ctypes.c_double),
                                           _p(self._acc, ctypes.c_double))

    def fallbacks(self):
        return int(self.c.lib.chk_cpr_fallbacks(self.h))

    def refresh(self, val):
        v = np.ascontiguousarray(val, np.float64).reshape(-1)
        return int(self.c.lib.chk_cpr_refresh(self.h, _p(v, ctypes.c_double)))

    def apply(self, r):
        r = np.ascontiguousarray(r, np.float64)
        z = np.empty(self.N)
        self.c.lib.chk_cpr_apply(self.h, _p(r, ctypes.c_double), _p(z, ctypes.c_double))
        return z

    def amg_apply(self, r):
        r = np.ascontiguousarray(r, np.float64)
        z = np.empty(self.n)
        self.c.lib.chk_cpr_amg_apply(self.h, _p(r, ctypes.c_double), _p(z, ctypes.c_double))
        return z

    def weights(self):
        w = np.empty(self.N)
        self.c.lib.chk_cpr_weights(self.h, _p(w, ctypes.c_double))
        return w

    def levels(self):
        out = []
        for l in range(int(self.c.lib.chk_cpr_nlevels(self.h))):
            n, nnz = ctypes.c_int(), ctypes.c_int()
            self.c.lib.chk_cpr_level(self.h, l, ctypes.byref(n), ctypes.byref(nnz))
            out.append((n.value, nnz.value))
        return out

    def level_matrix(self, l):
        n, nnz = self.levels()[l]
        rp = np.empty(n + 1, np.int32)
        ci = np.empty(nnz, np.int32)
        v = np.empty(nnz)
        self.c.lib.chk_cpr_level_matrix(self.h, l, _p(rp, ctypes.c_int), _p(ci, ctypes.c_int), _p(v, ctypes.c_double))
        return rp, ci, v

    def aggregates(self, l):
        n, _ = self.levels()[l]
        agg = np.empty(n, np.int32)
        nc = self.c.lib.chk_cpr_aggregates(self.h, l, _p(agg, ctypes.c_int))
        return int(nc), agg


def oracle() -> Checker:
    return Checker(ORACLE_LIB)


def reference(fast: bool = False) -> Checker | None:
    """The compiled reference, or None where it was never built."""
    path = REF_FAST_LIB if fast else REF_LIB
    return Checker(path) if os.path.exists(path) else None
