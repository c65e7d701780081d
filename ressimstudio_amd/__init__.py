"""ressimstudio_amd - the linear-solve hot path of ResSimStudio's black-oil Newton loop on MI355X (gfx950).

The product is ``libressim_hip.so`` (C-ABI declared in include/ressim_hip.h, HIP kernels in csrc/hip/).  This
module is its Python mirror of the reference's solver interface, for tests, benchmarks and scripting:

  * :class:`BSR`                  the block matrix (ressim::BSR, engine/include/ressim/bsr.h:23-37)
  * :class:`HipBlockSolver`       ressim::GpuBlockSolver (bsr.h:143-151): ``solve(A, b, tol, max_it)`` returns
                                  ``(ok, x, iterations)`` - false on device failure or non-finite results, the
                                  caller then falls back to its CPU path (simulator.cpp:3366, 3376-3377)
  * :class:`System`               a device-resident system: pattern, values, preconditioner and Krylov method
  * :func:`bicgstab`, :func:`gmres`  the Krylov entry points (linsolve.h:99-106) running on the device

There is no CPU fallback: every call goes through the HIP library and raises when it is not built or no GPU is
present.
"""
from __future__ import annotations

import ctypes
import json
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libressim_hip.so")

PREC_BILU0, PREC_MCBILU0, PREC_CPR, PREC_CPR_MC, PREC_CPR_BJ = 1, 2, 3, 4, 5
KRYLOV_BICGSTAB, KRYLOV_FGMRES = 0, 1
FP64, MIXED, MIXED_IR, MIXED_STORE = 0, 1, 2, 3
SMOOTH_GS, SMOOTH_JACOBI = 0, 1
PRECONDITIONERS = {"bilu0": PREC_BILU0, "mcbilu0": PREC_MCBILU0, "cpr": PREC_CPR, "cpr_mc": PREC_CPR_MC, "cpr_bj": PREC_CPR_BJ}

OK, ERR_ARG, ERR_DEVICE, ERR_NOT_CONVERGED, ERR_BREAKDOWN, ERR_NONFINITE, ERR_UNSUPPORTED = 0, -1, -2, -3, -4, -5, -6

EXPORTED = [
    "rsh_version", "rsh_last_error", "rsh_device_count", "rsh_default_options", "rsh_context_create", "rsh_context_destroy",
    "rsh_context_synchronize", "rsh_system_create", "rsh_system_destroy", "rsh_system_set_values", "rsh_system_set_values_device",
    "rsh_system_set_accumulation", "rsh_system_set_accumulation_device", "rsh_system_setup", "rsh_solve", "rsh_solve_device", "rsh_spmv", "rsh_precond_apply", "rsh_amg_apply", "rsh_cpr_weights",
    "rsh_bilu0_factors", "rsh_ilu_order", "rsh_amg_levels", "rsh_system_info", "rsh_device_alloc", "rsh_device_free", "rsh_memcpy_h2d",
    "rsh_memcpy_d2h", "rsh_time_kernel", "rsh_batch_create", "rsh_batch_destroy", "rsh_batch_set_values",
    "rsh_batch_set_values_device", "rsh_batch_setup", "rsh_batch_solve", "rsh_batch_solve_device", "rsh_batch_latencies",
    "rsh_batch_info", "rsh_server_create", "rsh_server_create_batched", "rsh_server_destroy", "rsh_server_add_model", "rsh_server_submit", "rsh_server_wait",
    "rsh_server_stats", "rsh_wells_create", "rsh_wells_destroy", "rsh_wells_system", "rsh_wells_set_values", "rsh_wells_solve",
    "rsh_wells_reduced_values", "rsh_context_device_seconds", "rsh_solve_device_noalloc", "rsh_system_set_spmv_lanes",
    "rsh_batch_set_spmv_lanes",
]


class RshError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Options(ctypes.Structure):
    _fields_ = [("precond", ctypes.c_int), ("krylov", ctypes.c_int), ("precision", ctypes.c_int), ("smoother", ctypes.c_int),
                ("restart", ctypes.c_int), ("max_outer", ctypes.c_int), ("inner_tol", ctypes.c_double), ("reuse_amg", ctypes.c_int),
                ("smooth_steps", ctypes.c_int), ("jacobi_omega", ctypes.c_double), ("ilu_lean", ctypes.c_int),
                ("coarse_smooth_steps", ctypes.c_int), ("ilu_schedule", ctypes.c_int), ("ilu_tile_dims", ctypes.c_int),
                ("retry_fp32", ctypes.c_int), ("ilu_rb_tiles", ctypes.c_int)]


class Request(ctypes.Structure):
    _fields_ = [("model", ctypes.c_int), ("values", ctypes.c_void_p), ("values_on_device", ctypes.c_int), ("perturb", ctypes.c_double),
                ("perturb_seed", ctypes.c_ulonglong), ("rhs", ctypes.c_void_p), ("rhs_on_device", ctypes.c_int), ("x", ctypes.c_void_p),
                ("x_on_device", ctypes.c_int), ("tol", ctypes.c_double), ("max_it", ctypes.c_int)]


class Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("iterations", ctypes.c_int), ("rel_res", ctypes.c_double), ("queue_ms", ctypes.c_double),
                ("service_ms", ctypes.c_double), ("latency_ms", ctypes.c_double)]


_LIB = None
_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


def lib():
    """Load the HIP library (raises when it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.rsh_version.restype = ctypes.c_int
        L.rsh_last_error.restype = ctypes.c_char_p
        L.rsh_device_count.restype = ctypes.c_int
        L.rsh_default_options.argtypes = [ctypes.POINTER(Options)]
        L.rsh_context_create.argtypes = [ctypes.c_int, ctypes.POINTER(_vp)]
        L.rsh_context_destroy.argtypes = [_vp]
        L.rsh_context_synchronize.argtypes = [_vp]
        L.rsh_context_device_seconds.argtypes = [_vp, _dp]
        L.rsh_system_create.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _ip, _ip, ctypes.POINTER(_vp)]
        L.rsh_system_destroy.argtypes = [_vp]
        L.rsh_system_set_values.argtypes = [_vp, _dp]
        L.rsh_system_set_values_device.argtypes = [_vp, _vp]
        L.rsh_system_setup.argtypes = [_vp, ctypes.POINTER(Options)]
        L.rsh_solve.argtypes = [_vp, _dp, _dp, ctypes.c_double, ctypes.c_int, _ip, _dp]
        L.rsh_solve_device.argtypes = [_vp, _vp, _vp, ctypes.c_double, ctypes.c_int, _ip, _dp]
        L.rsh_solve_device_noalloc.argtypes = [_vp, _vp, _vp, ctypes.c_double, ctypes.c_int, _ip, _dp]
        L.rsh_spmv.argtypes = [_vp, _dp, _dp]
        L.rsh_system_set_spmv_lanes.argtypes = [_vp, ctypes.c_int]
        L.rsh_batch_set_spmv_lanes.argtypes = [_vp, ctypes.c_int]
        L.rsh_precond_apply.argtypes = [_vp, _dp, _dp]
        L.rsh_amg_apply.argtypes = [_vp, _dp, _dp]
        L.rsh_cpr_weights.argtypes = [_vp, _dp]
        L.rsh_system_set_accumulation.argtypes = [_vp, _dp]
        L.rsh_system_set_accumulation_device.argtypes = [_vp, _vp]
        L.rsh_bilu0_factors.argtypes = [_vp, _dp]
        L.rsh_ilu_order.argtypes = [_vp, _ip]
        L.rsh_amg_levels.argtypes = [_vp, _ip, _ip, _ip, ctypes.c_int]
        L.rsh_system_info.argtypes = [_vp, ctypes.c_char_p, ctypes.c_int]
        L.rsh_device_alloc.argtypes = [_vp, ctypes.c_size_t, ctypes.POINTER(_vp)]
        L.rsh_device_free.argtypes = [_vp, _vp]
        L.rsh_memcpy_h2d.argtypes = [_vp, _vp, _vp, ctypes.c_size_t]
        L.rsh_memcpy_d2h.argtypes = [_vp, _vp, _vp, ctypes.c_size_t]
        L.rsh_time_kernel.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp]
        _pp = ctypes.POINTER(_vp)
        L.rsh_batch_create.argtypes = [_vp, ctypes.c_int, _ip, _ip, _pp, _pp, ctypes.POINTER(_vp)]
        L.rsh_batch_destroy.argtypes = [_vp]
        L.rsh_batch_set_values.argtypes = [_vp, _pp]
        L.rsh_batch_set_values_device.argtypes = [_vp, _vp]
        L.rsh_batch_setup.argtypes = [_vp, ctypes.POINTER(Options)]
        L.rsh_batch_solve.argtypes = [_vp, _pp, _pp, _dp, ctypes.c_int, _ip, _dp, _ip]
        L.rsh_batch_solve_device.argtypes = [_vp, _vp, _vp, _dp, ctypes.c_int, _ip, _dp, _ip]
        L.rsh_batch_latencies.argtypes = [_vp, _dp]
        L.rsh_batch_info.argtypes = [_vp, ctypes.c_char_p, ctypes.c_int]
        L.rsh_server_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.POINTER(_vp)]
        L.rsh_server_create_batched.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_longlong, ctypes.c_int, ctypes.POINTER(_vp)]
        L.rsh_server_destroy.argtypes = [_vp]
        L.rsh_server_add_model.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _ip, _ip, _dp, ctypes.POINTER(Options), _ip]
        L.rsh_server_submit.argtypes = [_vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_longlong)]
        L.rsh_server_wait.argtypes = [_vp, ctypes.c_longlong, ctypes.POINTER(Result)]
        L.rsh_server_stats.argtypes = [_vp, ctypes.c_char_p, ctypes.c_int]
        L.rsh_wells_create.argtypes = [_vp, ctypes.c_int, _ip, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]
        L.rsh_wells_destroy.argtypes = [_vp]
        L.rsh_wells_system.argtypes = [_vp, ctypes.POINTER(_vp)]
        L.rsh_wells_set_values.argtypes = [_vp, _vp, ctypes.c_int]
        L.rsh_wells_solve.argtypes = [_vp, _vp, _vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _ip, _dp]
        L.rsh_wells_reduced_values.argtypes = [_vp, _dp]
        _LIB = L
    return _LIB


def _check(rc: int, what: str = "") -> int:
    if rc < 0:
        raise RshError(rc, f"{what}: {lib().rsh_last_error().decode()}")
    return rc


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def device_count() -> int:
    return int(lib().rsh_device_count())


def default_options(**kw) -> Options:
    o = Options()
    lib().rsh_default_options(ctypes.byref(o))
    for k, v in kw.items():
        if k == "precond" and isinstance(v, str):
            v = PRECONDITIONERS[v]
        setattr(o, k, v)
    return o


class Context:
    """One HIP stream on one device (re-entrant use: one context per thread)."""

    def __init__(self, device: int = 0):
        h = _vp()
        _check(lib().rsh_context_create(device, ctypes.byref(h)), "rsh_context_create")
        self.h = h

    def synchronize(self):
        _check(lib().rsh_context_synchronize(self.h), "rsh_context_synchronize")

    def close(self):
        if getattr(self, "h", None):
            lib().rsh_context_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


_DEFAULT_CTX = None


def default_context() -> Context:
    global _DEFAULT_CTX
    if _DEFAULT_CTX is None:
        _DEFAULT_CTX = Context(0)
    return _DEFAULT_CTX


@dataclass
class BSR:
    """Block CSR matrix: n block rows, nb x nb row-major blocks, sorted columns per row, diagonal present."""
    n: int
    nb: int
    row_ptr: np.ndarray
    col: np.ndarray
    val: np.ndarray

    @classmethod
    def of(cls, s) -> "BSR":
        return cls(int(s.n), int(s.nb), np.ascontiguousarray(s.row_ptr, np.int32), np.ascontiguousarray(s.col, np.int32),
                   np.ascontiguousarray(s.val, np.float64).reshape(-1))

    @property
    def nnzb(self) -> int:
        return int(self.row_ptr[-1])

    def size(self) -> int:
        return self.n * self.nb


class System:
    """A device-resident block system (pattern + values + preconditioner), the unit the solver keeps across
    Newton iterations."""

    def __init__(self, A, ctx: Context | None = None):
        A = A if isinstance(A, BSR) else BSR.of(A)
        self.ctx = ctx or default_context()
        self.n, self.nb, self.N = A.n, A.nb, A.n * A.nb
        self.row_ptr, self.col = A.row_ptr, A.col
        h = _vp()
        _check(lib().rsh_system_create(self.ctx.h, A.n, A.nb, _p(A.row_ptr, ctypes.c_int), _p(A.col, ctypes.c_int), ctypes.byref(h)),
               "rsh_system_create")
        self.h = h
        self.set_values(A.val)
        self.options = None

    def close(self):
        if getattr(self, "h", None):
            lib().rsh_system_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_values(self, val):
        v = np.ascontiguousarray(val, np.float64).reshape(-1)
        _check(lib().rsh_system_set_values(self.h, _p(v, ctypes.c_double)), "rsh_system_set_values")

    def setup(self, options: Options | None = None, **kw):
        o = options or default_options(**kw)
        _check(lib().rsh_system_setup(self.h, ctypes.byref(o)), "rsh_system_setup")
        self.options = o
        return self

    def solve(self, b, tol=1e-8, max_it=1000, raise_on_fail=True):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.N)
        its = ctypes.c_int(0)
        rel = ctypes.c_double(0)
        rc = lib().rsh_solve(self.h, _p(b, ctypes.c_double), _p(x, ctypes.c_double), tol, max_it, ctypes.byref(its), ctypes.byref(rel))
        if rc < 0 and (raise_on_fail or rc not in (ERR_NOT_CONVERGED, ERR_BREAKDOWN)):
            _check(rc, "rsh_solve")
        return x, int(its.value), float(rel.value), rc == OK

    def spmv(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty(self.N)
        _check(lib().rsh_spmv(self.h, _p(x, ctypes.c_double), _p(y, ctypes.c_double)), "rsh_spmv")
        return y

    def set_spmv_lanes(self, lanes: int):
        """Rows per wavefront of the FP64 SpMV: 1, 2 or 4 lanes per row (64, 32 or 16 rows)."""
        _check(lib().rsh_system_set_spmv_lanes(self.h, int(lanes)), "rsh_system_set_spmv_lanes")
        return self

    def precond(self, r):
        r = np.ascontiguousarray(r, np.float64)
        z = np.empty(self.N)
        _check(lib().rsh_precond_apply(self.h, _p(r, ctypes.c_double), _p(z, ctypes.c_double)), "rsh_precond_apply")
        return z

    def amg_apply(self, r):
        r = np.ascontiguousarray(r, np.float64)
        z = np.empty(self.n)
        _check(lib().rsh_amg_apply(self.h, _p(r, ctypes.c_double), _p(z, ctypes.c_double)), "rsh_amg_apply")
        return z

    def set_accumulation(self, acc):
        """true-IMPES CPR weights from n accumulation blocks (None: quasi-IMPES again); used by the next setup()"""
        if acc is None:
            _check(lib().rsh_system_set_accumulation(self.h, None), "rsh_system_set_accumulation")
            return self
        a = np.ascontiguousarray(acc, np.float64).reshape(-1)
        if a.size != self.n * self.nb * self.nb:
            raise ValueError(f"accumulation blocks: {a.size} values, expected {self.n * self.nb * self.nb}")
        _check(lib().rsh_system_set_accumulation(self.h, _p(a, ctypes.c_double)), "rsh_system_set_accumulation")
        return self

    def cpr_weights(self):
        w = np.empty(self.N)
        _check(lib().rsh_cpr_weights(self.h, _p(w, ctypes.c_double)), "rsh_cpr_weights")
        return w

    def bilu0_factors(self):
        lu = np.empty(int(self.row_ptr[-1]) * self.nb * self.nb)
        _check(lib().rsh_bilu0_factors(self.h, _p(lu, ctypes.c_double)), "rsh_bilu0_factors")
        return lu

    def ilu_order(self):
        perm = np.empty(self.n, np.int32)
        _check(lib().rsh_ilu_order(self.h, _p(perm, ctypes.c_int)), "rsh_ilu_order")
        return perm

    def info(self) -> dict:
        buf = ctypes.create_string_buffer(1 << 16)
        _check(lib().rsh_system_info(self.h, buf, len(buf)), "rsh_system_info")
        return json.loads(buf.value.decode())

    def time_kernel(self, kind: int, reps: int = 20):
        ms = ctypes.c_double(0)
        by = ctypes.c_double(0)
        _check(lib().rsh_time_kernel(self.h, kind, reps, ctypes.byref(ms), ctypes.byref(by)), "rsh_time_kernel")
        return float(ms.value), float(by.value)


def _ptrs(arrays):
    return (_vp * len(arrays))(*[a.ctypes.data for a in arrays])


class Batch:
    """Many independent block systems solved together (rsh_batch_*): one block-diagonal resident system per block
    size, per-system BiCGSTAB scalars and convergence, one captured graph per BiCGSTAB iteration."""

    def __init__(self, systems, ctx: Context | None = None):
        mats = [A if isinstance(A, BSR) else BSR.of(A) for A in systems]
        self.ctx = ctx or default_context()
        self.mats = mats
        self.sizes = [A.n * A.nb for A in mats]
        n = np.array([A.n for A in mats], np.int32)
        nb = np.array([A.nb for A in mats], np.int32)
        self._rp = [A.row_ptr for A in mats]
        self._ci = [A.col for A in mats]
        h = _vp()
        _check(lib().rsh_batch_create(self.ctx.h, len(mats), _p(n, ctypes.c_int), _p(nb, ctypes.c_int), _ptrs(self._rp), _ptrs(self._ci),
                                      ctypes.byref(h)), "rsh_batch_create")
        self.h = h
        self.set_values([A.val for A in mats])

    def close(self):
        if getattr(self, "h", None):
            lib().rsh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_values(self, vals):
        vs = [np.ascontiguousarray(v, np.float64).reshape(-1) for v in vals]
        _check(lib().rsh_batch_set_values(self.h, _ptrs(vs)), "rsh_batch_set_values")

    def setup(self, options: Options | None = None, **kw):
        o = options or default_options(**{"precond": "mcbilu0", **kw})
        _check(lib().rsh_batch_setup(self.h, ctypes.byref(o)), "rsh_batch_setup")
        self.options = o
        return self

    def solve(self, bs, tol=1e-8, max_it=1000, raise_on_fail=True):
        """Returns (xs, iterations, rel_res, status) per system."""
        bs = [np.ascontiguousarray(b, np.float64) for b in bs]
        xs = [np.zeros(k) for k in self.sizes]
        m = len(bs)
        tols = np.full(m, tol) if np.isscalar(tol) else np.ascontiguousarray(tol, np.float64)
        its = np.zeros(m, np.int32)
        rel = np.zeros(m)
        st = np.zeros(m, np.int32)
        rc = lib().rsh_batch_solve(self.h, _ptrs(bs), _ptrs(xs), _p(tols, ctypes.c_double), max_it, _p(its, ctypes.c_int), _p(rel, ctypes.c_double),
                                   _p(st, ctypes.c_int))
        if rc < 0 and (raise_on_fail or rc not in (ERR_NOT_CONVERGED, ERR_BREAKDOWN, ERR_NONFINITE)):
            _check(rc, "rsh_batch_solve")
        return xs, its, rel, st

    def set_spmv_lanes(self, lanes: int):
        """SpMV lanes per row of every group from the next setup: 0 = autotuned (default), or 1, 2, 4."""
        _check(lib().rsh_batch_set_spmv_lanes(self.h, int(lanes)), "rsh_batch_set_spmv_lanes")
        return self

    def latencies(self):
        ms = np.zeros(len(self.sizes))
        _check(lib().rsh_batch_latencies(self.h, _p(ms, ctypes.c_double)), "rsh_batch_latencies")
        return ms

    def info(self) -> dict:
        buf = ctypes.create_string_buffer(1 << 16)
        _check(lib().rsh_batch_info(self.h, buf, len(buf)), "rsh_batch_info")
        return json.loads(buf.value.decode())


def perturbed_values(base, row_ptr, amp: float, seed: int) -> np.ndarray:
    """The values a served request with ``perturb = amp`` and ``perturb_seed = seed`` solves with: block row i of
    the BSR values scaled by 1 + amp * u_i, u_i uniform in [-1, 1) from a SplitMix64 hash of (seed, i) (the
    server's kPerturbRows)."""
    rp = np.asarray(row_ptr, np.int64)
    n = rp.size - 1
    b = np.ascontiguousarray(base, np.float64).reshape(int(rp[-1]), -1)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) + np.arange(n, dtype=np.uint64) + np.uint64(1)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0) * 2.0 - 1.0
    f = 1.0 + amp * u
    return (b * np.repeat(f, np.diff(rp))[:, None]).reshape(-1)


class Server:
    """Serving loop (rsh_server_*): resident models, FIFO admission by worker slot, model availability and HBM
    budget, one HIP stream per slot."""

    def __init__(self, slots: int = 4, hbm_budget: float = 200e9, device: int = 0, batched: bool = False, group_cells: int = 200000,
                 group_systems: int = 16):
        """batched: requests served in batched rounds through the batch engine (rsh_server_create_batched): models grouped
        at the first request (group_cells cells / group_systems models at most per group)."""
        h = _vp()
        if batched:
            _check(lib().rsh_server_create_batched(device, slots, float(hbm_budget), int(group_cells), int(group_systems), ctypes.byref(h)),
                   "rsh_server_create_batched")
        else:
            _check(lib().rsh_server_create(device, slots, float(hbm_budget), ctypes.byref(h)), "rsh_server_create")
        self.h = h
        self._keep = {}
        self._models = []

    def add_model(self, A, options: Options | None = None, **kw) -> int:
        A = A if isinstance(A, BSR) else BSR.of(A)
        o = options or default_options(**kw)
        mid = ctypes.c_int(0)
        _check(lib().rsh_server_add_model(self.h, A.n, A.nb, _p(A.row_ptr, ctypes.c_int), _p(A.col, ctypes.c_int), _p(A.val, ctypes.c_double),
                                          ctypes.byref(o), ctypes.byref(mid)), "rsh_server_add_model")
        self._models.append(A)
        return int(mid.value)

    def submit(self, model: int, rhs, *, values=None, perturb: float = 0.0, seed: int = 0, tol: float = 1e-8, max_it: int = 1000,
               want_x: bool = True) -> int:
        rhs = np.ascontiguousarray(rhs, np.float64)
        x = np.zeros(rhs.size) if want_x else None
        vals = None if values is None else np.ascontiguousarray(values, np.float64).reshape(-1)
        r = Request(model=model, values=None if vals is None else vals.ctypes.data, values_on_device=0, perturb=perturb, perturb_seed=seed,
                    rhs=rhs.ctypes.data, rhs_on_device=0, x=None if x is None else x.ctypes.data, x_on_device=0, tol=tol, max_it=max_it)
        t = ctypes.c_longlong(0)
        _check(lib().rsh_server_submit(self.h, ctypes.byref(r), ctypes.byref(t)), "rsh_server_submit")
        self._keep[int(t.value)] = (rhs, x, vals)
        return int(t.value)

    def wait(self, ticket: int) -> dict:
        res = Result()
        _check(lib().rsh_server_wait(self.h, ticket, ctypes.byref(res)), "rsh_server_wait")
        rhs, x, vals = self._keep.pop(ticket)
        return {"status": res.status, "iterations": res.iterations, "rel_res": res.rel_res, "queue_ms": res.queue_ms,
                "service_ms": res.service_ms, "latency_ms": res.latency_ms, "x": x}

    def stats(self) -> dict:
        buf = ctypes.create_string_buffer(1 << 16)
        _check(lib().rsh_server_stats(self.h, buf, len(buf)), "rsh_server_stats")
        return json.loads(buf.value.decode())

    def close(self):
        if getattr(self, "h", None):
            lib().rsh_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Wells:
    """Well Schur reduction on the device (rsh_wells_*): the engine's scalar Jacobian (cells in blocks of nv, then
    one BHP row per well) in, cells and BHPs out; ``setup`` / ``info`` act on the reduced resident system."""

    def __init__(self, row_ptr, col, na: int, nv: int, nw: int, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.row_ptr = np.ascontiguousarray(row_ptr, np.int32)
        self.col = np.ascontiguousarray(col, np.int32)
        self.n = len(self.row_ptr) - 1
        h = _vp()
        _check(lib().rsh_wells_create(self.ctx.h, self.n, _p(self.row_ptr, ctypes.c_int), _p(self.col, ctypes.c_int), na, nv, nw, ctypes.byref(h)),
               "rsh_wells_create")
        self.h = h
        sh = _vp()
        _check(lib().rsh_wells_system(self.h, ctypes.byref(sh)), "rsh_wells_system")
        self.sys_h = sh

    def close(self):
        if getattr(self, "h", None):
            lib().rsh_wells_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_values(self, jval):
        v = np.ascontiguousarray(jval, np.float64)
        _check(lib().rsh_wells_set_values(self.h, v.ctypes.data, 0), "rsh_wells_set_values")

    def setup(self, options: Options | None = None, **kw):
        o = options or default_options(**kw)
        _check(lib().rsh_system_setup(self.sys_h, ctypes.byref(o)), "rsh_system_setup")
        return self

    def solve(self, b_full, tol=1e-8, max_it=1000):
        b = np.ascontiguousarray(b_full, np.float64)
        x = np.zeros(self.n)
        its = ctypes.c_int(0)
        rel = ctypes.c_double(0)
        rc = lib().rsh_wells_solve(self.h, b.ctypes.data, x.ctypes.data, 0, tol, max_it, ctypes.byref(its), ctypes.byref(rel))
        _check(rc, "rsh_wells_solve")
        return x, int(its.value), float(rel.value)

    def reduced_values(self, nnzb: int, nb: int):
        out = np.empty(nnzb * nb * nb)
        _check(lib().rsh_wells_reduced_values(self.h, _p(out, ctypes.c_double)), "rsh_wells_reduced_values")
        return out


class HipBlockSolver:
    """ressim::GpuBlockSolver on the device (bsr.h:143-151).  Systems are cached by sparsity pattern so a Newton
    loop only uploads new values; the AMG hierarchy is refreshed (amgReuse semantics) while the pattern holds."""

    def __init__(self, precond: str = "cpr", precision: int = FP64, krylov: int = KRYLOV_BICGSTAB, smoother: int = SMOOTH_GS,
                 ctx: Context | None = None, reuse_amg: bool = True):
        self.ctx = ctx or default_context()
        self.opts = default_options(precond=precond, precision=precision, krylov=krylov, smoother=smoother, reuse_amg=int(reuse_amg))
        self._sys = {}
        self.device_solves = 0
        self.failures = 0

    def name(self) -> str:
        return "MI355X HIP block solver"

    def solve(self, A, b, tol: float, max_it: int):
        A = A if isinstance(A, BSR) else BSR.of(A)
        key = (A.n, A.nb, A.row_ptr.tobytes(), A.col.tobytes())
        try:
            s = self._sys.get(key)
            if s is None:
                s = System(A, self.ctx)
                self._sys[key] = s
            else:
                s.set_values(A.val)
            s.setup(self.opts)
            x, its, rel, ok = s.solve(b, tol, max_it, raise_on_fail=False)
            self.device_solves += 1
            if not ok or not np.all(np.isfinite(x)):
                self.failures += 1
                return False, x, its
            return True, x, its
        except RshError:
            self.failures += 1
            return False, np.zeros(A.size()), 0


def bicgstab(A, b, precond: str = "bilu0", tol: float = 1e-8, max_it: int = 1000, **kw):
    s = System(A).setup(precond=precond, krylov=KRYLOV_BICGSTAB, **kw)
    return s.solve(b, tol, max_it, raise_on_fail=False)


def gmres(A, b, precond: str = "bilu0", tol: float = 1e-8, max_it: int = 1000, restart: int = 30, **kw):
    s = System(A).setup(precond=precond, krylov=KRYLOV_FGMRES, restart=restart, **kw)
    return s.solve(b, tol, max_it, raise_on_fail=False)
