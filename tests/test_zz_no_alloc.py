"""rsh_solve_device_noalloc: the resident BiCGSTAB solve (rsh_solve_device) for callers that cannot allocate, its
Krylov workspace in the library's static device scratch (64 MiB per GPU).  Through the C-ABI: two calls in a row (the
second on a smaller system) and a third on the first system again, each bit-identical to rsh_solve_device on the
same system and vectors; a system beyond the scratch and a Krylov method it does not run are refused."""
import ctypes

import numpy as np
import pytest

import ressimstudio_amd as rs
from ressimstudio_amd import gen

pytestmark = pytest.mark.gpu

TOL = 1e-8


def _lib():
    return rs.lib()


class _Dev:
    """device copies of host vectors on a context (freed at close)"""

    def __init__(self, ctx):
        self.ctx, self.ptrs = ctx, []

    def put(self, a):
        a = np.ascontiguousarray(a, np.float64)
        p = ctypes.c_void_p()
        rs._check(rs.lib().rsh_device_alloc(self.ctx.h, a.nbytes, ctypes.byref(p)), "alloc")
        rs._check(rs.lib().rsh_memcpy_h2d(self.ctx.h, p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes), "h2d")
        self.ptrs.append(p)
        return p

    def get(self, p, n):
        out = np.empty(n)
        rs._check(rs.lib().rsh_memcpy_d2h(self.ctx.h, out.ctypes.data_as(ctypes.c_void_p), p, out.nbytes), "d2h")
        return out

    def close(self):
        for p in self.ptrs:
            rs.lib().rsh_device_free(self.ctx.h, p)
        self.ptrs = []


def _solve(sysd, dev, b, noalloc):
    L = _lib()
    bp, xp = dev.put(b), dev.put(np.full(b.size, np.nan))
    its, rel = ctypes.c_int(0), ctypes.c_double(0)
    f = L.rsh_solve_device_noalloc if noalloc else L.rsh_solve_device
    rc = f(sysd.h, bp, xp, TOL, 2000, ctypes.byref(its), ctypes.byref(rel))
    return rc, dev.get(xp, b.size), int(its.value), float(rel.value)


@pytest.mark.parametrize("precision", [rs.FP64, rs.MIXED_STORE])
def test_noalloc_solve_equals_resident_solve(orc, precision):
    big = gen.reservoir(30, 24, 8, 3, model=1, sigma=2.0, kvkh=0.1, wells=3, seed=401)
    small = gen.reservoir(14, 12, 5, 3, model=0, sigma=0.7, wells=1, seed=402)
    ctx = rs.default_context()
    dev = _Dev(ctx)
    try:
        systems = [rs.System(s).setup(precond="cpr_mc", smoother=rs.SMOOTH_JACOBI, precision=precision) for s in (big, small)]
        for k, (s, sysd) in enumerate([(big, systems[0]), (small, systems[1]), (big, systems[0])]):
            print(f"no-alloc solve call {k + 1}: {s.n} cells, precision {precision}", flush=True)
            rc0, x0, its0, rel0 = _solve(sysd, dev, s.rhs, noalloc=False)
            rc1, x1, its1, rel1 = _solve(sysd, dev, s.rhs, noalloc=True)
            assert rc0 == rs.OK and rc1 == rs.OK, rs.lib().rsh_last_error()
            assert np.array_equal(x0, x1) and its0 == its1 and rel0 == rel1, k
            r = s.rhs - orc.spmv(s, x1)
            assert np.linalg.norm(r) <= 1.01 * TOL * np.linalg.norm(s.rhs)
    finally:
        dev.close()


def test_noalloc_refuses_what_it_cannot_run():
    ctx = rs.default_context()
    dev = _Dev(ctx)
    try:
        # 80 x 80 x 56 cells x 3 unknowns = 1,075,200 > the 1,047,552 unknowns eight vectors may hold in 64 MiB
        huge = rs.System(gen.reservoir(80, 80, 56, 3, model=0, sigma=1.0, wells=1, seed=403))
        rc, _, _, _ = _solve(huge, dev, np.ones(80 * 80 * 56 * 3), noalloc=True)
        assert rc == rs.ERR_ARG and b"exceed the scratch" in rs.lib().rsh_last_error()
        del huge
        s = gen.small()
        fg = rs.System(s).setup(precond="bilu0", krylov=rs.KRYLOV_FGMRES)
        rc, _, _, _ = _solve(fg, dev, s.rhs, noalloc=True)
        assert rc == rs.ERR_UNSUPPORTED
    finally:
        dev.close()
