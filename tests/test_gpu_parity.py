"""Parity of the HIP path (through the C-ABI) with the oracle restatement, on seeded systems small enough for the
oracle to finish in seconds.  Kernel-level outputs that keep the reference's order of operations are compared
bit for bit; Krylov solves by iteration count and by solution error against the FP64 oracle solution
(<= 1e-6 relative, the path's tolerance), with the true relative residual <= the requested 1e-8."""
import numpy as np
import pytest

import ressimstudio_amd as rs
from ressimstudio_amd import gen

pytestmark = pytest.mark.gpu

TOL = 1e-8


def systems():
    return {
        "small": gen.small(),
        "laplace": gen.block_laplace(40, 30, 3, 11),
        "chain": gen.chain(),
        "nb2": gen.reservoir(12, 10, 3, 2, model=0, sigma=1.5, wells=2, seed=5),
        "nb4": gen.reservoir(10, 9, 3, 4, model=0, sigma=0.8, wells=1, seed=6),
        "nb1": gen.reservoir(16, 12, 4, 1, model=0, sigma=1.0, wells=2, seed=8),
        "layered": gen.reservoir(12, 30, 14, 3, model=1, sigma=2.5, kvkh=0.1, wells=5, seed=7),
        # tall wells: the Schur-reduced well rows couple 40-60 perforated cells all-to-all, so the multicolour
        # order ends in a long run of tiny colours (one cell per well) swept through dense groups
        "deepwell": gen.reservoir(24, 24, 60, 3, model=1, sigma=2.0, kvkh=0.1, wells=3, seed=9),
        "deepwell_nb2": gen.reservoir(12, 12, 40, 2, model=0, sigma=1.0, wells=2, seed=10),
        "deepwell_nb4": gen.reservoir(10, 10, 48, 4, model=0, sigma=1.0, wells=2, seed=11),
        # 60,000 cells: the first coarse AMG level is large enough for its own row-slice layout
        "medium": gen.reservoir(60, 50, 20, 3, model=1, sigma=2.5, kvkh=0.1, wells=4, seed=12),
    }


@pytest.fixture(scope="module")
def sysd():
    return systems()


def _true_rel(s, x, b):
    import oracle
    r = b - oracle.oracle().spmv(s, x)
    return np.linalg.norm(r) / np.linalg.norm(b)


def test_device_present():
    assert rs.device_count() >= 1


@pytest.mark.parametrize("name", ["small", "laplace", "chain", "nb2", "nb4", "nb1", "layered"])
def test_spmv_matches_oracle(sysd, orc, name):
    s = sysd[name]
    d = rs.System(s)
    x = np.random.default_rng(1).standard_normal(s.n * s.nb)
    y = d.spmv(x)
    yo = orc.spmv(s, x)
    info = d.info()
    assert info["spmv_col16"] == 1     # these grids fit the 16-bit relative columns
    if info["spmv_long_rows"] == 0:
        assert np.array_equal(y, yo), "rows handled by one lane keep the reference's order: bit-identical"
    assert np.max(np.abs(y - yo)) <= 1e-13 * np.max(np.abs(yo))


# natural-order sweeps: one launch per level, or tiled (one launch per tile level, one workgroup per tile walking its
# own levels) with the cost model's tiles and with fixed tile shapes (bx | by << 8 | bz << 16)
SCHEDULES = {"levels": (1, 0), "tiles": (2, 0), "tiles444": (2, 4 | 4 << 8 | 4 << 16), "tiles832": (2, 8 | 3 << 8 | 2 << 16),
             "tiles1": (2, 1 | 1 << 8 | 1 << 16)}


@pytest.mark.parametrize("sched", list(SCHEDULES))
@pytest.mark.parametrize("name", ["small", "laplace", "chain", "nb2", "nb4", "nb1", "layered", "deepwell"])
def test_bilu0_factors_and_apply(sysd, orc, name, sched):
    s = sysd[name]
    schedule, dims = SCHEDULES[sched]
    d = rs.System(s).setup(precond="bilu0", ilu_schedule=schedule, ilu_tile_dims=dims)
    info = d.info()
    assert info["ilu_tiled"] == (schedule == 2), info
    o = orc.bilu0(s, 1)
    lu, luo = d.bilu0_factors(), o.factors()
    assert np.array_equal(lu, luo), "level-scheduled factorisation reproduces the natural-order factors bit for bit"
    r = np.cos(0.7 * np.arange(s.n * s.nb))
    z, zo = d.precond(r), o.apply(r)
    assert np.max(np.abs(z - zo)) <= 1e-12 * np.max(np.abs(zo))
    if schedule == 2 and info["ilu_long_rows_L"] == 0 and info["ilu_long_rows_U"] == 0:
        # every row swept by one lane in its slot order: the oracle's arithmetic, bit for bit
        assert np.array_equal(z, zo)


@pytest.mark.parametrize("fp", ["mixed", "store32"])
@pytest.mark.parametrize("name", ["layered", "deepwell", "nb2", "nb4"])
def test_bilu0_tiled_fp32_forms(sysd, orc, name, fp):
    # FP32 factors with FP32 vectors (the mixed inner preconditioner), and FP32-stored factors with FP64 arithmetic:
    # the tiled sweeps against the level-scheduled ones in the same precision
    s = sysd[name]
    precision = rs.MIXED if fp == "mixed" else rs.MIXED_STORE
    r = np.cos(0.7 * np.arange(s.n * s.nb))
    zl = rs.System(s).setup(precond="bilu0", precision=precision, ilu_schedule=1).precond(r)
    dt = rs.System(s).setup(precond="bilu0", precision=precision, ilu_schedule=2)
    assert dt.info()["ilu_tiled"] == 1
    zt = dt.precond(r)
    zo = orc.bilu0(s, 1).apply(r)
    tol = 1e-4 if fp == "mixed" else 1e-12
    assert np.max(np.abs(zt - zl)) <= tol * np.max(np.abs(zl))
    assert np.max(np.abs(zt - zo)) <= (1e-4 if fp == "mixed" else 1e-3) * np.max(np.abs(zo))


@pytest.mark.parametrize("name", ["small", "laplace", "layered", "nb2", "deepwell", "deepwell_nb2", "deepwell_nb4"])
def test_mcbilu0_apply(sysd, orc, name):
    s = sysd[name]
    d = rs.System(s).setup(precond="mcbilu0")
    o = orc.mcbilu0(s)
    if name.startswith("deepwell"):
        info = d.info()
        assert min(info["ilu_dense_groups"]) > 0, info     # the well-clique colours go through the dense groups
    r = np.sin(0.37 * np.arange(s.n * s.nb)) + 0.5
    z, zo = d.precond(r), o.apply(r)
    assert np.max(np.abs(z - zo)) <= 1e-12 * np.max(np.abs(zo))


@pytest.mark.parametrize("name", ["deepwell", "deepwell_nb2", "deepwell_nb4"])
def test_mcbilu0_dense_groups_fp32(sysd, orc, name):
    # FP32 factors and vectors (the mixed-precision inner preconditioner) through the dense groups: FP32 rounding only
    s = sysd[name]
    d = rs.System(s).setup(precond="mcbilu0", precision=rs.MIXED)
    r = np.sin(0.37 * np.arange(s.n * s.nb)) + 0.5
    z, zo = d.precond(r), orc.mcbilu0(s).apply(r)
    assert np.max(np.abs(z - zo)) <= 1e-4 * np.max(np.abs(zo))


@pytest.mark.parametrize("name", ["small", "laplace", "layered", "nb2", "nb4", "chain"])
def test_cpr_pieces(sysd, orc, name):
    s = sysd[name]
    d = rs.System(s).setup(precond="cpr", smoother=rs.SMOOTH_GS)
    o = orc.cpr(s)
    assert np.array_equal(d.cpr_weights(), o.weights())
    levels = d.info()["amg_levels"]
    assert levels == [n for n, _ in o.levels()]
    rp = np.cos(0.11 * np.arange(s.n))
    xa, xo = d.amg_apply(rp), o.amg_apply(rp)
    assert np.max(np.abs(xa - xo)) <= 1e-10 * np.max(np.abs(xo))
    r = np.cos(0.7 * np.arange(s.n * s.nb))
    z, zo = d.precond(r), o.apply(r)
    assert np.max(np.abs(z - zo)) <= 1e-10 * np.max(np.abs(zo))


@pytest.mark.parametrize("prec", ["bilu0", "mcbilu0", "cpr"])
@pytest.mark.parametrize("name", ["small", "layered", "nb2", "chain"])
def test_bicgstab_matches_oracle(sysd, orc, name, prec):
    s = sysd[name]
    d = rs.System(s).setup(precond=prec, krylov=rs.KRYLOV_BICGSTAB)
    x, its, rel, ok = d.solve(s.rhs, TOL, 2000, raise_on_fail=False)
    xo, its_o, rel_o, ok_o = orc.krylov(s, s.rhs, prec, "bicgstab", TOL, 2000)
    if not ok_o:
        # the reference does not converge here either (MC-BILU0 on the coupled chain): the device must report the
        # same failure so the caller falls back, rather than return a wrong answer
        assert not ok and its == 2000
        return
    assert ok
    # identical arithmetic per row; the dot products and the well (long) rows are summed in a different order,
    # which a slowly converging BiCGSTAB amplifies into a few percent of iterations
    assert abs(its - its_o) <= max(1, its_o // 5)
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
    assert _true_rel(s, x, s.rhs) <= 10 * TOL


@pytest.mark.parametrize("restart", [30, 60])
@pytest.mark.parametrize("prec", ["bilu0", "cpr"])
@pytest.mark.parametrize("name", ["small", "layered"])
def test_fgmres_matches_oracle_solution(sysd, orc, name, prec, restart):
    # 60: the engine's fallback GMRES(60) (simulator.cpp:3395, 3402), orthogonalised in chunks of 32 vectors
    s = sysd[name]
    d = rs.System(s).setup(precond=prec, krylov=rs.KRYLOV_FGMRES, restart=restart)
    x, its, rel, ok = d.solve(s.rhs, TOL, 2000)
    xo, its_o, _, ok_o = orc.krylov(s, s.rhs, prec, "gmres", TOL, 2000, restart=restart)
    assert ok and ok_o
    assert abs(its - its_o) <= max(2, its_o // 10)
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
    assert _true_rel(s, x, s.rhs) <= 10 * TOL


def test_fgmres_restart_limit(sysd):
    # restarts up to 256 are built; a longer one is refused, never silently shortened
    s = sysd["small"]
    with pytest.raises(rs.RshError) as e:
        rs.System(s).setup(precond="bilu0", krylov=rs.KRYLOV_FGMRES, restart=257)
    assert e.value.code == rs.ERR_UNSUPPORTED


MIXED_MODES = {"fp32prec_bicgstab": (rs.MIXED, rs.KRYLOV_BICGSTAB), "fp32prec_fgmres": (rs.MIXED, rs.KRYLOV_FGMRES),
               "refine_bicgstab": (rs.MIXED_IR, rs.KRYLOV_BICGSTAB), "fp32store_bicgstab": (rs.MIXED_STORE, rs.KRYLOV_BICGSTAB)}


@pytest.mark.parametrize("mode", list(MIXED_MODES))
@pytest.mark.parametrize("prec", ["bilu0", "cpr", "cpr_mc"])
@pytest.mark.parametrize("name", ["small", "layered", "nb4", "deepwell"])
def test_mixed_precision_error_vs_fp64(sysd, orc, name, prec, mode):
    s = sysd[name]
    precision, krylov = MIXED_MODES[mode]
    d = rs.System(s).setup(precond=prec, precision=precision, krylov=krylov)
    x, its, rel, ok = d.solve(s.rhs, TOL, 3000)
    xo, _, _, ok_o = orc.krylov(s, s.rhs, "cpr", "bicgstab", 1e-12, 3000)
    assert ok and ok_o
    assert _true_rel(s, x, s.rhs) <= 1.01 * TOL
    err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    print(f"{mode} {name}/{prec}: {its} its, {d.info()['last_outer']} refinements, error vs FP64 {err:.2e}")
    assert err <= 1e-6


def test_refinement_never_returns_a_diverged_iterate(sysd):
    # the FP32-operator inner solves of RSH_MIXED_IR may diverge on hard systems; a correction that does not reduce
    # the FP64 residual is taken back, so the result is either converged or a finite iterate no worse than x = 0
    s = sysd["deepwell"]
    d = rs.System(s).setup(precond="mcbilu0", precision=rs.MIXED_IR)
    x, its, rel, ok = d.solve(s.rhs, TOL, 2000, raise_on_fail=False)
    true_rel = _true_rel(s, x, s.rhs)
    assert np.all(np.isfinite(x))
    if ok:
        assert true_rel <= 1.01 * TOL
    else:
        assert true_rel <= 1.0 + 1e-12 and abs(true_rel - rel) <= 1e-6 * max(rel, 1e-30) + 1e-12


def test_jacobi_smoother_converges(sysd, orc):
    s = sysd["layered"]
    d = rs.System(s).setup(precond="cpr", smoother=rs.SMOOTH_JACOBI)
    x, its, rel, ok = d.solve(s.rhs, TOL, 2000)
    xo, its_o, _, _ = orc.krylov(s, s.rhs, "cpr", "bicgstab", 1e-12, 2000)
    assert ok
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)


@pytest.mark.parametrize("name", ["small", "layered", "nb2", "nb4", "laplace"])
def test_device_cpr_block_jacobi(sysd, orc, name):
    # RSH_PREC_CPR_BJ with the Jacobi V-cycle is DeviceCprAmg (bsr.cpp:1235-1371), which the reference runs in float32:
    # the FP64 apply agrees with the float32 oracle (itself bit-identical to the compiled reference) to FP32 rounding
    s = sysd[name]
    d = rs.System(s).setup(precond="cpr_bj", smoother=rs.SMOOTH_JACOBI)
    o = orc.devcpr(s)
    r = np.cos(0.7 * np.arange(s.n * s.nb))
    z, zo = d.precond(r), o.apply(r)
    assert np.max(np.abs(z - zo)) <= 2e-5 * np.max(np.abs(zo))
    x, its, rel, ok = d.solve(s.rhs, TOL, 3000)
    xo, _, _, ok_o = orc.krylov(s, s.rhs, "cpr", "bicgstab", 1e-12, 3000)
    assert ok and ok_o and _true_rel(s, x, s.rhs) <= 1.01 * TOL
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)


@pytest.mark.parametrize("name", ["layered", "nb4", "deepwell"])
def test_fp32_hierarchy(sysd, orc, name):
    # RSH_MIXED with the Jacobi V-cycle: level values, dinv and the pressure column stored in FP32 (sums in FP64);
    # one V-cycle differs from the FP64 hierarchy's by FP32 rounding only, and the solve meets the FP64 bars
    s = sysd[name]
    d64 = rs.System(s).setup(precond="cpr_mc", smoother=rs.SMOOTH_JACOBI, smooth_steps=2)
    d32 = rs.System(s).setup(precond="cpr_mc", smoother=rs.SMOOTH_JACOBI, smooth_steps=2, precision=rs.MIXED)
    rp = np.cos(0.11 * np.arange(s.n))
    a64, a32 = d64.amg_apply(rp), d32.amg_apply(rp)
    assert np.max(np.abs(a32 - a64)) <= 1e-5 * np.max(np.abs(a64))
    if len(d64.info()["amg_levels"]) > 1:
        assert not np.array_equal(a32, a64)  # the FP32 copies are the ones used (one level = the FP64 coarsest solve)
    x, its, rel, ok = d32.solve(s.rhs, TOL, 3000)
    xo, _, _, ok_o = orc.krylov(s, s.rhs, "cpr", "bicgstab", 1e-12, 3000)
    assert ok and ok_o and _true_rel(s, x, s.rhs) <= 1.01 * TOL
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)


@pytest.mark.parametrize("steps", [2, 3])
def test_jacobi_smoother_steps(sysd, orc, steps):
    # further damped-Jacobi steps per level (modes 2 and 3 of the fused kernels): same solution, fewer iterations
    s = sysd["layered"]
    d1 = rs.System(s).setup(precond="cpr_mc", smoother=rs.SMOOTH_JACOBI)
    _, its1, _, ok1 = d1.solve(s.rhs, TOL, 2000)
    d = rs.System(s).setup(precond="cpr_mc", smoother=rs.SMOOTH_JACOBI, smooth_steps=steps)
    x, its, rel, ok = d.solve(s.rhs, TOL, 2000)
    xo, _, _, _ = orc.krylov(s, s.rhs, "cpr", "bicgstab", 1e-12, 2000)
    assert ok and ok1 and its <= its1
    assert _true_rel(s, x, s.rhs) <= 1.01 * TOL
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)


def test_refresh_equals_fresh_setup(sysd, orc):
    s = sysd["layered"]
    d = rs.System(s).setup(precond="cpr", reuse_amg=1)
    v2 = s.val * 1.5
    d.set_values(v2)
    d.setup(precond="cpr", reuse_amg=1)
    fresh = rs.System(s)
    fresh.set_values(v2)
    fresh.setup(precond="cpr", reuse_amg=0)
    r = np.cos(0.3 * np.arange(s.n * s.nb))
    z1, z2 = d.precond(r), fresh.precond(r)
    assert np.max(np.abs(z1 - z2)) <= 1e-10 * np.max(np.abs(z2))


def test_zero_rhs_and_errors(sysd):
    s = sysd["small"]
    d = rs.System(s).setup(precond="bilu0")
    x, its, rel, ok = d.solve(np.zeros(s.n * s.nb), TOL, 10)
    assert ok and its == 0 and np.all(x == 0)
    with pytest.raises(rs.RshError):
        rs.System(rs.BSR(2, 3, np.array([0, 1, 2], np.int32), np.array([1, 0], np.int32), np.zeros(18)))


@pytest.mark.parametrize("precond,precision", [("bilu0", rs.FP64), ("cpr_mc", rs.FP64), ("cpr_mc", rs.MIXED), ("cpr_mc", rs.MIXED_IR)])
def test_failures_are_reported_not_hidden(sysd, precond, precision):
    # the boundary's error behaviour (GpuBlockSolver::solve returns false, and the engine re-solves on its CPU path,
    # simulator.cpp:3366, 3376-3377): a non-finite right-hand side gives RSH_ERR_NONFINITE, too few iterations
    # RSH_ERR_NOT_CONVERGED; the Python hook mirror returns ok = False for both; the system keeps working afterwards
    s = sysd["small"]
    d = rs.System(s).setup(precond=precond, precision=precision, smoother=rs.SMOOTH_JACOBI, smooth_steps=2)
    bad = s.rhs.copy()
    bad[5] = np.nan
    with pytest.raises(rs.RshError) as e:
        d.solve(bad, TOL, 200)
    assert e.value.code == rs.ERR_NONFINITE and "non-finite right-hand side" in str(e.value)
    x, its, rel, ok = d.solve(s.rhs, 1e-14, 2, raise_on_fail=False)
    assert not ok and its <= 2
    x, its, rel, ok = d.solve(s.rhs, TOL, 500)
    assert ok and _true_rel(s, x, s.rhs) <= 1.01 * TOL
    if precision == rs.FP64:
        # a non-finite matrix value: reported as such, and the solve stops at its next residual check
        v = s.val.copy()
        v[7, 0, 0] = np.inf
        d2 = rs.System(gen.BlockSystem(n=s.n, nb=s.nb, row_ptr=s.row_ptr, col=s.col, val=v, rhs=s.rhs))
        d2.setup(precond=precond, smoother=rs.SMOOTH_JACOBI, smooth_steps=2)
        with pytest.raises(rs.RshError) as e:
            d2.solve(s.rhs, TOL, 200)
        assert e.value.code in (rs.ERR_NONFINITE, rs.ERR_BREAKDOWN)
    hook = rs.HipBlockSolver(precond)
    ok, x, its = hook.solve(s, bad, TOL, 200)
    assert not ok
    ok, x, its = hook.solve(s, s.rhs, 1e-14, 2)
    assert not ok
    ok, x, its = hook.solve(s, s.rhs, TOL, 500)
    assert ok


def test_hook_adapter(sysd, orc):
    s = sysd["small"]
    hook = rs.HipBlockSolver("cpr")
    ok, x, its = hook.solve(s, s.rhs, TOL, 500)
    xo, its_o, _, _ = orc.krylov(s, s.rhs, "cpr", "bicgstab", TOL, 500)
    assert ok and abs(its - its_o) <= 1
    ok2, x2, _ = hook.solve(s, s.rhs, TOL, 500)     # cached system: values re-uploaded, hierarchy refreshed
    assert ok2 and np.array_equal(x, x2)
    assert hook.device_solves == 2 and hook.failures == 0


def test_medium_system_jacobi_cpr(sysd, orc):
    # a 60,000-cell system (five AMG levels): DeviceCprAmg's apply agrees with the float32 oracle to FP32 rounding,
    # and the Jacobi-smoothed CPR solves with the FP64 and the FP32-stored preconditioner agree with the FP64 oracle
    # solution
    s = sysd["medium"]
    d = rs.System(s).setup(precond="cpr_bj", smoother=rs.SMOOTH_JACOBI)
    info = d.info()
    assert info["amg_sliced"][0] == 1 and len(info["amg_levels"]) >= 3
    r = np.cos(0.7 * np.arange(s.n * s.nb))
    z, zo = d.precond(r), orc.devcpr(s).apply(r)
    assert np.max(np.abs(z - zo)) <= 2e-5 * np.max(np.abs(zo))
    xo, _, _, ok_o = orc.krylov(s, s.rhs, "cpr", "bicgstab", 1e-12, 3000)
    assert ok_o
    for precision in (rs.FP64, rs.MIXED_STORE):
        d = rs.System(s).setup(precond="cpr_mc", smoother=rs.SMOOTH_JACOBI, smooth_steps=2, precision=precision)
        x, its, rel, ok = d.solve(s.rhs, TOL, 3000)
        assert ok and _true_rel(s, x, s.rhs) <= 1.01 * TOL
        assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
        print(f"medium cpr_mc precision {precision}: {its} its")


def test_mixed_refinement_recovers_from_a_diverged_inner_solve(orc):
    # a served request that made the FP32-preconditioned inner BiCGSTAB return a non-finite correction: the
    # refinement keeps the previous iterate, falls back to the FP64 preconditioner and converges
    import dataclasses
    m = gen.serving_models(10)[9]
    pv = rs.perturbed_values(m.val, m.row_ptr, 0.05, 9002)
    s = dataclasses.replace(m, val=pv.reshape(m.val.shape), full=None, handle=None)
    d = rs.System(s).setup(precond="cpr_mc", smoother=rs.SMOOTH_JACOBI, smooth_steps=2, precision=rs.MIXED)
    x, its, rel, ok = d.solve(s.rhs, TOL, 3000)
    info = d.info()
    print(f"{its} its, {info['last_outer']} refinements, FP64 fallback {info['last_fallback_fp64']}")
    assert ok and np.all(np.isfinite(x))
    assert np.linalg.norm(s.rhs - orc.spmv(s, x)) <= 1.01 * TOL * np.linalg.norm(s.rhs)


@pytest.mark.parametrize("prec", ["bilu0", "mcbilu0", "cpr_mc", "cpr_bj"])
@pytest.mark.parametrize("name", ["layered", "deepwell", "nb4"])
def test_fp32_storage_preconditioner_is_linear(sysd, name, prec):
    # RSH_MIXED_STORE keeps every preconditioner array in FP32 but computes in FP64 with FP64 vectors: the apply is an
    # exactly linear operator (to FP64 rounding), unlike FP32 arithmetic, and stays close to the FP64 apply
    s = sysd[name]
    kw = dict(precond=prec, smoother=rs.SMOOTH_JACOBI, smooth_steps=2)
    d = rs.System(s).setup(precision=rs.MIXED_STORE, **kw)
    d64 = rs.System(s).setup(precision=rs.FP64, **kw)
    rng = np.random.default_rng(3)
    r1, r2 = rng.standard_normal(s.n * s.nb), rng.standard_normal(s.n * s.nb)
    z1, z2, z12 = d.precond(r1), d.precond(r2), d.precond(2.0 * r1 - 3.0 * r2)
    lin = np.linalg.norm(z12 - (2.0 * z1 - 3.0 * z2)) / np.linalg.norm(z12)
    near = np.linalg.norm(z1 - d64.precond(r1)) / np.linalg.norm(z1)
    print(f"{name}/{prec}: linearity {lin:.1e}, FP32 storage vs FP64 {near:.1e}")
    assert lin <= 1e-10
    assert near <= 1e-3


@pytest.mark.parametrize("name", ["layered", "deepwell", "deepwell_nb2", "deepwell_nb4"])
def test_mcbilu0_factors_bit_identical(sysd, orc, name):
    # the multicolour factors, with the well-clique runs eliminated group by group (right-looking), equal the
    # natural-order BILU0 of P A P^T bit for bit (MC-BILU0 = BILU0(PAP^T), c18_tests.inc:205-215)
    s = sysd[name]
    d = rs.System(s).setup(precond="mcbilu0")
    if name.startswith("deepwell"):
        assert d.info()["ilu_factor_groups"] > 0
    k, colour, perm = orc.mcbilu0(s).colouring()
    iperm = np.empty_like(perm)
    iperm[perm] = np.arange(s.n)
    rp, cols, vals = [0], [], []
    for inew in range(s.n):
        i = perm[inew]
        p0, p1 = int(s.row_ptr[i]), int(s.row_ptr[i + 1])
        c = iperm[s.col[p0:p1]]
        o = np.argsort(c, kind="stable")
        cols.append(c[o])
        vals.append(s.val[p0:p1][o])
        rp.append(rp[-1] + p1 - p0)
    ps = gen.BlockSystem(n=s.n, nb=s.nb, row_ptr=np.array(rp, np.int32), col=np.concatenate(cols).astype(np.int32),
                         val=np.concatenate(vals), rhs=np.zeros(s.n * s.nb))
    assert np.array_equal(d.bilu0_factors(), orc.bilu0(ps, 1).factors())


def _colour_permuted(orc, s):
    """P A P^T in the oracle's multicolour order, and the order itself (new row -> old row)"""
    _, _, perm = orc.mcbilu0(s).colouring()
    iperm = np.empty_like(perm)
    iperm[perm] = np.arange(s.n)
    rp, cols, vals = [0], [], []
    for inew in range(s.n):
        i = perm[inew]
        p0, p1 = int(s.row_ptr[i]), int(s.row_ptr[i + 1])
        c = iperm[s.col[p0:p1]]
        o = np.argsort(c, kind="stable")
        cols.append(c[o])
        vals.append(s.val[p0:p1][o])
        rp.append(rp[-1] + p1 - p0)
    ps = gen.BlockSystem(n=s.n, nb=s.nb, row_ptr=np.array(rp, np.int32), col=np.concatenate(cols).astype(np.int32),
                         val=np.concatenate(vals), rhs=np.zeros(s.n * s.nb))
    return ps, perm


@pytest.mark.parametrize("name", ["small", "layered", "nb2", "nb4", "deepwell"])
def test_lean_bilu0_factors_and_apply(sysd, orc, name):
    # rsh_options.ilu_lean, the reference's C45 lean factorisation (BILU0::setLean, bilu0RowF bsr.cpp:216-263): the
    # device factors equal the oracle's lean factors bit for bit (every finished row rounded to float32, later rows
    # eliminated against the rounded values; the oracle's lean equals the compiled reference's), and the apply
    # (float32 factors, FP64 vectors and arithmetic) equals the oracle's lean apply to FP64 rounding
    s = sysd[name]
    d = rs.System(s).setup(precond="bilu0", ilu_lean=1)
    o = orc.bilu0(s, 4)
    lu = d.bilu0_factors()
    assert np.array_equal(lu, o.factors())
    assert np.array_equal(lu, lu.astype(np.float32).astype(np.float64))
    assert not np.array_equal(lu, orc.bilu0(s, 1).factors())
    r = np.cos(0.7 * np.arange(s.n * s.nb))
    z, zo = d.precond(r), o.apply(r)
    assert np.max(np.abs(z - zo)) <= 1e-12 * np.max(np.abs(zo))


@pytest.mark.parametrize("name", ["layered", "deepwell", "deepwell_nb4"])
def test_lean_mcbilu0_factors_and_apply(sysd, orc, name):
    # lean multicolour factors (the well-clique groups eliminated right-looking, each member's row rounded when it is
    # finished) equal the lean BILU0 of P A P^T bit for bit; the apply, with the dense group inverses formed in FP64
    # from the rounded factors, equals the lean sweeps of P A P^T to FP64 rounding
    s = sysd[name]
    d = rs.System(s).setup(precond="mcbilu0", ilu_lean=1)
    ps, perm = _colour_permuted(orc, s)
    o = orc.bilu0(ps, 4)
    assert np.array_equal(d.bilu0_factors(), o.factors())
    r = np.sin(0.37 * np.arange(s.n * s.nb)) + 0.5
    nb = s.nb
    rp_ = r.reshape(s.n, nb)[perm].reshape(-1)
    zp = o.apply(rp_).reshape(s.n, nb)
    zo = np.empty_like(zp)
    zo[perm] = zp
    zo = zo.reshape(-1)
    z = d.precond(r)
    assert np.max(np.abs(z - zo)) <= 1e-12 * np.max(np.abs(zo))


@pytest.mark.parametrize("name", ["layered", "deepwell", "medium"])
@pytest.mark.parametrize("precond,precision", [("bilu0", rs.FP64), ("cpr_mc", rs.FP64), ("cpr_mc", rs.MIXED_STORE)])
def test_lean_solves(sysd, orc, name, precond, precision):
    # lean factors inside the FP64 BiCGSTAB (alone and as CPR's second stage, also with the FP32-stored CPR):
    # true residual <= 1e-8 and the solution within 1e-6 of the FP64 oracle's
    s = sysd[name]
    xo, _, _, ok_o = orc.krylov(s, s.rhs, "cpr", "bicgstab", 1e-12, 3000)
    assert ok_o
    d = rs.System(s).setup(precond=precond, precision=precision, ilu_lean=1, smoother=rs.SMOOTH_JACOBI, smooth_steps=2, jacobi_omega=1.0)
    x, its, rel, ok = d.solve(s.rhs, TOL, 3000)
    assert ok and _true_rel(s, x, s.rhs) <= 1.01 * TOL
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
    if precond == "bilu0":
        _, its_o, _, _ = orc.krylov(s, s.rhs, "bilu0_lean", "bicgstab", TOL, 3000)
        assert abs(its - its_o) <= max(3, its_o // 5), (its, its_o)


def _selftest_matrix(nx, ny, nz, nb, seed):
    """The structure of the reference app's device self-test matrix (testMatrix, app/src/gpu_compute.cpp:720-749):
    7-point couplings sqrt(k_c k_o) with log-uniform permeabilities (kz 0.1 x), diagonal 1.05 x the coupling sum
    scaled 1 + 0.2 e on its diagonal and 0.03 off it, off-diagonal blocks -t on the diagonal (0.7 for e > 0).  The
    reference draws with std::uniform_real_distribution, whose values are implementation-defined, so this draws the
    same distributions from numpy: the check is the property (convergence, error < 1e-8), not the matrix."""
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    kx = 10.0 ** rng.uniform(-1, 1, n)
    kz = 0.1 * 10.0 ** rng.uniform(-1, 1, n)
    rp, cols, vals = [0], [], []
    eye = np.eye(nb)
    offd = -np.diag([1.0] + [0.7] * (nb - 1))
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                c = (k * ny + j) * nx + i
                row = []
                if k > 0: row.append((c - nx * ny, np.sqrt(kz[c] * kz[c - nx * ny])))
                if j > 0: row.append((c - nx, np.sqrt(kx[c] * kx[c - nx])))
                if i > 0: row.append((c - 1, np.sqrt(kx[c] * kx[c - 1])))
                if i < nx - 1: row.append((c + 1, np.sqrt(kx[c] * kx[c + 1])))
                if j < ny - 1: row.append((c + nx, np.sqrt(kx[c] * kx[c + nx])))
                if k < nz - 1: row.append((c + nx * ny, np.sqrt(kz[c] * kz[c + nx * ny])))
                row.append((c, 1.05 * sum(t for _, t in row)))
                row.sort()
                for o, t in row:
                    cols.append(o)
                    if o == c:
                        vals.append(np.where(eye > 0, t * (1 + 0.2 * np.arange(nb))[:, None] * eye, 0.03 * t))
                    else:
                        vals.append(t * offd)
                rp.append(len(cols))
    return gen.BlockSystem(n=n, nb=nb, row_ptr=np.array(rp, np.int32), col=np.array(cols, np.int32),
                           val=np.array(vals, np.float64), rhs=np.zeros(n * nb))


@pytest.mark.parametrize("precond", ["mcbilu0", "cpr_bj", "cpr_mc"])
def test_device_selftest_known_answer(orc, precond):
    # the reference app's device self-test (selfTest, app/src/gpu_compute.cpp:759-797): a 120 x 100 x 3 system with a
    # known solution x_t = cos(0.003 e) + 0.2 (e mod 3), solved through the FP64 refinement around the FP32 device
    # solve (refineWithDevice to 1e-10, 8 outer steps, 2000 inner iterations): converged, max-norm error < 1e-8
    s = _selftest_matrix(120, 100, 1, 3, 5)
    e = np.arange(s.n * s.nb)
    xt = np.cos(0.003 * e) + 0.2 * (e % 3)
    b = orc.spmv(s, xt)
    d = rs.System(s).setup(precond=precond, precision=rs.MIXED_IR, max_outer=8, smoother=rs.SMOOTH_JACOBI)
    x, its, rel, ok = d.solve(b, 1e-10, 2000)
    assert ok
    assert np.max(np.abs(x - xt)) / np.max(np.abs(xt)) < 1e-8


def _c38_matrix():
    """The reference's own C38 test system (testC38FloatIlu, c38_tests.inc:159-186), built with its portable 64-bit
    LCG (so the same values as the reference draws): a 30 x 30 x 8 grid, nb 3, weaker couplings across layers."""
    nx, ny, nz, nb = 30, 30, 8, 3
    state = [12345]

    def rnd():
        state[0] = (state[0] * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        return float((state[0] >> 11) & 0xFFFFF) / float(0xFFFFF)

    def cid(i, j, k):
        return (k * ny + j) * nx + i

    n = nx * ny * nz
    rp, cols, vals = [0], [], []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                row = []
                if k > 0: row.append(cid(i, j, k - 1))
                if j > 0: row.append(cid(i, j - 1, k))
                if i > 0: row.append(cid(i - 1, j, k))
                row.append(cid(i, j, k))
                if i < nx - 1: row.append(cid(i + 1, j, k))
                if j < ny - 1: row.append(cid(i, j + 1, k))
                if k < nz - 1: row.append(cid(i, j, k + 1))
                r = cid(i, j, k)
                for c in row:
                    blk = []
                    for e in range(9):
                        if c == r:
                            blk.append(0.3 * (rnd() - 0.5))
                        else:
                            v = 1.0 + rnd() if e % 4 == 0 else 0.2 * (rnd() - 0.5)
                            blk.append(-v * (0.1 if c // (nx * ny) != r // (nx * ny) else 1.0))
                    if c == r:
                        for e in range(3):
                            blk[e * 4] += 7.5 + rnd()
                    cols.append(c)
                    vals.append(np.array(blk).reshape(3, 3))
                rp.append(len(cols))
    rhs = np.array([rnd() - 0.5 for _ in range(n * nb)])
    return gen.BlockSystem(n=n, nb=nb, row_ptr=np.array(rp, np.int32), col=np.array(cols, np.int32), val=np.array(vals),
                           rhs=rhs)


def test_c38_float_factors(orc):
    # the reference's C38 check (c38_tests.inc:186-211) on its own matrix: the BILU0 apply from float32-stored factors
    # (FP64 vectors and arithmetic) equals the FP64 one to float rounding (|dz|/|z| < 1e-5), also after a
    # re-factorisation of new values on the same pattern (values x 2); beyond it, both applies equal the oracle's
    # (FP64, and C38's float factors) to FP64 rounding
    s = _c38_matrix()
    d64 = rs.System(s).setup(precond="bilu0")
    d32 = rs.System(s).setup(precond="bilu0", precision=rs.MIXED_STORE)
    for scale in (1.0, 2.0):
        if scale != 1.0:
            v = s.val * scale
            d64.set_values(v)
            d64.setup(precond="bilu0")
            d32.set_values(v)
            d32.setup(precond="bilu0", precision=rs.MIXED_STORE)
            s = gen.BlockSystem(n=s.n, nb=s.nb, row_ptr=s.row_ptr, col=s.col, val=v, rhs=s.rhs)
        z1, z2 = d64.precond(s.rhs), d32.precond(s.rhs)
        assert np.linalg.norm(z1 - z2) / np.linalg.norm(z1) < 1e-5
        zo1, zo2 = orc.bilu0(s, 1).apply(s.rhs), orc.bilu0(s, 2).apply(s.rhs)
        assert np.max(np.abs(z1 - zo1)) <= 1e-12 * np.max(np.abs(zo1))
        assert np.max(np.abs(z2 - zo2)) <= 1e-12 * np.max(np.abs(zo2))


@pytest.mark.parametrize("name", ["layered", "medium"])
def test_coarse_smoothing_steps_option(sysd, orc, name):
    # rsh_options.coarse_smooth_steps: Jacobi steps on the levels below the finest (0: smooth_steps); another count
    # changes the apply and still converges to the FP64 oracle's solution
    s = sysd[name]
    xo, _, _, ok_o = orc.krylov(s, s.rhs, "cpr", "bicgstab", 1e-12, 3000)
    assert ok_o
    d = rs.System(s)
    r = np.cos(0.3 * np.arange(s.n * s.nb))
    zs = []
    for cs in (0, 2, 1):
        d.setup(precond="cpr_mc", smoother=rs.SMOOTH_JACOBI, smooth_steps=2, jacobi_omega=1.0, coarse_smooth_steps=cs)
        zs.append(d.precond(r))
        x, its, rel, ok = d.solve(s.rhs, TOL, 3000)
        assert ok and _true_rel(s, x, s.rhs) <= 1.01 * TOL
        assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
    assert np.array_equal(zs[0], zs[1])          # 0 means smooth_steps
    assert not np.array_equal(zs[0], zs[2])


@pytest.mark.parametrize("name", ["layered", "deepwell", "medium"])
def test_jacobi_weight_option(sysd, orc, name):
    # rsh_options.jacobi_omega: the V-cycle's damping weight (default DeviceCprAmg's 0.6667); weight 1 converges to the
    # FP64 oracle solution like the default, and a changed weight changes the apply (the captured apply is rebuilt)
    s = sysd[name]
    xo, _, _, ok_o = orc.krylov(s, s.rhs, "cpr", "bicgstab", 1e-12, 3000)
    assert ok_o
    d = rs.System(s)
    r = np.cos(0.3 * np.arange(s.n * s.nb))
    zs = []
    for w in (0.0, 1.0):
        d.setup(precond="cpr_mc", smoother=rs.SMOOTH_JACOBI, smooth_steps=2, jacobi_omega=w)
        zs.append(d.precond(r))
        x, its, rel, ok = d.solve(s.rhs, TOL, 3000)
        assert ok and _true_rel(s, x, s.rhs) <= 1.01 * TOL
        assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
        print(f"{name} weight {w or 0.6667}: {its} its")
    assert not np.array_equal(zs[0], zs[1])


def test_wide_layers_use_32bit_columns(orc):
    # 200 x 200 x 2 cells: the vertical neighbour is 40,000 block rows away, beyond the slice layout's 16-bit relative
    # columns, so the SpMV, the sliced pressure smoother and the CPR second stage read 32-bit columns: the SpMV against
    # the oracle's, the MC-BILU0 apply to 1e-12, and CPR-BiCGSTAB (FP64 and FP32-stored) against the FP64 oracle
    # solution (<= 1e-6) with the true relative residual <= 1e-8
    s = gen.reservoir(200, 200, 2, 3, model=0, sigma=1.0, wells=2, seed=77)
    d = rs.System(s)
    info = d.info()
    assert info["spmv_col16"] == 0, info
    x = np.random.default_rng(3).standard_normal(s.n * s.nb)
    y, yo = d.spmv(x), orc.spmv(s, x)
    if info["spmv_long_rows"] == 0:
        assert np.array_equal(y, yo)
    assert np.max(np.abs(y - yo)) <= 1e-13 * np.max(np.abs(yo))
    d.setup(precond="mcbilu0")
    r = np.sin(0.37 * np.arange(s.n * s.nb)) + 0.5
    z, zo = d.precond(r), orc.mcbilu0(s).apply(r)
    assert np.max(np.abs(z - zo)) <= 1e-12 * np.max(np.abs(zo))
    xo, _, _, ok_o = orc.krylov(s, s.rhs, "cpr", "bicgstab", 1e-10, 3000)
    assert ok_o
    for precision in (rs.FP64, rs.MIXED_STORE):
        dc = rs.System(s).setup(precond="cpr_mc", precision=precision, smoother=rs.SMOOTH_JACOBI, smooth_steps=2,
                                jacobi_omega=1.0)
        assert dc.info()["spmv_col16"] == 0
        x, its, rel, ok = dc.solve(s.rhs, TOL, 1000)
        assert ok and _true_rel(s, x, s.rhs) <= 1.01 * TOL
        assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)


FLOAT_FORMS = {"bilu0f": dict(precond="bilu0"), "mcbilu0": dict(precond="mcbilu0"),
               "devcpr": dict(precond="cpr_bj", smoother=rs.SMOOTH_JACOBI)}


@pytest.mark.parametrize("kind", list(FLOAT_FORMS))
@pytest.mark.parametrize("name", ["small", "nb2", "nb4"])
def test_mixed_ir_against_the_float_device_solve(sysd, orc, name, kind):
    # RSH_MIXED_IR against the reference's float32 device solve (CpuFloatBlockSolver::solve, bsr.cpp:1130-1215) and
    # its FP64 refinement (refineWithDevice, bsr.cpp:1373-1396), restated in the oracle and pinned bit for bit to the
    # compiled reference.  The device keeps the inner Krylov vectors and dot products in FP64 (the reference's are
    # float32), with the FP32 matrix and preconditioner of the same kind, so the inner solves agree to their
    # tolerance, not bit for bit: the first correction (one refinement step) reaches the inner tolerance 1e-4 as the
    # float solve does and is no further from the exact solution than 10x the float solve's error; the refined
    # solutions both meet 1e-8 and agree to 1e-6.
    s = sysd[name]
    xe, _, _, ok = orc.krylov(s, s.rhs, "cpr", "bicgstab", 1e-13, 3000)
    assert ok
    one = rs.System(s).setup(precision=rs.MIXED_IR, max_outer=1, **FLOAT_FORMS[kind])
    d1, _, _, _ = one.solve(s.rhs, TOL, 3000, raise_on_fail=False)
    f1, _, fin = orc.float_solve(s, s.rhs, kind, 1e-4, 3000)
    assert fin
    res = lambda x: float(np.linalg.norm(s.rhs - orc.spmv(s, x)) / np.linalg.norm(s.rhs))
    err = lambda x: float(np.linalg.norm(x - xe) / np.linalg.norm(xe))
    assert res(d1) <= 2e-4 and res(f1) <= 2e-4, (res(d1), res(f1))
    assert err(d1) <= 10 * err(f1) + 1e-9, (err(d1), err(f1))
    full = rs.System(s).setup(precision=rs.MIXED_IR, **FLOAT_FORMS[kind])
    x, its, rel, ok = full.solve(s.rhs, TOL, 3000)
    xr, its_r, rel_r, conv = orc.refine_float(s, s.rhs, kind, TOL, 10, 3000)
    assert ok and conv and res(x) <= 1.01 * TOL and rel_r <= TOL
    assert np.linalg.norm(x - xr) <= 1e-6 * np.linalg.norm(xr)
    print(f"{name}/{kind}: device outer {full.info()['last_outer']} inner its {its}; reference outer-loop inner its {its_r}")


@pytest.mark.parametrize("precision", [rs.FP64, rs.MIXED_STORE])
def test_sliced_coarse_levels_bit_identical(sysd, monkeypatch, precision):
    # coarse AMG levels above a row threshold are smoothed through row slices (one row per lane); each row's sum is
    # formed in the chains of the CSR step's G lanes and folded in its butterfly order, so V-cycles, CPR applies and
    # whole solves equal those of the G-lanes-per-row CSR step bit for bit.  The threshold is lowered so that the
    # 60,000-cell system's coarse levels (about 11,000 and 1,600 rows) take the slices, long rows included.
    s = sysd["medium"]
    r = np.cos(0.3 * np.arange(s.n * s.nb))
    out = {}
    for rows in ("1000000000", "1000"):
        monkeypatch.setenv("RSH_AMG_SLICE_ROWS", rows)
        d = rs.System(s).setup(precond="cpr_mc", smoother=rs.SMOOTH_JACOBI, smooth_steps=2, jacobi_omega=1.0, precision=precision)
        info = d.info()
        x, its, rel, ok = d.solve(s.rhs, TOL, 3000)
        assert ok
        out[rows] = (info, d.amg_apply(r[: s.n]), d.precond(r), x, its)
    csr, sl = out["1000000000"], out["1000"]
    assert all(v[0] == 0 for v in csr[0]["amg_slices"][1:])
    sliced = [l for l, v in enumerate(sl[0]["amg_slices"]) if l > 0 and v[0] > 0]
    assert len(sliced) >= 1 and any(sl[0]["amg_slices"][l][1] > 0 for l in sliced)   # a sliced level with long rows
    assert np.array_equal(csr[1], sl[1])
    assert np.array_equal(csr[2], sl[2])
    assert np.array_equal(csr[3], sl[3]) and csr[4] == sl[4]
    print(f"precision {precision}: levels {sl[0]['amg_levels']} sliced {sliced} slices {sl[0]['amg_slices']}, {sl[4]} its")
