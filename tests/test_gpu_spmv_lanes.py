"""SpMV with 1, 2 or 4 lanes per row (64, 32 or 16 rows per wavefront) through the C-ABI: the product against the
oracle's BSR::spmv (one lane per row bit for bit where no row is long; 2 and 4 lanes add each row's blocks in
another order, so within 1e-13 of the largest entry), Krylov solves with the split rows against the oracle's FP64
solve (<= 1e-6, true relative residual <= 1e-8), and the batch autotuner's choice of lanes per block-size group."""
import numpy as np
import pytest

import ressimstudio_amd as rs
from ressimstudio_amd import gen

pytestmark = pytest.mark.gpu

TOL = 1e-8

SYSTEMS = {
    "small": gen.small,
    "nb2": lambda: gen.reservoir(12, 10, 3, 2, model=0, sigma=1.5, wells=2, seed=5),
    "nb4": lambda: gen.reservoir(10, 9, 3, 4, model=0, sigma=0.8, wells=1, seed=6),
    "nb1": lambda: gen.reservoir(16, 12, 4, 1, model=0, sigma=1.0, wells=2, seed=8),
    "layered": lambda: gen.reservoir(12, 30, 14, 3, model=1, sigma=2.5, kvkh=0.1, wells=5, seed=7),
    # tall wells: long rows (perforated cells after the Schur fill) beside the slices
    "deepwell": lambda: gen.reservoir(24, 24, 60, 3, model=1, sigma=2.0, kvkh=0.1, wells=3, seed=9),
}



def batch_systems():
    # uneven sizes, all three block sizes, wells with Schur fill: (nx, ny, nz, nb, sigma, wells, seed)
    spec = [(10, 10, 3, 2, 0.5, 1, 101), (24, 18, 6, 3, 2.0, 2, 102), (12, 10, 4, 4, 1.0, 1, 103), (30, 20, 8, 3, 1.5, 3, 104),
            (16, 16, 5, 2, 2.5, 2, 105), (8, 9, 3, 3, 0.8, 0, 106), (20, 12, 10, 3, 1.2, 2, 107), (14, 14, 4, 4, 2.0, 2, 108)]
    return [gen.reservoir(nx, ny, nz, nb, model=0, sigma=sg, wells=w, seed=sd) for nx, ny, nz, nb, sg, w, sd in spec]


@pytest.mark.parametrize("lanes", [1, 2, 4])
@pytest.mark.parametrize("name", list(SYSTEMS))
def test_spmv_lanes_match_oracle(orc, name, lanes):
    s = SYSTEMS[name]()
    d = rs.System(s)
    x = np.random.default_rng(3).standard_normal(s.n * s.nb)
    y1 = d.spmv(x)
    d.set_spmv_lanes(lanes)
    assert d.info()["spmv_lanes"] == lanes
    y = d.spmv(x)
    yo = orc.spmv(s, x)
    if lanes == 1:
        assert np.array_equal(y, y1)
    if lanes == 1 and d.info()["spmv_long_rows"] == 0:
        assert np.array_equal(y, yo)
    assert np.max(np.abs(y - yo)) <= 1e-13 * np.max(np.abs(yo))
    # new values after the layout change are packed into it too
    d.set_values(s.val * 1.5)
    assert np.max(np.abs(d.spmv(x) - 1.5 * yo)) <= 1e-13 * np.max(np.abs(yo))
    d.set_values(s.val)
    assert np.array_equal(d.spmv(x), y)
    d.set_spmv_lanes(1)
    assert d.info()["spmv_lanes"] == 1 and np.array_equal(d.spmv(x), y1)


def test_spmv_lanes_refused(orc):
    d = rs.System(gen.small())
    for bad in (0, 3, 8, -1):
        with pytest.raises(rs.RshError):
            d.set_spmv_lanes(bad)
    assert d.info()["spmv_lanes"] == 1


@pytest.mark.parametrize("lanes", [2, 4])
@pytest.mark.parametrize("precond", ["cpr", "mcbilu0"])
def test_solve_with_split_rows(orc, lanes, precond):
    s = SYSTEMS["layered"]()
    xo, _, _, ok_o = orc.krylov(s, s.rhs, "bilu0", "bicgstab", 1e-10, 3000)
    assert ok_o
    d = rs.System(s).set_spmv_lanes(lanes).setup(precond=precond, krylov=rs.KRYLOV_BICGSTAB)
    x, its, rel, ok = d.solve(s.rhs, TOL, 3000)
    assert ok and rel <= TOL
    assert np.linalg.norm(s.rhs - orc.spmv(s, x)) / np.linalg.norm(s.rhs) <= 1.01 * TOL
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
    x2, its2, _, _ = d.solve(s.rhs, TOL, 3000)
    assert its2 == its and np.array_equal(x2, x)       # deterministic


@pytest.mark.parametrize("lanes", [0, 1, 2, 4])
def test_batch_spmv_lanes(orc, lanes):
    systems = batch_systems()
    B = rs.Batch(systems).set_spmv_lanes(lanes).setup(precond="mcbilu0")
    info = B.info()
    chosen = [g["spmv_lanes"] for g in info["groups"]]
    assert all(c in (1, 2, 4) for c in chosen)
    if lanes:
        assert chosen == [lanes] * len(chosen)
    xs, its, rel, st = B.solve([s.rhs for s in systems], TOL, 3000)
    for k, s in enumerate(systems):
        assert st[k] == rs.OK, k
        assert np.linalg.norm(s.rhs - orc.spmv(s, xs[k])) / np.linalg.norm(s.rhs) <= 1.01 * TOL, k
        d = rs.System(s).setup(precond="mcbilu0", krylov=rs.KRYLOV_BICGSTAB)
        x1, _, _, ok = d.solve(s.rhs, TOL, 3000)
        assert ok and np.linalg.norm(xs[k] - x1) <= 1e-6 * np.linalg.norm(x1), k
    # another system's right-hand side changes nothing for the others (slices shared by two systems included)
    bs = [s.rhs.copy() for s in systems]
    bs[0] = bs[0] * 2.0
    xs2, its2, _, _ = B.solve(bs, TOL, 3000)
    for k in range(1, len(systems)):
        assert its2[k] == its[k] and np.array_equal(xs2[k], xs[k]), k
    # a fixed choice replaces the tuned one at the next set-up
    B.set_spmv_lanes(4).setup(precond="mcbilu0")
    assert [g["spmv_lanes"] for g in B.info()["groups"]] == [4] * len(chosen)
    xs3, _, _, st3 = B.solve([s.rhs for s in systems], TOL, 3000)
    assert all(v == rs.OK for v in st3)
