"""The serving loop in batched rounds (rsh_server_create_batched) through the C-ABI: the requests queued for one group
of models are solved together by the batch engine (one numeric set-up and one batched BiCGSTAB per round).  Every
result must be the same whether its request ran alone in a round or beside others (bit for bit: the batch's
preconditioner is block-diagonal and its Krylov scalars are per system), meet the 1e-8 true residual on its perturbed
matrix (recomputed with the oracle's SpMV) and match the oracle's solve to 1e-6."""
import numpy as np
import pytest

import ressimstudio_amd as rs
from ressimstudio_amd import gen

pytestmark = pytest.mark.gpu

TOL = 1e-8


def models():
    return [gen.reservoir(20, 16, 6, 3, model=0, sigma=1.0, wells=2, seed=201), gen.reservoir(30, 24, 8, 3, model=1, sigma=2.0, kvkh=0.1, wells=3, seed=202),
            gen.reservoir(12, 12, 5, 3, model=0, sigma=0.7, wells=1, seed=203), gen.reservoir(26, 18, 10, 3, model=0, sigma=1.5, wells=2, seed=204)]


@pytest.fixture(scope="module")
def mods():
    return models()


def _run(srv, ids, mods, reqs, together):
    out = {}
    if together:
        tickets = [(i, srv.submit(ids[m], mods[m].rhs, perturb=0.05, seed=sd, tol=TOL, max_it=2000)) for i, (m, sd) in enumerate(reqs)]
        for i, t in tickets:
            out[i] = srv.wait(t)
    else:
        for i, (m, sd) in enumerate(reqs):
            out[i] = srv.wait(srv.submit(ids[m], mods[m].rhs, perturb=0.05, seed=sd, tol=TOL, max_it=2000))
    return out


@pytest.mark.parametrize("precond,precision,cells", [("cpr_mc", rs.FP64, 10 ** 7), ("cpr_mc", rs.MIXED_STORE, 10 ** 7),
                                                    ("mcbilu0", rs.FP64, 10 ** 7), ("cpr_mc", rs.FP64, 3000)])
def test_batched_rounds_same_alone_and_in_the_mix(mods, orc, precond, precision, cells):
    srv = rs.Server(slots=2, batched=True, group_cells=cells, group_systems=16)
    kw = dict(smoother=rs.SMOOTH_JACOBI, smooth_steps=2, jacobi_omega=1.0) if precond.startswith("cpr") else {}
    ids = [srv.add_model(s, precond=precond, precision=precision, **kw) for s in mods]
    reqs = [(k % len(mods), 1000 + k) for k in range(3 * len(mods))]
    _run(srv, ids, mods, [(m, 1) for m in range(len(mods))], together=False)      # loads the groups
    st0 = srv.stats()
    assert st0["batched"] == 1 and sum(st0["group_models"]) == len(mods), st0
    if cells >= 10 ** 6:
        assert st0["groups"] == 1, st0
    else:
        assert st0["groups"] > 1, st0                                                # the small cell limit splits the models
    alone = _run(srv, ids, mods, reqs, together=False)
    st1 = srv.stats()
    mix = _run(srv, ids, mods, reqs, together=True)
    st2 = srv.stats()
    assert st1["rounds"] - st0["rounds"] == len(reqs), (st0, st1)                   # one request per round
    if cells >= 10 ** 6:
        assert st2["rounds"] - st1["rounds"] < len(reqs), (st1, st2)                # rounds of several requests
    with pytest.raises(rs.RshError):
        srv.add_model(mods[0], precond=precond)                                      # grouped at the first request
    for i, (m, sd) in enumerate(reqs):
        a, b = alone[i], mix[i]
        assert a["status"] == rs.OK and b["status"] == rs.OK, (a["status"], b["status"])
        assert np.array_equal(a["x"], b["x"]) and a["iterations"] == b["iterations"], i
        s = mods[m]
        v = rs.perturbed_values(s.val, s.row_ptr, 0.05, sd)
        sp = gen.BlockSystem(n=s.n, nb=s.nb, row_ptr=s.row_ptr, col=s.col, val=v.reshape(s.val.shape), rhs=s.rhs)
        r = s.rhs - orc.spmv(sp, a["x"])
        assert np.linalg.norm(r) <= 1.01 * TOL * np.linalg.norm(s.rhs), i
        xo, _, _, ok = orc.krylov(sp, s.rhs, "cpr", "bicgstab", 1e-12, 3000)
        assert ok and np.linalg.norm(a["x"] - xo) <= 1e-6 * np.linalg.norm(xo), i
    srv.close()


def test_batched_server_argument_checks():
    with pytest.raises(rs.RshError):
        rs.Server(slots=2, batched=True, group_cells=0)
    with pytest.raises(rs.RshError):
        rs.Server(slots=2, batched=True, group_systems=0)


def test_batched_groups_beyond_the_budget_are_refused(mods):
    # a one-byte budget: the first group is admitted (nothing else resident), every other group is refused at once
    # (groups are not evicted) rather than left waiting; the admitted group keeps serving
    srv = rs.Server(slots=1, batched=True, group_cells=3000, group_systems=16, hbm_budget=1.0)
    ids = [srv.add_model(s, precond="mcbilu0") for s in mods]
    res = [srv.wait(srv.submit(ids[m], mods[m].rhs, tol=TOL, max_it=2000)) for m in range(len(mods))]
    st = srv.stats()
    assert st["groups"] == 3 and sorted(st["group_models"]) == [1, 1, 2], st     # sizes 720 + 1920 | 4680 | 5760 cells
    assert res[0]["status"] == rs.OK and res[2]["status"] == rs.OK                 # models 0 and 2: the first group
    assert res[1]["status"] == rs.ERR_UNSUPPORTED and res[3]["status"] == rs.ERR_UNSUPPORTED
    again = srv.wait(srv.submit(ids[0], mods[0].rhs, tol=TOL, max_it=2000))
    assert again["status"] == rs.OK and np.array_equal(again["x"], res[0]["x"])
    srv.close()
