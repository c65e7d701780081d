"""The serving loop (rsh_server_*) through the C-ABI: requests from several resident models (each a different
Newton iteration: block rows scaled by +-5 % on the device), solved alone one after another and then all submitted at
once so that they overlap on the device; every result must be the same alone and in the mix (the kernels are
deterministic: bit for bit), meet the 1e-8 true residual on its perturbed matrix and match the oracle's solve."""
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


def _requests(mods):
    # (model, perturbation seed): three Newton iterations of every model, interleaved as ensemble members would arrive
    return [(k % len(mods), 1000 + k) for k in range(3 * len(mods))]


def _run(srv, ids, mods, reqs, together, **kw):
    out = {}
    if together:
        tickets = [(i, srv.submit(ids[m], mods[m].rhs, perturb=0.05, seed=sd, tol=TOL, max_it=2000)) for i, (m, sd) in enumerate(reqs)]
        for i, t in tickets:
            out[i] = srv.wait(t)
    else:
        for i, (m, sd) in enumerate(reqs):
            out[i] = srv.wait(srv.submit(ids[m], mods[m].rhs, perturb=0.05, seed=sd, tol=TOL, max_it=2000))
    return out


@pytest.mark.parametrize("precision", [rs.FP64, rs.MIXED])
def test_same_result_alone_and_in_the_mix(mods, orc, precision):
    srv = rs.Server(slots=4)
    ids = [srv.add_model(s, precond="cpr_mc", smoother=rs.SMOOTH_JACOBI, precision=precision) for s in mods]
    reqs = _requests(mods)
    # first Newton iteration of every model: builds its AMG hierarchy, later iterations refresh it (AMG::refresh)
    _run(srv, ids, mods, [(m, 1) for m in range(len(mods))], together=False)
    alone = _run(srv, ids, mods, reqs, together=False)
    mix = _run(srv, ids, mods, reqs, together=True)
    st = srv.stats()
    assert st["peak_running"] >= 2, st            # requests did overlap on the device
    for i, (m, sd) in enumerate(reqs):
        a, b = alone[i], mix[i]
        assert a["status"] == rs.OK and b["status"] == rs.OK
        assert np.array_equal(a["x"], b["x"]) and a["iterations"] == b["iterations"], i
        s = mods[m]
        v = rs.perturbed_values(s.val, s.row_ptr, 0.05, sd)
        sp = gen.BlockSystem(n=s.n, nb=s.nb, row_ptr=s.row_ptr, col=s.col, val=v.reshape(s.val.shape), rhs=s.rhs)
        r = s.rhs - orc.spmv(sp, a["x"])
        assert np.linalg.norm(r) <= 1.01 * TOL * np.linalg.norm(s.rhs), i
        xo, _, _, ok = orc.krylov(sp, s.rhs, "cpr", "bicgstab", 1e-12, 3000)
        assert ok and np.linalg.norm(a["x"] - xo) <= 1e-6 * np.linalg.norm(xo), i
    srv.close()


def test_hbm_budget_evicts_idle_models(mods):
    # MC-BILU0 (no AMG hierarchy whose aggregates would come from the loading request): a reloaded model solves
    # exactly as a resident one
    ref = rs.Server(slots=2)
    ids = [ref.add_model(s, precond="mcbilu0") for s in mods[:3]]
    reqs = [(k % 3, 500 + k) for k in range(6)]
    base = _run(ref, ids, mods, reqs, together=False)
    foot = ref.stats()["footprints"]
    ref.close()
    # a budget that holds the two largest models but not all three: serving them in turn must evict and reload
    big = sorted(foot)
    srv = rs.Server(slots=2, hbm_budget=big[2] + 1.05 * big[1])
    ids = [srv.add_model(s, precond="mcbilu0") for s in mods[:3]]
    got = _run(srv, ids, mods, reqs, together=True)
    st = srv.stats()
    assert st["evictions"] >= 1 and st["hbm_used"] <= st["hbm_budget"], st
    for i in range(len(reqs)):
        assert got[i]["status"] == rs.OK and np.array_equal(got[i]["x"], base[i]["x"]), i
    srv.close()


def test_perturbation_matches_host_formula(mods):
    # the host restatement of the device perturbation used by the checks above
    s = mods[2]
    srv = rs.Server(slots=1)
    mid = srv.add_model(s, precond="bilu0")
    r = srv.wait(srv.submit(mid, s.rhs, perturb=0.05, seed=77, tol=TOL, max_it=2000))
    v = rs.perturbed_values(s.val, s.row_ptr, 0.05, 77)
    d = rs.System(rs.BSR(s.n, s.nb, s.row_ptr, s.col, v)).setup(precond="bilu0")
    x, its, rel, ok = d.solve(s.rhs, TOL, 2000)
    assert ok and np.array_equal(x, r["x"]) and its == r["iterations"]
    srv.close()


def test_concurrent_callers(mods):
    # the reference's ensemble callers (runBatch, study.cpp:345-361) are threads that each submit their own Newton
    # iterations and wait for them; four Python threads do so against one server (ctypes releases the GIL), and
    # every result must equal the same request served alone
    import threading
    srv = rs.Server(slots=4)
    ids = [srv.add_model(s, precond="cpr_mc", smoother=rs.SMOOTH_JACOBI) for s in mods]
    _run(srv, ids, mods, [(m, 1) for m in range(len(mods))], together=False)
    per_thread = [[(t, 3000 + 10 * t + k) for k in range(4)] for t in range(4)]
    alone = {(m, sd): _run(srv, ids, mods, [(m, sd)], together=False)[0] for reqs in per_thread for (m, sd) in reqs}
    got, errors = {}, []

    def caller(reqs):
        try:
            for m, sd in reqs:
                got[(m, sd)] = srv.wait(srv.submit(ids[m], mods[m].rhs, perturb=0.05, seed=sd, tol=TOL, max_it=2000))
        except Exception as e:       # surfaced below
            errors.append(e)

    th = [threading.Thread(target=caller, args=(r,)) for r in per_thread]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not errors and not any(t.is_alive() for t in th), errors
    for key, a in alone.items():
        b = got[key]
        assert a["status"] == rs.OK and b["status"] == rs.OK, key
        assert np.array_equal(a["x"], b["x"]) and a["iterations"] == b["iterations"], key
    srv.close()


def test_model_larger_than_the_budget_is_served_alone(mods):
    # FGMRES(30) keeps 61 Krylov vectors: the measured footprint exceeds the estimate the model was admitted with and
    # the budget itself.  Its requests must still be served (alone, once nothing else is resident), never left queued
    ref = rs.Server(slots=1)
    mid = ref.add_model(mods[1], precond="bilu0", krylov=rs.KRYLOV_FGMRES, restart=30)
    base = ref.wait(ref.submit(mid, mods[1].rhs, perturb=0.05, seed=5, tol=TOL, max_it=2000))
    foot = ref.stats()["footprints"][mid]
    ref.close()
    srv = rs.Server(slots=2, hbm_budget=0.8 * foot)
    i_big = srv.add_model(mods[1], precond="bilu0", krylov=rs.KRYLOV_FGMRES, restart=30)
    i_small = srv.add_model(mods[2], precond="bilu0")
    t = [srv.submit(i_big, mods[1].rhs, perturb=0.05, seed=5, tol=TOL, max_it=2000),
         srv.submit(i_small, mods[2].rhs, perturb=0.05, seed=6, tol=TOL, max_it=2000),
         srv.submit(i_big, mods[1].rhs, perturb=0.05, seed=5, tol=TOL, max_it=2000)]
    res = [srv.wait(k) for k in t]
    assert all(r["status"] == rs.OK for r in res), res
    assert np.array_equal(res[0]["x"], base["x"]) and np.array_equal(res[2]["x"], base["x"])
    srv.close()


def test_worker_error_reaches_the_waiting_thread(mods):
    # a request that fails inside a worker (here: an option refused at setup) reports its status, and
    # rsh_last_error() in the waiting thread carries the worker's message
    srv = rs.Server(slots=1)
    mid = srv.add_model(mods[2], precond="bilu0", krylov=rs.KRYLOV_FGMRES, restart=300)
    r = srv.wait(srv.submit(mid, mods[2].rhs, tol=TOL, max_it=100))
    assert r["status"] == rs.ERR_UNSUPPORTED
    assert "restart" in rs.lib().rsh_last_error().decode()
    srv.close()


def test_fp32_stall_is_paid_once(orc):
    # model 9 of the serving workload (155,526 cells): its FP32-preconditioned inner BiCGSTAB stalls (the refinement
    # then falls back to the FP64 preconditioner: about 365 iterations instead of 65).  The system remembers the
    # stall, so its next Newton iteration starts with the FP64 preconditioner; retry_fp32=1 starts with FP32 again.
    m = gen.serving_model(9)
    assert m.n == 155526
    kw = dict(precond="cpr_mc", smoother=rs.SMOOTH_JACOBI, smooth_steps=2, jacobi_omega=1.0, precision=rs.MIXED)
    d = rs.System(m).setup(**kw)
    x, its, rel, ok = d.solve(m.rhs, TOL, 3000)
    inf = d.info()
    assert ok and inf["last_fallback_fp64"] == 1 and inf["fp32_stalled"] == 1, inf["last_inner"]
    for retry in (0, 1):
        v = rs.perturbed_values(m.val, m.row_ptr, 0.05, 9002)
        sp = gen.BlockSystem(n=m.n, nb=m.nb, row_ptr=m.row_ptr, col=m.col, val=v.reshape(m.val.shape), rhs=m.rhs)
        d.set_values(v)
        d.setup(retry_fp32=retry, **kw)
        x2, its2, rel2, ok2 = d.solve(m.rhs, TOL, 3000)
        inf2 = d.info()
        true_rel = np.linalg.norm(m.rhs - orc.spmv(sp, x2)) / np.linalg.norm(m.rhs)
        print(f"retry_fp32={retry}: {its2} its (first solve {its}), inner {inf2['last_inner']}, fallback {inf2['last_fallback_fp64']}")
        assert ok2 and true_rel <= 1.01 * TOL
        if retry == 0:
            assert inf2["last_fallback_fp64"] == 0 and its2 < its / 3
        else:
            assert inf2["last_fallback_fp64"] == 1
