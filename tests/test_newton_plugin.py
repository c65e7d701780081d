"""The reference engine's Newton loop on its demo waterflood (makeDemoCase, 6 report steps) with the MI355X solver
registered in the engine's existing hook (setDistributedSolver) through include/ressim_hip_solver.hpp, against
the same engine with its CPU solver.  The harness binary (oracle/_ref/newton_harness) is built from the reference
sources where they are mounted; the tests skip where it was never built.

The reference's own plug-in tests assert FOPT within 1e-5 / 1e-6 relative and no fallbacks
(run_tests.cpp:3771-3773, c29_tests.inc:295)."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "newton_harness")


def run_harness(*args, timeout=300):
    if not os.path.exists(HARNESS):
        pytest.skip("oracle/_ref/newton_harness not built (needs the reference sources)")
    out = subprocess.run([HARNESS, *map(str, args)], capture_output=True, text=True, timeout=timeout, check=True).stdout
    runs = {}
    for line in out.splitlines():
        line = line.strip()
        if line.startswith("{"):
            d = json.loads(line)
            runs[d.get("run", "meta")] = d
    return runs


def _compare_steps(cpu, dev, tol):
    assert len(cpu["steps"]) == len(dev["steps"])
    for a, b in zip(cpu["steps"], dev["steps"]):
        assert a["t"] == b["t"]
        for k in ("fopt", "fpr", "fwpt"):
            assert abs(a[k] - b[k]) <= tol * max(abs(a[k]), 1.0), (k, a, b)


def test_hook_failure_falls_back_to_the_cpu_path():
    # without a device every hook call fails and the engine re-solves on its CPU path: same answers, all fallbacks
    import ressimstudio_amd as rs
    if rs.device_count() > 0:
        pytest.skip("a device is present (covered by the GPU test)")
    r = run_harness(3, "cpr", "gs")
    cpu, dev = r["cpu"], r["device"]
    assert dev["dd_solves"] > 0 and dev["dd_fallbacks"] == dev["dd_solves"]
    assert dev["newton"] == cpu["newton"]
    _compare_steps(cpu, dev, 1e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("precond,smoother", [("cpr", "gs"), ("cpr_mc", "jacobi"), ("bilu0", "gs")])
def test_demo_waterflood_with_device_solver(precond, smoother):
    r = run_harness(6, precond, smoother)
    cpu, dev = r["cpu"], r["device"]
    print(json.dumps({"cpu": {k: cpu[k] for k in ("newton", "linear")}, "device": {k: dev[k] for k in ("newton", "linear", "dd_solves")},
                      "device_seconds": r["meta"]["device_seconds"]}))
    assert dev["dd_solves"] > 0 and dev["dd_fallbacks"] == 0, r["meta"]
    # identical Newton convergence: the same number of Newton iterations in every report step
    assert [s["newton"] for s in dev["steps"]] == [s["newton"] for s in cpu["steps"]]
    _compare_steps(cpu, dev, 1e-5)
