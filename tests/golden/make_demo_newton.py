"""Writes tests/golden/demo_newton_cpu.json: the reference engine's own CPU run of its demo waterflood (makeDemoCase,
6 report steps, block CPR, quasi-IMPES), with the per-report-step summary and the per-Newton-iteration CNV / MB
trajectory the engine prints with RESSIM_DEBUG (simulator.cpp:3207-3214), from oracle/_ref/newton_harness (the engine
compiled from the reference sources by oracle/Makefile).

    python tests/golden/make_demo_newton.py
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "tests"))
from test_newton_plugin import HARNESS, harness_runs  # noqa: E402


def main():
    if not os.path.exists(HARNESS):
        sys.exit("oracle/_ref/newton_harness is not built (needs the reference sources)")
    runs, traj = harness_runs(6, "cpr", "gs")
    out = {"case": "makeDemoCase (case.cpp:427-480), schedule cut to 6 report steps, blockSolver, linearSolver 1 (CPR), cprWeights 0",
           "summary": runs["cpu"], "newton_trajectory": traj["cpu"]}
    path = os.path.join(ROOT, "tests", "golden", "demo_newton_cpu.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"{path}: {len(traj['cpu'])} Newton solves, {sum(len(t) for t in traj['cpu'])} iterations")


if __name__ == "__main__":
    main()
