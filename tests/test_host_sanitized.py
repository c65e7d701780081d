"""The library's host-side set-up under the host sanitizers (AddressSanitizer + UndefinedBehaviorSanitizer): level
schedules, colouring and the colour-permuted pattern, row-slice layouts and level runs, the aggregation AMG set-up and
the well Schur plan, built with g++ without device code (ressimstudio_amd/host_check, built by build.py) and run on the
test systems of the GPU suite, including the tall-well systems with the long dense well rows.  Every result is also
checked for what it must satisfy (valid level order and colouring, every entry placed once, AMG levels shrinking, the
well plan reproducing the reduced system).  CPU only."""
import os
import subprocess

import numpy as np
import pytest

from ressimstudio_amd import gen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "ressimstudio_amd", "host_check")


def _systems():
    return [gen.small(), gen.chain(), gen.block_laplace(40, 30, 3, 11), gen.reservoir(12, 10, 3, 2, model=0, sigma=1.5, wells=2, seed=5),
            gen.reservoir(10, 9, 3, 4, model=0, sigma=0.8, wells=1, seed=6), gen.reservoir(16, 12, 4, 1, model=0, sigma=1.0, wells=2, seed=8),
            gen.reservoir(12, 30, 14, 3, model=1, sigma=2.5, kvkh=0.1, wells=5, seed=7),
            gen.reservoir(24, 24, 60, 3, model=1, sigma=2.0, kvkh=0.1, wells=3, seed=9),
            gen.reservoir(12, 12, 40, 2, model=0, sigma=1.0, wells=2, seed=10)]


def _write(path, systems):
    with open(path, "wb") as f:
        f.write(np.int32(len(systems)).tobytes())
        for s in systems:
            f.write(np.array([s.n, s.nb], np.int32).tobytes())
            f.write(np.ascontiguousarray(s.row_ptr, np.int32).tobytes())
            f.write(np.ascontiguousarray(s.col, np.int32).tobytes())
            f.write(np.ascontiguousarray(s.val, np.float64).tobytes())
            j = s.full
            f.write(np.int32(1 if j else 0).tobytes())
            if j:
                f.write(np.array([j["n"], j["na"], j["nv"], j["nw"]], np.int32).tobytes())
                f.write(np.ascontiguousarray(j["row_ptr"], np.int32).tobytes())
                f.write(np.ascontiguousarray(j["col"], np.int32).tobytes())
                f.write(np.ascontiguousarray(j["val"], np.float64).tobytes())


def test_host_setup_under_sanitizers(tmp_path):
    if not os.path.exists(BIN):
        from ressimstudio_amd import build
        build.build_host_check()
    path = str(tmp_path / "systems.bin")
    systems = _systems()
    _write(path, systems)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    p = subprocess.run([BIN, path], capture_output=True, text=True, timeout=600, env=env)
    report = p.stdout + p.stderr
    assert p.returncode == 0, report[-4000:]
    assert "AddressSanitizer" not in report and "runtime error" not in report and "LeakSanitizer" not in report, report[-4000:]
    assert report.count("system ") == len(systems) and "0 failed checks" in report, report[-2000:]
