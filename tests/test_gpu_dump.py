"""A real engine Jacobian (the reference's demo waterflood, first Newton iteration, RESSIM_DUMP format) solved on the
device through the C-ABI with every preconditioner: solution against the oracle's FP64 solve (<= 1e-6), true
relative residual <= 1e-8, and the recovered full solution (BHPs included) satisfies the engine's scalar system."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import ressimstudio_amd as rs
from ressimstudio_amd import dump

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
TOL = 1e-8


@pytest.fixture(scope="module")
def demo():
    z = np.load(os.path.join(HERE, "golden", "demo_jacobian.npz"))
    return {k: (int(z[k]) if z[k].ndim == 0 else z[k]) for k in z.files}


@pytest.mark.parametrize("opts", [dict(precond="cpr", smoother=rs.SMOOTH_GS), dict(precond="cpr_mc", smoother=rs.SMOOTH_JACOBI, smooth_steps=2),
                                  dict(precond="bilu0"), dict(precond="mcbilu0"),
                                  dict(precond="cpr_mc", smoother=rs.SMOOTH_JACOBI, smooth_steps=2, precision=rs.MIXED)],
                         ids=["cpr_gs", "cpr_mc_j2", "bilu0", "mcbilu0", "cpr_mc_mixed"])
def test_engine_jacobian(demo, orc, opts):
    s = dump.to_system(demo)
    d = rs.System(s).setup(**opts)
    x, its, rel, ok = d.solve(s.rhs, TOL, 2000)
    xo, its_o, _, ok_o = orc.krylov(s, s.rhs, "cpr", "bicgstab", 1e-12, 2000)
    assert ok and ok_o
    r = s.rhs - orc.spmv(s, x)
    assert np.linalg.norm(r) <= 1.01 * TOL * np.linalg.norm(s.rhs)
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
    J = sp.csr_matrix((demo["val"], demo["col"], demo["row_ptr"]), shape=(demo["n"], demo["n"]))
    xf = s.recover(demo["rhs"], x)
    assert np.linalg.norm(J @ xf - demo["rhs"]) <= 1e-6 * np.linalg.norm(demo["rhs"])


def test_true_impes_weights(demo, orc):
    # rsh_system_set_accumulation: the device's true-IMPES weights from the engine's accumulation blocks equal the
    # oracle's bit for bit (also with cells forced to fall back to quasi-IMPES, counted alike); the CPR built on them
    # solves the engine's system; clearing the blocks restores the quasi-IMPES weights
    s = dump.to_system(demo)
    acc = demo["acc"]
    ill = acc.copy()
    ill[::7, :, 1] = ill[::7, :, 0] * (1 + 1e-13)
    ill[3::11] = 0.0
    d = rs.System(s)
    for a in (acc, ill):
        d.set_accumulation(a).setup(precond="cpr_mc", smoother=rs.SMOOTH_JACOBI, smooth_steps=2)
        o = orc.cpr(s, a)
        assert np.array_equal(d.cpr_weights(), o.weights())
        info = d.info()
        assert info["cpr_true_impes"] == 1 and info["cpr_impes_fallbacks"] == o.fallbacks()
        x, its, rel, ok = d.solve(s.rhs, TOL, 2000)
        assert ok and np.linalg.norm(s.rhs - orc.spmv(s, x)) <= 1.01 * TOL * np.linalg.norm(s.rhs)
    d.set_accumulation(None).setup(precond="cpr_mc", smoother=rs.SMOOTH_JACOBI, smooth_steps=2)
    assert np.array_equal(d.cpr_weights(), orc.cpr(s).weights())
    assert d.info()["cpr_true_impes"] == 0
