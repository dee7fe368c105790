"""The oracle against golden vectors written by the reference's own code
(tests/golden/make_golden.py).  Runs without the reference build."""
import os

import numpy as np
import pytest

from oracle.bind import OracleDesign

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_vectors.npz")


@pytest.fixture(scope="module")
def g():
    return np.load(GOLD, allow_pickle=False)


def test_sor_golden(oracle, g):
    a, st = oracle.poisson_sor(g["sor_D"], g["sor_phi0"], 37, max_update_tol=1e-300)
    assert np.array_equal(a, g["sor_phi_37"])
    b, _ = oracle.poisson_sor(g["sor_D"], g["sor_phi0"], 100000, max_update_tol=1e-7)
    assert np.array_equal(b, g["sor_phi_conv"])


def test_grid_golden(oracle, g):
    A, B = g["grid_A"], g["grid_B"]
    gx, gy = oracle.gradient(A)
    assert np.array_equal(gx, g["grad_x"]) and np.array_equal(gy, g["grad_y"])
    assert np.array_equal(oracle.divergence(A, B), g["div"])
    assert np.array_equal(oracle.subtract_average(A), g["subavg"])
    assert np.array_equal(oracle.scale_minmax(A), g["minmax"])
    vals = np.array([oracle.bilinear(A, x, y) for x, y in g["bil_xy"]])
    assert np.array_equal(vals, g["bil_val"])


def test_design_golden(oracle, g):
    M, N, W = 10, 40, 0.5
    d = OracleDesign(oracle, M, M, N, N, W, W, 1.5)
    d.init(g["des_image"])
    assert np.allclose(d.vec(2), g["des_target_areas"], rtol=1e-13, atol=0)
    for k in range(1, 4):
        rc, s = d.transport_step()
        assert rc == 0
        assert s == pytest.approx(g[f"des_step_{k}"][0], rel=1e-10)
        assert np.abs(d.vec(0) - g[f"des_points_{k}"][:, :2]).max() < 1e-13
    assert np.abs(d.grid(0) - g["des_phi_3"]).max() < 1e-10
    d.height_step()
    d.height_step()
    assert np.abs(d.vec(1) - g["des_z_h"]).max() < 1e-13
    assert np.abs(d.grid(1) - g["des_h"]).max() < 1e-12
