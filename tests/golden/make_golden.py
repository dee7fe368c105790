"""Write the golden vectors of tests/golden/ref_vectors.npz by running the
reference's own code (oracle/_ref/libcdref.so, built from the reference
sources by oracle/Makefile) on small seeded inputs.

    python tests/golden/make_golden.py

Contents (all float64):
  sor_D, sor_phi0, sor_phi_37     40x33 right-hand side, warm start, 37 sweeps (threads 1)
  sor_phi_conv                    same, to the reference's max-update 1e-7
  grid_A, grid_B, grad_x, grad_y, div, subavg, minmax, bil_xy, bil_val
  des_image                       40x40 image (M = 10, N = 40, W = 0.5, f = 1.5)
  des_target_areas                after initialize_solvers
  des_points_k, des_step_k        target points / step size after transport step k (k=1..3)
  des_phi_3                       phi after 3 transport steps
  des_z_h, des_h                  source heights / h after 3 transport + 2 height steps
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.bind import RefDesign, RefLib  # noqa: E402
from poisson_caustic_design_amd import synth  # noqa: E402


def main():
    R = RefLib()
    rng = np.random.default_rng(20240601)
    out = {}
    D = rng.standard_normal((40, 33)); D -= D.mean()
    p0 = 0.1 * rng.standard_normal((40, 33))
    out["sor_D"], out["sor_phi0"] = D, p0
    out["sor_phi_37"] = R.poisson_solver(D, p0, 37, 1e-300, 1)
    out["sor_phi_conv"] = R.poisson_solver(D, p0, 100000, 1e-7, 1)
    A = rng.standard_normal((17, 23)); B = rng.standard_normal((17, 23))
    out["grid_A"], out["grid_B"] = A, B
    out["grad_x"], out["grad_y"] = R.gradient(A)
    out["div"] = R.divergence(A, B)
    out["subavg"] = R.subtract_average(A)
    out["minmax"] = R.scale_minmax(A)
    xy = np.column_stack([rng.uniform(-2, 25, 32), rng.uniform(-2, 19, 32)])
    out["bil_xy"] = xy
    out["bil_val"] = np.array([R.bilinear(A, x, y) for x, y in xy])
    M, N, W = 10, 40, 0.5
    img = synth.resize_nearest(synth.gray_from_u8(synth.gauss5(11, 64)), N, N)
    out["des_image"] = img
    d = RefDesign(R, M, M, N, N, W, W, 1.5)
    d.init(img)
    out["des_target_areas"] = d.vector(1)
    for k in range(1, 4):
        out[f"des_step_{k}"] = np.array([d.transport_step()])
        out[f"des_points_{k}"] = d.target_points()
    out["des_phi_3"] = d.grid(0)
    d.height_step()
    d.height_step()
    out["des_z_h"] = d.source_points()[:, 2].copy()
    out["des_h"] = d.grid(1)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_vectors.npz")
    np.savez_compressed(path, **out)
    print(path, sum(v.nbytes for v in out.values()), "bytes")


if __name__ == "__main__":
    main()
