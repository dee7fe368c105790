"""Writes tests/golden/c1_design.npz: the oracle's (FP64, red-black SOR to
relative residual 1e-10) complete design for the C1 configuration -- 100x100
mesh, gauss5 seed 0 (256x256 u8 target, resized to the 400x400 grid), width
0.5, focal 1.5, <=50 transport steps stopping below 0.01f, then 3 height
steps (main.cpp:243-262).  The oracle is pinned bit-for-bit to the
reference's own code (tests/test_oracle_*.py); the GPU design is checked
against this file at this size by tests/test_gpu_design.py.

    python tests/golden/make_c1_design.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.bind import RED_BLACK, STOP_RESIDUAL, Oracle, OracleDesign  # noqa: E402
from poisson_caustic_design_amd import synth  # noqa: E402

TOL = 1e-10


def main():
    img, prm = synth.target_for("gauss5", 0)
    M, N, W, H, f = prm["mesh_res"], prm["res"], prm["width"], prm["height"], prm["focal_l"]
    od = OracleDesign(Oracle(), M, M, N, N, W, H, f)
    od.set_solver(STOP_RESIDUAL, TOL, TOL, 2000000, RED_BLACK, 10)
    od.init(img)
    conv = float(np.float32(0.01))
    steps, t0 = [], time.time()
    for _ in range(50):
        rc, s = od.transport_step()
        assert rc == 0, rc
        steps.append(s)
        print(f"OT {len(steps) - 1}: step {s:.6f} sweeps {od.last_sweeps()} "
              f"residual {od.last_residual():.2e} {time.time() - t0:.0f}s", flush=True)
        if s < conv:
            break
    for _ in range(3):
        rc, u = od.height_step()
        assert rc == 0, rc
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_design.npz")
    np.savez_compressed(out, target_xy=od.vec(0), heights=od.vec(1), steps=np.array(steps),
                        tol=TOL, mesh_res=M, grid=N, width=W, focal_l=f)
    print("wrote", out, f"{time.time() - t0:.0f}s")


if __name__ == "__main__":
    main()
