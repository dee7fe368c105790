"""Writes tests/golden/c1_reference_flow.npz: the reference's own
Caustic_design (src/*.cpp compiled where they lie into oracle/_ref/libcdref.so)
run through the flow of its CLI with default solver settings (main.cpp:224-266:
max-update stops 1e-7 / 1e-8, omega 2/(1+3.14159265/width), --threads 1, the
main.cpp default), on the C1 configuration: 100x100 mesh, gauss5 seed 0
(256x256 u8 target read as RGB, resized to the 400x400 grid), width 0.5,
focal 1.5, thickness 0.1, conv_tres 0.01f.  The GPU tests compare the CLI
drop-ins' designs with it and with the converged oracle design
(c1_design.npz) to bound the stop rule's own error.

    python tests/golden/make_c1_reference_flow.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.bind import RefDesign, RefLib  # noqa: E402
from poisson_caustic_design_amd import synth  # noqa: E402


def main():
    img, prm = synth.target_for("gauss5", 0)
    M, N, W, H, f = prm["mesh_res"], prm["res"], prm["width"], prm["height"], prm["focal_l"]
    rd = RefDesign(RefLib(), M, M, N, N, W, H, f, float(np.float32(0.1)), 1)
    rd.init(img)
    conv = float(np.float32(0.01))
    steps, t0 = [], time.time()
    for itr in range(50):
        s = rd.transport_step()
        steps.append(s)
        print(f"OT {itr}: step {s:.6f} {time.time() - t0:.0f}s", flush=True)
        if s < conv:
            break
    for _ in range(3):
        rd.height_step()
    tp, sp = rd.target_points(), rd.source_points()
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_reference_flow.npz")
    np.savez_compressed(out, target_xy=tp[:, :2], heights=sp[:, 2], steps=np.array(steps),
                        mesh_res=M, grid=N, width=W, focal_l=f)
    print("wrote", out, f"{time.time() - t0:.0f}s")


if __name__ == "__main__":
    main()
