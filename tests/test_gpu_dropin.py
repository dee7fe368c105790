"""The C++ drop-ins used the way the reference's own main.cpp uses them: a
program compiled unchanged against include/caustic_design.h and
include/solver.h (tests/dropin/dropin_main.cpp) produces the same lens as the
same calls through the C ABI, and the free poisson_solver converges."""
import os
import subprocess

import numpy as np
import pytest

from poisson_caustic_design_amd import capi

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(__file__), "dropin", "dropin_main")


def test_dropin_program_matches_c_abi(ctx, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("tests/dropin/dropin_main not built (python -m poisson_caustic_design_amd.build)")
    M = 16
    r = subprocess.run([EXE, str(M), str(tmp_path) + "/"], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout[-600:], r.stderr[-300:])
    assert r.returncode == 0
    vals = dict(line.split(" ", 1) for line in r.stdout.strip().splitlines() if " " in line)
    assert int(vals["mesh_points"]) == M * M
    assert float(vals["poisson_max_residual"]) < 1e-8
    # the same design through the C ABI with the reference's default solves
    N = 4 * M
    img = np.fromfile(tmp_path / "image.f64", dtype=np.float64).reshape(N, N)
    cd = capi.CausticDesign(ctx)
    cd.set_mesh_resolution(M, M)
    cd.set_domain_resolution(N, N)
    cd.set_mesh_size(1.0, 1.0)
    cd.set_lens_focal_length(1.5)
    cd.set_lens_thickness(0.2)
    cd.initialize_solvers(img)
    steps = 0
    for _ in range(50):
        steps += 1
        if cd.perform_transport_iteration() < float(np.float32(0.01)):
            break
    for i in range(3):
        cd.perform_height_map_iteration(i)
    assert steps == int(vals["steps"])
    assert float(vals["h_abs_sum"]) == pytest.approx(np.abs(cd.get("h")).sum(), rel=1e-14)
    cd.save_solid_obj_source(str(tmp_path / "api.obj"))
    assert (tmp_path / "output.obj").read_bytes() == (tmp_path / "api.obj").read_bytes()
    cd.close()
