"""Full designs at the benchmark configurations' real sizes, checked through
size-independent properties (the oracle cannot run these sizes in a test):

* every Poisson solve ends at FP64 relative residual <= 1e-6;
* the target areas sum to W*H (polygon_utils.cpp:382-386);
* boundary vertices stay on their edges and corners stay fixed
  (mesh.cpp:489-512; interior vertices are not clamped, as in the reference);
* heights are >= 0 after the min-shift (mesh.cpp:726-740);
* the OT loop obeys the reference's stop rule (main.cpp:243-256);
* the learned and plain forms end at the same design.

C2: 256^2 mesh, 1024^2 grid, 12 sharp-edged rectangles (`rects12`), plain
and learned (`weights/learned_grid1024.npz`).  C5: 1024^2 mesh, 4096^2 grid,
clipped log-normal noise (`lognormal`), plain and learned
(`weights/learned_grid4096.npz`)."""
import os

import numpy as np
import pytest

from poisson_caustic_design_amd import capi, learned, synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONV = float(np.float32(0.01))


def run_design(ctx, config, seed, weights=None):
    img, prm = synth.target_for(config, seed)
    M, N, W, H = prm["mesh_res"], prm["res"], prm["width"], prm["height"]
    net = None
    if weights:
        net = learned.load_net(ctx, os.path.join(ROOT, "poisson_caustic_design_amd", "weights",
                                                 weights))
        ctx.set_net(net)
    try:
        cd = capi.CausticDesign(ctx)
        cd.set_mesh_resolution(M, M)
        cd.set_domain_resolution(N, N)
        cd.set_mesh_size(W, H)
        cd.set_lens_focal_length(prm["focal_l"])
        cd.set_lens_thickness(0.1)
        o = capi.default_opts(tol=1e-6, learned_init=int(net is not None))
        cd.set_solve_opts(o, o)
        cd.initialize_solvers(img)
        areas = cd.target_areas.copy()
        # the first OT step alone (its positions are compared before the OT
        # loop amplifies solve differences), then the reference driver loop
        first = cd.perform_transport_iteration()
        res1 = cd.last_solve_stats()[0].rel_residual
        tp1 = cd.get("target_points").copy()
        rs = cd.run(49 if first >= CONV else 0, CONV, 3)
        rs.transport_steps += 1
        out = {"rs": rs, "areas": areas, "tp": cd.get("target_points"), "tp1": tp1, "res1": res1,
               "sp": cd.get("source_points"), "M": M, "W": W, "H": H}
        cd.close()
    finally:
        if net is not None:
            ctx.set_net(None)
            net.close()
    return out


def check_properties(d):
    rs, M, W, H = d["rs"], d["M"], d["W"], d["H"]
    assert rs.status == capi.OK
    assert max(rs.max_phi_residual, rs.max_h_residual, d["res1"]) <= 1e-6
    assert rs.transport_steps == 50 or rs.last_step < CONV
    assert abs(d["areas"].sum() - W * H) <= 1e-12 * W * H
    tp = d["tp"].reshape(M, M, 3)
    assert np.all(tp[:, 0, 0] == 0.0) and np.all(tp[:, -1, 0] == tp[0, -1, 0])
    assert np.all(tp[0, :, 1] == 0.0) and np.all(tp[-1, :, 1] == tp[-1, 0, 1])
    for cy in (0, -1):
        for cx in (0, -1):
            assert tp[cy, cx, 0] == (0.0 if cx == 0 else tp[0, -1, 0])
    assert np.all(np.isfinite(tp))
    z = d["sp"][:, 2]
    assert np.all(np.isfinite(z)) and z.min() >= 0.0 and z.max() > 0.0


# final-design agreement is asserted where the OT loop does not amplify solve
# differences: on the log-normal C5 target each OT step multiplies a position
# difference by ~100 (plain SOR at 1e-6 and at 1e-9 end 4e-3*W apart there),
# so only the first step is compared
@pytest.mark.parametrize("config,seed,weights,final_close",
                         [("rects12", 0, "learned_grid1024.npz", True),
                          ("lognormal", 2, "learned_grid4096.npz", False)])
def test_full_size_design_plain_and_learned(ctx, config, seed, weights, final_close):
    plain = run_design(ctx, config, seed)
    lrn = run_design(ctx, config, seed, weights)
    for d in (plain, lrn):
        check_properties(d)
    assert lrn["rs"].learned_applied > 0
    W = plain["W"]
    dp1 = np.abs(plain["tp1"][:, :2] - lrn["tp1"][:, :2]).max() / W
    dp = np.abs(plain["tp"][:, :2] - lrn["tp"][:, :2]).max() / W
    zr = plain["sp"][:, 2].max()
    dz = np.abs(plain["sp"][:, 2] - lrn["sp"][:, 2]).max() / zr
    print(f"{config} seed {seed}: OT steps {plain['rs'].transport_steps}/"
          f"{lrn['rs'].transport_steps}, sweeps {plain['rs'].phi_sweeps + plain['rs'].h_sweeps}/"
          f"{lrn['rs'].phi_sweeps + lrn['rs'].h_sweeps}, learned iterations "
          f"{lrn['rs'].learned_applied}, first-step difference {dp1:.2e} W, final design "
          f"difference {dp:.2e} W, {dz:.2e} of max height")
    assert dp1 <= 1e-7
    if final_close:
        assert plain["rs"].transport_steps == lrn["rs"].transport_steps
        assert dp <= 1e-6 and dz <= 1e-4
