"""The drop-in's reference-semantics path at real sizes: the max-update stop
rule (solver.cpp:92-146, stop at :142) with the reference's omega
(solver.cpp:71), mixed precision (FP32 sweeps on a BF16-rounded right-hand
side, FP64 defect correction) against the same iteration in FP64 sweeps.

Under the max-update rule the blocked sweeps read an FP32 right-hand side
(the FP64 residual of the last check rounded to FP32): the rule cannot see
smooth error, and the smooth part of a BF16 rounding of the initial residual
(2^-9 relative per cell) would stay in the result when the stop falls within
the first batch of sweeps (an FP64 restatement of that case ends 20x further
from the exact solution than the FP64 solve: 3.3e-5 against 1.7e-6 at this
size and warm start).  A stop is also confirmed on a fresh FP64 residual.
These tests show that the mixed result is as accurate as the FP64 one at
N = 1024, from a cold start and from a warm start whose whole solve lies in
the first batch of sweeps."""
import numpy as np
import pytest

from poisson_caustic_design_amd import capi, synth
from tests.helpers import dct_solve, smooth_rhs

pytestmark = pytest.mark.gpu


def mf_err(x, ref):
    return float(np.abs((x - x.mean()) - (ref - ref.mean())).max())


@pytest.mark.parametrize("start", ["cold", "warm"])
def test_max_update_mixed_as_accurate_as_fp64_at_1024(ctx, start):
    rng = np.random.default_rng(7)
    n = 1024
    b = smooth_rhs(rng, n, n)
    ex = dct_solve(b)
    phi0 = np.zeros_like(b)
    if start == "warm":
        # a converged solution perturbed at high frequency: the first sweeps'
        # updates exceed the threshold, and it is met again within one batch
        phi0 = ex + 1e-5 * rng.standard_normal(b.shape)
    kw = dict(stop=capi.STOP_MAX_UPDATE, tol=1e-7, omega=-1.0)
    if start == "warm":
        # one batch of sweeps between FP64 checks holds the whole solve
        kw["check_every"] = 4096
    g64, s64 = ctx.poisson_solve(b, phi0, capi.default_opts(precision=capi.PREC_FP64, **kw))
    gm, sm = ctx.poisson_solve(b, phi0, capi.default_opts(precision=capi.PREC_MIXED, **kw))
    e64, em = mf_err(g64, ex), mf_err(gm, ex)
    print(f"{start}: FP64 {s64.sweeps} sweeps err {e64:.3e}; mixed {sm.sweeps} sweeps, "
          f"{sm.checks} checks, err {em:.3e}")
    assert s64.status == capi.OK and sm.status == capi.OK
    assert sm.last_max_update < 1e-7
    if start == "warm":
        assert s64.sweeps < 4096  # the threshold is met within the first batch
    assert em <= 2.0 * e64 + 1e-12 * np.abs(ex).max()


def design_defaults(ctx, img, prm, precision, tol_scale=1.0):
    """Caustic_design with the reference's solve defaults (max-update 1e-7 for
    phi, 1e-8 for h, the reference's omega: cd_design_params_default), two
    transport steps and one height step."""
    M, N = prm["mesh_res"], prm["res"]
    cd = capi.CausticDesign(ctx)
    cd.set_mesh_resolution(M, M)
    cd.set_domain_resolution(N, N)
    cd.set_mesh_size(prm["width"], prm["height"])
    cd.set_lens_focal_length(prm["focal_l"])
    cd.set_lens_thickness(0.1)
    p = capi.default_opts(stop=capi.STOP_MAX_UPDATE, tol=1e-7 * tol_scale, omega=-1.0,
                          precision=precision)
    h = capi.default_opts(stop=capi.STOP_MAX_UPDATE, tol=1e-8 * tol_scale, omega=-1.0,
                          precision=precision)
    cd.set_solve_opts(p, h)
    cd.initialize_solvers(img)
    for _ in range(2):
        cd.perform_transport_iteration()
    cd.perform_height_map_iteration(0)
    out = (cd.get("target_points")[:, :2], cd.get("source_points")[:, 2])
    cd.close()
    return out


def test_design_defaults_mixed_within_fp64_tolerance_sensitivity_at_1024(ctx):
    """At the C2 size (256^2 mesh, 1024^2 grid, `rects12`), the mixed-precision
    design on the reference's default stop rule differs from the FP64-sweep
    design by no more than the FP64 design differs from itself at half the
    thresholds (the stop rule's own sensitivity)."""
    img, prm = synth.target_for("rects12", 0)
    W = prm["width"]
    tp64, z64 = design_defaults(ctx, img, prm, capi.PREC_FP64)
    tp64h, z64h = design_defaults(ctx, img, prm, capi.PREC_FP64, 0.5)
    tpm, zm = design_defaults(ctx, img, prm, capi.PREC_MIXED)
    zr = np.abs(z64).max()
    ctrl_p, ctrl_z = np.abs(tp64h - tp64).max() / W, np.abs(z64h - z64).max() / zr
    dp, dz = np.abs(tpm - tp64).max() / W, np.abs(zm - z64).max() / zr
    print(f"mixed vs FP64: positions {dp:.3e} W, heights {dz:.3e}; FP64 at half the "
          f"thresholds: {ctrl_p:.3e} W, {ctrl_z:.3e}")
    assert dp <= max(ctrl_p, 1e-12) and dz <= max(ctrl_z, 1e-10)
