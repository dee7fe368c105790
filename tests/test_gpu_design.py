"""GPU parity of the design pipeline (Caustic_design through the C ABI)
against the FP64 oracle restatement of caustic_design.cpp."""
import numpy as np
import pytest

from oracle.bind import RED_BLACK, STOP_RESIDUAL, OracleDesign
from poisson_caustic_design_amd import capi
from tests.helpers import small_image

pytestmark = pytest.mark.gpu


def make_pair(ctx, oracle, M, tol=1e-12, precision=capi.PREC_MIXED, seed=3, W=0.5):
    N = 4 * M
    img = small_image(seed, N)
    cd = capi.CausticDesign(ctx)
    cd.set_mesh_resolution(M, M)
    cd.set_domain_resolution(N, N)
    cd.set_mesh_size(W, W)
    cd.set_lens_focal_length(1.5)
    cd.set_lens_thickness(0.1)
    o = capi.default_opts(tol=tol, precision=precision)
    cd.set_solve_opts(o, o)
    cd.initialize_solvers(img)
    od = OracleDesign(oracle, M, M, N, N, W, W, 1.5)
    od.set_solver(STOP_RESIDUAL, tol, tol, 1000000, RED_BLACK, 10)
    od.init(img)
    return cd, od


def test_init_parity(ctx, oracle):
    cd, od = make_pair(ctx, oracle, 16)
    assert np.array_equal(cd.pixels, od.grid(6))
    ta = cd.target_areas
    assert np.abs(ta - od.vec(2)).max() <= 1e-14 * np.abs(od.vec(2)).max()
    tp = cd.get("target_points")
    assert np.array_equal(tp[:, :2], od.vec(0)) and (tp[:, 2] == 0).all()


@pytest.mark.parametrize("precision", [capi.PREC_MIXED, capi.PREC_FP64])
def test_transport_steps_parity(ctx, oracle, precision):
    cd, od = make_pair(ctx, oracle, 16, precision=precision)
    W = 0.5
    for it in range(4):
        s_gpu = cd.perform_transport_iteration()
        rc, s_orc = od.transport_step()
        assert rc == 0
        assert np.abs(cd.errors - od.vec(3)).max() < 1e-9, it
        r = cd.raster
        assert np.abs(r - od.grid(2)).max() < 1e-9 * np.abs(od.grid(2)).max(), it
        tp = cd.get("target_points")[:, :2]
        assert np.abs(tp - od.vec(0)).max() / W < 1e-10, it
        assert s_gpu == pytest.approx(s_orc, rel=1e-6)
        phi_s, _ = cd.last_solve_stats()
        assert phi_s.rel_residual <= 1e-12


def test_height_steps_parity(ctx, oracle):
    cd, od = make_pair(ctx, oracle, 12)
    for _ in range(3):
        cd.perform_transport_iteration()
        od.transport_step()
    for it in range(3):
        u_gpu = cd.perform_height_map_iteration(it)
        rc, u_orc = od.height_step()
        assert rc == 0
        z = cd.get("source_points")[:, 2]
        zo = od.vec(1)
        assert np.abs(z - zo).max() <= 1e-9 * max(np.abs(zo).max(), 1e-30), it
        # h is defined up to the Neumann constant: compare mean-free
        h, ho = cd.h, od.grid(1)
        assert np.abs((h - h.mean()) - (ho - ho.mean())).max() <= 1e-9 * np.abs(ho).max(), it
        nx = cd.get("normals")
        assert np.abs(nx[0] - od.vec(4)).max() <= 1e-12 * np.abs(od.vec(4)).max()


def test_run_loop_matches_manual_steps(ctx, oracle):
    a, _ = make_pair(ctx, oracle, 12, tol=1e-10)
    b, _ = make_pair(ctx, oracle, 12, tol=1e-10)
    rs = a.run(5, 0.0, 2)
    assert rs.transport_steps == 5 and rs.height_steps == 2
    for _ in range(5):
        b.perform_transport_iteration()
    for _ in range(2):
        b.perform_height_map_iteration()
    assert np.array_equal(a.get("target_points"), b.get("target_points"))
    assert np.array_equal(a.get("source_points"), b.get("source_points"))


def test_raster_miss_is_an_error_not_exit(ctx, oracle):
    cd, _ = make_pair(ctx, oracle, 10)
    tp = cd.get("target_points")
    M = 10
    # pull the bottom-right corner vertex inside: samples near that corner are uncovered
    tp[M * M - 1, :2] = [0.5 * 0.9, 0.5 * 0.9]
    cd.set("target_points", tp)
    with pytest.raises(capi.CDError) as ei:
        cd.perform_transport_iteration()
    assert ei.value.status == capi.RASTER_MISS


def test_laplacian_smoothing_parity(ctx, oracle):
    M = 12
    cd, _ = make_pair(ctx, oracle, M)
    rng = np.random.default_rng(21)
    tp = cd.get("target_points")
    tp[:, :2] += 0.003 * rng.standard_normal((M * M, 2))
    cd.set("target_points", tp)
    f = float(np.float32(0.1))
    cd.laplacian_smoothing(f)
    exp = oracle.laplacian_smoothing(tp[:, :2], M, M, f)
    assert np.array_equal(cd.get("target_points")[:, :2], exp)


def test_raster_miss_recovery_smooths_and_continues(ctx, oracle):
    """caustic_design.cpp:215-218: a miss smooths the target mesh (factor 0.1f)
    and the step reports NaN; the driver loop keeps going."""
    M = 10
    cd, _ = make_pair(ctx, oracle, M)
    tp = cd.get("target_points")
    tp[M * M - 1, :2] = [0.5 * 0.9, 0.5 * 0.9]
    cd.set("target_points", tp)
    cd.set_miss_recovery(True)
    step = cd.perform_transport_iteration()
    assert np.isnan(step)
    exp = oracle.laplacian_smoothing(tp[:, :2], M, M, float(np.float32(0.1)))
    assert np.array_equal(cd.get("target_points")[:, :2], exp)
    rs = cd.run(3, float(np.float32(0.01)), 0, allow=(capi.OK, capi.RASTER_MISS))
    assert rs.transport_steps >= 1


def test_c1_design_matches_oracle_fixture(ctx):
    """The complete C1 design (100x100 mesh, 400x400 grid, gauss5 seed 0,
    <=50 OT steps below 0.01f, 3 height steps) on the GPU against the oracle's
    FP64 design (tests/golden/c1_design.npz, both solving to residual 1e-10)."""
    import os
    from poisson_caustic_design_amd import synth
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "c1_design.npz"))
    img, prm = synth.target_for("gauss5", 0)
    M, N, W = prm["mesh_res"], prm["res"], prm["width"]
    cd = capi.CausticDesign(ctx)
    cd.set_mesh_resolution(M, M)
    cd.set_domain_resolution(N, N)
    cd.set_mesh_size(W, prm["height"])
    cd.set_lens_focal_length(prm["focal_l"])
    cd.set_lens_thickness(0.1)
    o = capi.default_opts(tol=float(g["tol"]))
    cd.set_solve_opts(o, o)
    cd.initialize_solvers(img)
    conv = float(np.float32(0.01))
    steps = []
    for _ in range(50):
        steps.append(cd.perform_transport_iteration())
        if steps[-1] < conv:
            break
    for i in range(3):
        cd.perform_height_map_iteration(i)
    ref_steps = g["steps"]
    assert len(steps) == len(ref_steps)
    assert np.allclose(steps, ref_steps, rtol=1e-6)
    dxy = np.abs(cd.get("target_points")[:, :2] - g["target_xy"]).max() / W
    z, zr = cd.get("source_points")[:, 2], g["heights"]
    dz = np.abs(z - zr).max() / np.abs(zr).max()
    print(f"C1: {len(steps)} OT steps, max |dxy|/W {dxy:.2e}, max |dz|/max|z| {dz:.2e}")
    assert dxy < 1e-11 and dz < 1e-10  # measured 1.0e-13 and 7.4e-13
    cd.close()


def test_fold_count_matches_host_counter(ctx, oracle, rng):
    """cd_design_count_folds on host-forced meshes: none on the initial mesh,
    two after an interior vertex is pushed past its right neighbour (the
    hand-computed case of tests/test_folds.py), and the host counter's count
    on random perturbations."""
    from tests.helpers import fold_count
    M, W = 12, 0.5
    cd, _ = make_pair(ctx, oracle, M)
    tp = cd.get("target_points").copy()
    assert cd.count_folds() == 0 == fold_count(tp, M, M)
    dx = W / (M - 1)
    forced = tp.copy()
    v = 5 * M + 5
    forced[v, 0] += 1.5 * dx
    cd.set("target_points", forced)
    assert cd.count_folds() == 2 == fold_count(forced, M, M)
    for amp in (0.2, 0.6, 1.2):
        p = tp.copy()
        p[:, :2] += amp * dx * rng.standard_normal((M * M, 2))
        cd.set("target_points", p)
        assert cd.count_folds() == fold_count(p, M, M), amp


def test_folding_step_is_reported(ctx, oracle):
    """A transport step with a step size 400x the reference's folds the mesh:
    the run reports the inverted triangles (the count the host counter finds
    on the final mesh) and, with the fold check on, the step fails with
    CD_ERR_FOLDED."""
    from tests.helpers import fold_count
    M, W = 12, 0.5
    N = 4 * M
    img = small_image(3, N)

    def make(fail):
        cd = capi.CausticDesign(ctx)
        cd.set_mesh_resolution(M, M)
        cd.set_domain_resolution(N, N)
        cd.set_mesh_size(W, W)
        cd.set_lens_focal_length(1.5)
        cd.set_lens_thickness(0.1)
        o = capi.default_opts(tol=1e-10)
        cd.set_solve_opts(o, o)
        cd.step_size = 400 * float(np.float32(0.05))
        cd.initialize_solvers(img)
        cd.set_fold_check(fail)
        return cd

    cd = make(False)
    rs = cd.run(3, 0.0, 0, allow=(capi.OK, capi.RASTER_MISS))
    tp = cd.get("target_points")
    assert rs.folded_steps >= 1 and rs.first_folded_step == 0
    if rs.status == capi.OK:
        assert rs.folded_triangles == fold_count(tp, M, M) > 0
    cd = make(True)
    with pytest.raises(capi.CDError) as e:
        cd.perform_transport_iteration()
    assert e.value.status == capi.FOLDED
    assert cd.count_folds() == fold_count(cd.get("target_points"), M, M) > 0


def test_reference_step_on_smooth_target_does_not_fold(ctx, oracle):
    cd, _ = make_pair(ctx, oracle, 16)
    cd.set_fold_check(True)
    rs = cd.run(5, 0.0, 1)
    assert rs.status == capi.OK and rs.folded_steps == 0 and rs.first_folded_step == -1
