"""Folded target meshes on the GPU in the reference's first-hit mode
(cd_design_set_first_hit(d, 1)): the raster of a transport step and the
inverse map of a height step equal the reference's own (oracle/_ref, whose
BVH query, bvh.cpp:196-266, picks among overlapping triangles), where the
default lowest-index mode does not."""
import numpy as np
import pytest

from poisson_caustic_design_amd import capi
from tests.helpers import small_image

pytestmark = pytest.mark.gpu

M, W = 16, 0.5
N = 4 * M


def _design(ctx, img, mode):
    cd = capi.CausticDesign(ctx)
    cd.set_mesh_resolution(M, M)
    cd.set_domain_resolution(N, N)
    cd.set_mesh_size(W, W)
    cd.set_lens_focal_length(1.5)
    cd.set_lens_thickness(0.1)
    o = capi.default_opts(tol=1e-10, precision=capi.PREC_FP64)
    cd.set_solve_opts(o, o)
    cd.initialize_solvers(img)
    cd.set_first_hit(mode)
    return cd


def _folded(tp, seed):
    """Interior vertices moved up to 1.2 cell widths: inverted, overlapping
    triangles; the boundary stays, so every sample is still covered."""
    tp = tp.copy()
    inner = np.zeros((M, M), bool)
    inner[1:-1, 1:-1] = True
    rng = np.random.default_rng(seed)
    tp[inner.ravel(), :2] += rng.uniform(-1.2, 1.2, (inner.sum(), 2)) * W / (M - 1)
    return tp


@pytest.mark.parametrize("seed", [0, 1])
def test_first_hit_mode_matches_the_reference_on_folded_meshes(ctx, reflib, seed):
    from oracle.bind import RefDesign
    print(f"first-hit parity: M={M} N={N} seed={seed}", flush=True)
    img = small_image(3, N)
    rd = RefDesign(reflib, M, M, N, N, W, W, 1.5, 0.1, 1)
    rd.init(img)
    tp = _folded(rd.target_points(), seed)
    rd.set_target_points(tp)
    cds = [_design(ctx, img, mode) for mode in (1, 0)]
    for cd in cds:
        cd.set("target_points", tp)
        assert cd.count_folds() > 0
    # inverse map of the folded mesh (mesh.cpp:348-409), computed by a height step
    ref_inv, n = rd.inverted_map()
    assert n == M * M
    invs = []
    for cd in cds:
        cd.perform_height_map_iteration()
        invs.append(cd.get("inverse_map")[:, :2])
    assert np.abs(invs[0] - ref_inv).max() <= 1e-12 * W
    # raster of the folded mesh's errors (mesh.cpp:234-288), first thing of a transport step
    rd.transport_step()
    ref_r = rd.grid(2)
    rs = []
    for cd in cds:
        cd.perform_transport_iteration()
        rs.append(cd.raster)
    scale = np.abs(ref_r).max()
    assert np.abs(rs[0] - ref_r).max() <= 1e-9 * scale
    # the lowest-index mode picks other triangles where they overlap
    assert np.abs(rs[1] - ref_r).max() > 1e-6 * scale
    assert np.abs(invs[1] - ref_inv).max() > 1e-9 * W


def test_first_hit_mode_rejects_unknown_modes(ctx):
    cd = capi.CausticDesign(ctx)
    with pytest.raises(capi.CDError):
        cd.set_first_hit(2)
    cd = _design(ctx, small_image(3, N), 0)
    assert cd.L.cd_design_set_first_hit(cd.d, 2) == capi.BAD_ARG
    assert cd.L.cd_design_set_first_hit(cd.d, -1) == capi.BAD_ARG
    assert cd.L.cd_design_set_first_hit(None, 1) == capi.BAD_ARG
