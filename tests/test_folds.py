"""Host fold counter (tests/helpers.fold_count) pinned by hand-computed cases.

The reference never detects folds (mesh.cpp:485-532 steps interior vertices
unclamped); the library counts inverted triangles after every transport step
(cd_design_count_folds, cd_run_stats.folded_triangles), and the GPU tests
compare those counts with this counter."""
import numpy as np

from tests.helpers import fold_count, mesh_triangles


def regular(mx, my, W=2.0, H=2.0):
    xs, ys = np.meshgrid(np.linspace(0, W, mx), np.linspace(0, H, my))
    return np.stack([xs.ravel(), ys.ravel()], 1)


def test_triangulation_matches_reference_layout():
    # mesh.cpp:57-63 on a 3x3 mesh: 8 triangles, cell (0,0) first
    t = mesh_triangles(3, 3)
    assert t.shape == (8, 3)
    assert t[0].tolist() == [0, 1, 3] and t[1].tolist() == [3, 1, 4]
    assert t[7].tolist() == [7, 5, 8]


def test_regular_mesh_has_no_folds():
    for mx, my in ((3, 3), (7, 5), (32, 32)):
        assert fold_count(regular(mx, my), mx, my) == 0


def test_centre_vertex_past_its_right_neighbour_inverts_two_triangles():
    # 3x3 mesh, spacing 1: centre (1,1) moved to (2.5,1).  By hand,
    # (b-a)x(c-a) of its six triangles: {3,1,4} 2.5, {1,2,4} 1, {4,2,5} -0.5,
    # {3,4,6} 2.5, {6,4,7} 1, {4,5,7} -0.5 -> exactly two inverted
    xy = regular(3, 3)
    xy[4] = (2.5, 1.0)
    assert fold_count(xy, 3, 3) == 2


def test_collapsed_triangle_counts_as_folded():
    # centre moved onto the right neighbour: {4,2,5} and {4,5,7} have zero area
    xy = regular(3, 3)
    xy[4] = xy[5]
    assert fold_count(xy, 3, 3) == 2


def test_mirrored_mesh_inverts_every_triangle():
    mx = my = 6
    xy = regular(mx, my)
    xy[:, 0] = 2.0 - xy[:, 0]
    assert fold_count(xy, mx, my) == 2 * (mx - 1) * (my - 1)


def test_non_finite_vertex_counts_as_folded():
    xy = regular(3, 3)
    xy[0] = (np.nan, 0.0)
    assert fold_count(xy, 3, 3) == 1  # only triangle {0,1,3} uses vertex 0
