"""Pins the oracle's grid utilities and mesh geometry (oracle.c) against the
reference build and closed forms."""
import numpy as np
import pytest

from oracle.bind import OracleDesign, RefDesign
from tests.helpers import small_image


def test_grid_utils_vs_reference(oracle, reflib, rng):
    A = rng.standard_normal((20, 30))
    B = rng.standard_normal((20, 30))
    for x, y in zip(reflib.gradient(A), oracle.gradient(A)):
        assert np.array_equal(x, y)
    assert np.array_equal(reflib.divergence(A, B), oracle.divergence(A, B))
    assert np.array_equal(reflib.subtract_average(A), oracle.subtract_average(A))
    assert np.array_equal(reflib.scale_minmax(A), oracle.scale_minmax(A))
    An = A.copy(); An[3, 4] = np.nan
    assert np.array_equal(reflib.subtract_average(An), oracle.subtract_average(An), equal_nan=True)
    assert np.array_equal(reflib.scale_minmax(An), oracle.scale_minmax(An))
    for _ in range(50):
        x, y = rng.uniform(-3, 33), rng.uniform(-3, 23)
        assert reflib.bilinear(A, x, y) == oracle.bilinear(A, x, y)


def test_gradient_linear_field(oracle):
    h, w = 9, 13
    yy, xx = np.mgrid[0:h, 0:w].astype(float)
    gx, gy = oracle.gradient(3.0 * xx - 2.0 * yy)
    assert np.allclose(gx[:, 1:-1], 3.0) and np.allclose(gx[:, [0, -1]], 1.5)
    assert np.allclose(gy[1:-1, :], -2.0) and np.allclose(gy[[0, -1], :], -1.0)
    d = oracle.divergence(2.0 * xx, 5.0 * yy)
    assert np.allclose(d[1:-1, 1:-1], 7.0)
    assert (d[0] == 0).all() and (d[-1] == 0).all() and (d[:, 0] == 0).all() and (d[:, -1] == 0).all()


def test_bilinear_closed_form(oracle):
    h, w = 7, 11
    yy, xx = np.mgrid[0:h, 0:w].astype(float)
    f = 1.0 + 2.0 * xx + 3.0 * yy + 0.5 * xx * yy
    for x, y in [(2.25, 3.5), (0.0, 0.0), (9.75, 5.125)]:
        assert abs(oracle.bilinear(f, x, y) - (1 + 2 * x + 3 * y + 0.5 * x * y)) < 1e-12
    # clamp-to-edge outside the grid
    assert oracle.bilinear(f, -0.4, 2.0) == pytest.approx(f[2, 0])
    assert oracle.bilinear(f, 10.4, 6.3) == pytest.approx(f[6, 10])


def test_clip_area_rectangles(oracle, rng):
    for _ in range(40):
        a0, b0 = np.sort(rng.uniform(0, 1, 2)), np.sort(rng.uniform(0, 1, 2))
        c0, d0 = np.sort(rng.uniform(0, 1, 2)), np.sort(rng.uniform(0, 1, 2))
        sub = np.array([[a0[0], b0[0]], [a0[1], b0[0]], [a0[1], b0[1]], [a0[0], b0[1]]])  # CCW
        clip = np.array([[c0[0], d0[0]], [c0[0], d0[1]], [c0[1], d0[1]], [c0[1], d0[0]]])  # CW
        ox = max(0.0, min(a0[1], c0[1]) - max(a0[0], c0[0]))
        oy = max(0.0, min(b0[1], d0[1]) - max(b0[0], d0[0]))
        assert oracle.clip_area(sub, clip) == pytest.approx(ox * oy, abs=1e-15)


def test_integrate_polygon_vs_reference(oracle, reflib, rng):
    img = rng.uniform(0, 1, (40, 40))
    for _ in range(30):
        poly = rng.uniform(-0.05, 0.55, (4, 2))
        c = poly.mean(0)
        poly = poly[np.argsort(np.arctan2(poly[:, 1] - c[1], poly[:, 0] - c[0]))]
        assert reflib.integrate_polygon(img, 0.5, poly) == oracle.integrate_polygon(img, 0.5, poly)
    # uniform image: the integral equals the in-domain polygon area
    ones = np.ones((40, 40))
    tri = np.array([[0.1, 0.1], [0.4, 0.12], [0.2, 0.45]])
    u, v = tri[1] - tri[0], tri[2] - tri[0]
    area = 0.5 * abs(u[0] * v[1] - u[1] * v[0])
    assert oracle.integrate_polygon(ones, 0.5, tri) == pytest.approx(area, rel=1e-13)


def test_cell_areas_partition(oracle, rng):
    mx, my, W, H = 9, 7, 0.5, 0.4
    xy = oracle.mesh_points(mx, my, W, H)
    a = oracle.cell_areas(mx, my, xy)
    assert a.sum() == pytest.approx(W * H, rel=1e-14)
    # jitter interior vertices: areas still partition the domain and each equals
    # one third of the adjacent triangle areas
    j = xy.copy()
    inner = [(i * mx + k) for i in range(1, my - 1) for k in range(1, mx - 1)]
    j[inner] += rng.uniform(-0.2, 0.2, (len(inner), 2)) * np.array([W / (mx - 1), H / (my - 1)])
    a = oracle.cell_areas(mx, my, j)
    assert a.sum() == pytest.approx(W * H, rel=1e-13)
    tri = oracle.mesh_triangles(mx, my)
    p = j[tri]
    ta = 0.5 * ((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1])
                - (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))
    third = np.zeros(mx * my)
    for k in range(3):
        np.add.at(third, tri[:, k], ta / 3)
    assert np.allclose(a, third, rtol=1e-12, atol=1e-16)


def test_target_areas_uniform_image(oracle):
    mx, my, N, W = 6, 6, 24, 0.5
    t = oracle.target_areas(np.ones((N, N)), W, W, mx, my)
    cells = oracle.cell_areas(mx, my, oracle.mesh_points(mx, my, W, W))
    assert np.allclose(t, cells * (W * W) / cells.sum(), rtol=1e-13)


def test_raster_linear_field_exact(oracle, rng):
    mx, my, W, H, N = 8, 6, 0.5, 0.5, 30
    xy = oracle.mesh_points(mx, my, W, H)
    inner = [(i * mx + k) for i in range(1, my - 1) for k in range(1, mx - 1)]
    xy[inner] += rng.uniform(-0.15, 0.15, (len(inner), 2)) * np.array([W / (mx - 1), H / (my - 1)])
    vals = 2.0 * xy[:, 0] - 3.0 * xy[:, 1] + 0.25
    r, miss = oracle.raster(mx, my, xy, vals, N, N, W, H)
    assert miss == 0
    eps = 1e-8
    xs = np.arange(N) * (W - eps) / (N - 1) + 0.5 * eps
    ys = np.arange(N) * (H - eps) / (N - 1) + 0.5 * eps
    assert np.allclose(r, 2.0 * xs[None, :] - 3.0 * ys[:, None] + 0.25, atol=1e-12)


def test_inverse_map_identity_and_normals(oracle):
    mx, my, W, H = 7, 5, 0.5, 0.35
    xy = oracle.mesh_points(mx, my, W, H)
    inv, miss = oracle.inverse_map(mx, my, W, H, xy)
    assert miss == 0
    assert np.allclose(inv, xy, atol=1e-12)
    nx, ny = oracle.normals(mx, my, W, H, inv, np.zeros(mx * my), 1200.0)
    # the epsilon inset (mesh.cpp:356-360) displaces the identity map by ~1e-8
    assert np.abs(nx).max() < 1e-10 and np.abs(ny).max() < 1e-10
    t = np.stack([inv[:, 0] - xy[:, 0], inv[:, 1] - xy[:, 1], np.full(mx * my, 1200.0)], 1)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    assert np.allclose(nx, t[:, 0] / (t[:, 2] - 1.49), rtol=1e-9, atol=1e-24)
    # closed form for a displaced map with zero heights
    d = np.zeros_like(inv); d[:, 0] = 0.01
    nx, ny = oracle.normals(mx, my, W, H, xy + d, np.zeros(mx * my), 2.0)
    t = np.array([0.01, 0.0, 2.0]); t /= np.linalg.norm(t)
    assert np.allclose(nx, t[0] / (t[2] - 1.49), rtol=1e-13)


def test_set_heights_min_shift(oracle):
    mx, my, N, W = 5, 5, 20, 0.5
    hg = np.full((N, N), -0.3)
    z, upd = oracle.set_heights(hg, W, W, mx, my, np.zeros(mx * my))
    assert np.allclose(z, 0.0) and upd < 1e-30
    hg = np.full((N, N), 0.7)
    z, upd = oracle.set_heights(hg, W, W, mx, my, np.full(mx * my, 0.5))
    assert np.allclose(z, 0.7) and upd == pytest.approx(mx * my * 0.04)


@pytest.mark.parametrize("M,steps", [(8, 4), (12, 3)])
def test_design_trajectory_vs_reference(oracle, reflib, M, steps):
    N = 4 * M
    W = 0.5
    img = small_image(3, N)
    rd = RefDesign(reflib, M, M, N, N, W, W, 1.5)
    od = OracleDesign(oracle, M, M, N, N, W, W, 1.5)
    rd.init(img)
    od.init(img)
    assert np.array_equal(rd.grid(6), od.grid(6))
    assert np.allclose(rd.vector(1), od.vec(2), rtol=1e-13, atol=0)
    for _ in range(steps):
        s1 = rd.transport_step()
        rc, s2 = od.transport_step()
        assert rc == 0
        assert s1 == pytest.approx(s2, rel=1e-9)
        assert np.abs(rd.target_points()[:, :2] - od.vec(0)).max() < 1e-13
        assert np.abs(rd.grid(0) - od.grid(0)).max() < 1e-9 * max(1.0, np.abs(od.grid(0)).max())
    for _ in range(3):
        rd.height_step()
        rc, _ = od.height_step()
        assert rc == 0
        assert np.abs(rd.source_points()[:, 2] - od.vec(1)).max() < 1e-12
        assert np.abs(rd.grid(1) - od.grid(1)).max() < 1e-12


def test_laplacian_smoothing_matches_reference(oracle, reflib):
    """orc_laplacian_smoothing == Mesh::laplacian_smoothing (mesh.cpp:534-606)
    bit for bit on a perturbed mesh, and its closed forms: a regular grid's
    interior is a fixed point of the averaging; corners never move."""
    from oracle.bind import RefDesign
    mx, my = 9, 7
    W, H = 1.0, 0.75
    rd = RefDesign(reflib, mx, my, 4 * mx, 4 * my, W, H, 1.5)
    rd.init(np.ones((4 * my, 4 * mx)))  # builds the reference's Mesh
    rng = np.random.default_rng(11)
    pts = rd.target_points()
    pts[:, :2] += 0.01 * rng.standard_normal((mx * my, 2))
    rd.set_target_points(pts)
    f = float(np.float32(0.1))
    rd.laplacian_smoothing(f)
    ref = rd.target_points()[:, :2]
    ours = oracle.laplacian_smoothing(pts[:, :2], mx, my, f)
    assert np.array_equal(ours, ref)
    # closed forms on the regular grid
    xs, ys = np.meshgrid(np.linspace(0, W, mx), np.linspace(0, H, my))
    reg = np.stack([xs.ravel(), ys.ravel()], 1)
    sm = oracle.laplacian_smoothing(reg, mx, my, 0.5)
    assert np.allclose(sm, reg, atol=1e-15)
    pert = pts[:, :2].copy()
    one = oracle.laplacian_smoothing(pert, mx, my, 1.0)
    for c in (0, mx - 1, (my - 1) * mx, mx * my - 1):
        assert np.array_equal(one[c], pert[c])
    # interior vertex: exact 4-neighbour mean at f = 1
    i = 3 * mx + 4
    nb = pert[[i - 1, i + 1, i - mx, i + mx]]
    assert np.allclose(one[i], nb.mean(0), rtol=0, atol=1e-15)
