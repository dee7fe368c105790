"""ctypes bindings for the checkers under oracle/ (test infrastructure only).

``Oracle`` wraps liboracle.so, the FP64 C restatement of the reference hot path
(oracle.c).  ``RefLib`` wraps _ref/libcdref.so, the reference's own sources
compiled where they lie (built by oracle/Makefile when the reference tree is
present; the built library travels with the repository snapshot).

Only tests, the smoke check and the CPU-baseline leg of bench.py use this
module; the product library never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libcdref.so")
# the reference's unmodified main.cpp built against include/ and libcaustic_amd.so
REF_MAIN_DROPIN = os.path.join(HERE, "_ref", "ref_main_dropin")
REFERENCE_TREE = "/root/reference"

_D = C.c_double
_I = C.c_int
_P = C.c_void_p


def build(with_reference: bool = True) -> None:
    """Compile liboracle.so (and _ref/libcdref.so when the reference tree exists)."""
    targets = ["oracle"]
    if with_reference and os.path.isdir(os.path.join(REFERENCE_TREE, "src")):
        targets.append("ref")
        if os.path.exists(os.path.join(os.path.dirname(HERE), "poisson_caustic_design_amd",
                                       "libcaustic_amd.so")):
            targets.append("dropin")
    subprocess.run(["make", "-s", "-C", HERE, *targets], check=True)


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


class SorOpts(C.Structure):
    _fields_ = [
        ("max_iterations", _I),
        ("max_update_tol", _D),
        ("rel_residual_tol", _D),
        ("check_every", _I),
        ("ordering", _I),
        ("omega", _D),
    ]


class SorStats(C.Structure):
    _fields_ = [("sweeps", _I), ("last_max_update", _D), ("rel_residual", _D)]


LEXICOGRAPHIC, RED_BLACK = 0, 1
STOP_REFERENCE, STOP_RESIDUAL = 0, 1


class Oracle:
    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            build(with_reference=False)
        L = C.CDLL(path)
        self.L = L
        L.orc_poisson_sor.argtypes = [_P, _P, _I, _I, C.POINTER(SorOpts), C.POINTER(SorStats)]
        L.orc_residual.argtypes = [_P, _P, _I, _I, _P, C.POINTER(_D)]
        L.orc_residual.restype = _D
        L.orc_subtract_average.argtypes = [_P, _I, _I]
        L.orc_gradient.argtypes = [_P, _I, _I, _P, _P]
        L.orc_divergence.argtypes = [_P, _P, _I, _I, _P]
        L.orc_scale_minmax.argtypes = [_P, _I, _I, _D, _D, _P]
        L.orc_bilinear.argtypes = [_P, _I, _I, _D, _D]
        L.orc_bilinear.restype = _D
        L.orc_mesh_points.argtypes = [_I, _I, _D, _D, _P]
        L.orc_mesh_triangles.argtypes = [_I, _I, _P]
        L.orc_cell_areas.argtypes = [_I, _I, _P, _P]
        L.orc_target_areas.argtypes = [_P, _I, _I, _D, _D, _I, _I, _P]
        L.orc_integrate_polygon.argtypes = [_P, _I, _I, _D, _P, _I]
        L.orc_integrate_polygon.restype = _D
        L.orc_clip_area.argtypes = [_P, _I, _P, _I]
        L.orc_clip_area.restype = _D
        L.orc_raster.argtypes = [_I, _I, _P, _P, _I, _I, _D, _D, _P]
        L.orc_raster.restype = _I
        L.orc_vertex_step.argtypes = [_P, _I, _I, _D, _D, _I, _I, _P]
        L.orc_vertex_step.restype = _D
        L.orc_laplacian_smoothing.argtypes = [_P, _I, _I, _D]
        L.orc_inverse_map.argtypes = [_I, _I, _D, _D, _P, _P]
        L.orc_inverse_map.restype = _I
        L.orc_normals.argtypes = [_I, _I, _D, _D, _P, _P, _D, _D, _P, _P]
        L.orc_set_heights.argtypes = [_P, _I, _I, _D, _D, _I, _I, _P]
        L.orc_set_heights.restype = _D
        L.orc_design_create.argtypes = [_I, _I, _I, _I, _D, _D, _D]
        L.orc_design_create.restype = _P
        L.orc_design_destroy.argtypes = [_P]
        L.orc_design_set_solver.argtypes = [_P, _I, _D, _D, _I, _I, _I]
        L.orc_design_init.argtypes = [_P, _P]
        L.orc_design_transport_step.argtypes = [_P, C.POINTER(_D)]
        L.orc_design_height_step.argtypes = [_P, C.POINTER(_D)]
        L.orc_design_last_sweeps.argtypes = [_P]
        L.orc_design_last_residual.argtypes = [_P]
        L.orc_design_last_residual.restype = _D
        L.orc_design_grid.argtypes = [_P, _I]
        L.orc_design_grid.restype = C.POINTER(_D)
        L.orc_design_vec.argtypes = [_P, _I]
        L.orc_design_vec.restype = C.POINTER(_D)

    # ---- Poisson -----------------------------------------------------------
    def poisson_sor(self, D, phi, max_iterations=100000, max_update_tol=0.0,
                    rel_residual_tol=0.0, check_every=1, ordering=LEXICOGRAPHIC, omega=0.0):
        D = _f64(D)
        phi = _f64(phi).copy()
        h, w = D.shape
        o = SorOpts(max_iterations, max_update_tol, rel_residual_tol, check_every, ordering, omega)
        st = SorStats()
        self.L.orc_poisson_sor(_ptr(D), _ptr(phi), w, h, C.byref(o), C.byref(st))
        return phi, st

    def residual(self, D, phi):
        D, phi = _f64(D), _f64(phi)
        h, w = D.shape
        r = np.empty_like(D)
        bn = _D()
        rn = self.L.orc_residual(_ptr(D), _ptr(phi), w, h, _ptr(r), C.byref(bn))
        return r, rn, bn.value

    # ---- grid utils --------------------------------------------------------
    def subtract_average(self, a):
        a = _f64(a).copy()
        self.L.orc_subtract_average(_ptr(a), a.shape[1], a.shape[0])
        return a

    def gradient(self, a):
        a = _f64(a)
        gx, gy = np.empty_like(a), np.empty_like(a)
        self.L.orc_gradient(_ptr(a), a.shape[1], a.shape[0], _ptr(gx), _ptr(gy))
        return gx, gy

    def divergence(self, nx, ny):
        nx, ny = _f64(nx), _f64(ny)
        out = np.empty_like(nx)
        self.L.orc_divergence(_ptr(nx), _ptr(ny), nx.shape[1], nx.shape[0], _ptr(out))
        return out

    def scale_minmax(self, a, lo=0.0, hi=1.0):
        a = _f64(a)
        out = np.empty_like(a)
        self.L.orc_scale_minmax(_ptr(a), a.shape[1], a.shape[0], lo, hi, _ptr(out))
        return out

    def bilinear(self, img, x, y):
        img = _f64(img)
        return self.L.orc_bilinear(_ptr(img), img.shape[1], img.shape[0], x, y)

    # ---- mesh --------------------------------------------------------------
    def mesh_points(self, mx, my, W, H):
        xy = np.empty((my * mx, 2))
        self.L.orc_mesh_points(mx, my, W, H, _ptr(xy))
        return xy

    def mesh_triangles(self, mx, my):
        t = np.empty((2 * (mx - 1) * (my - 1), 3), dtype=np.int32)
        self.L.orc_mesh_triangles(mx, my, _ptr(t))
        return t

    def cell_areas(self, mx, my, xy):
        xy = _f64(xy)
        out = np.empty(mx * my)
        self.L.orc_cell_areas(mx, my, _ptr(xy), _ptr(out))
        return out

    def target_areas(self, image, W, H, mx, my):
        image = _f64(image)
        out = np.empty(mx * my)
        self.L.orc_target_areas(_ptr(image), image.shape[1], image.shape[0], W, H, mx, my, _ptr(out))
        return out

    def integrate_polygon(self, image, W, poly):
        image, poly = _f64(image), _f64(poly)
        return self.L.orc_integrate_polygon(_ptr(image), image.shape[1], image.shape[0], W,
                                            _ptr(poly), poly.shape[0])

    def clip_area(self, sub, clip):
        sub, clip = _f64(sub), _f64(clip)
        return self.L.orc_clip_area(_ptr(sub), sub.shape[0], _ptr(clip), clip.shape[0])

    def raster(self, mx, my, xy, values, nx, ny, W, H):
        xy, values = _f64(xy), _f64(values)
        out = np.empty((ny, nx))
        miss = self.L.orc_raster(mx, my, _ptr(xy), _ptr(values), nx, ny, W, H, _ptr(out))
        return out, miss

    def vertex_step(self, phi, W, H, mx, my, xy):
        phi = _f64(phi)
        xy = _f64(xy).copy()
        s = self.L.orc_vertex_step(_ptr(phi), phi.shape[1], phi.shape[0], W, H, mx, my, _ptr(xy))
        return xy, s

    def laplacian_smoothing(self, xy, mx, my, factor):
        xy = _f64(xy).copy()
        self.L.orc_laplacian_smoothing(_ptr(xy), mx, my, factor)
        return xy

    def inverse_map(self, mx, my, W, H, target_xy):
        target_xy = _f64(target_xy)
        inv = np.empty((mx * my, 2))
        miss = self.L.orc_inverse_map(mx, my, W, H, _ptr(target_xy), _ptr(inv))
        return inv, miss

    def normals(self, mx, my, W, H, inv_xy, src_z, focal, ior=1.49):
        inv_xy, src_z = _f64(inv_xy), _f64(src_z)
        nx, ny = np.empty(mx * my), np.empty(mx * my)
        self.L.orc_normals(mx, my, W, H, _ptr(inv_xy), _ptr(src_z), focal, ior, _ptr(nx), _ptr(ny))
        return nx, ny

    def set_heights(self, hgrid, W, H, mx, my, z):
        hgrid = _f64(hgrid)
        z = _f64(z).copy()
        u = self.L.orc_set_heights(_ptr(hgrid), hgrid.shape[1], hgrid.shape[0], W, H, mx, my, _ptr(z))
        return z, u


class OracleDesign:
    """FP64 design pipeline of the oracle (caustic_design.cpp orchestration)."""

    def __init__(self, orc: Oracle, mx, my, nx, ny, W, H, focal_l):
        self.o = orc
        self.L = orc.L
        self.mx, self.my, self.nx, self.ny = mx, my, nx, ny
        self.h = self.L.orc_design_create(mx, my, nx, ny, W, H, focal_l)

    def __del__(self):
        try:
            self.L.orc_design_destroy(self.h)
        except Exception:
            pass

    def set_solver(self, stop=STOP_REFERENCE, phi_tol=1e-7, h_tol=1e-8, max_sweeps=100000,
                   ordering=LEXICOGRAPHIC, check_every=1):
        self.L.orc_design_set_solver(self.h, stop, phi_tol, h_tol, max_sweeps, ordering, check_every)

    def init(self, image):
        image = _f64(image)
        assert image.shape == (self.ny, self.nx)
        self.L.orc_design_init(self.h, _ptr(image))

    def transport_step(self):
        s = _D()
        rc = self.L.orc_design_transport_step(self.h, C.byref(s))
        return rc, s.value

    def height_step(self):
        u = _D()
        rc = self.L.orc_design_height_step(self.h, C.byref(u))
        return rc, u.value

    def last_sweeps(self):
        return self.L.orc_design_last_sweeps(self.h)

    def last_residual(self):
        return self.L.orc_design_last_residual(self.h)

    def grid(self, which):
        p = self.L.orc_design_grid(self.h, which)
        return np.ctypeslib.as_array(p, shape=(self.ny, self.nx)).copy()

    def vec(self, which):
        V = self.mx * self.my
        n = 2 * V if which in (0, 6) else V
        p = self.L.orc_design_vec(self.h, which)
        a = np.ctypeslib.as_array(p, shape=(n,)).copy()
        return a.reshape(V, 2) if which in (0, 6) else a


class RefLib:
    """The reference's own code (src/*.cpp) compiled into _ref/libcdref.so."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = C.CDLL(path)
        self.L = L
        L.ref_poisson_solver.argtypes = [_P, _P, _I, _I, _I, _D, _I]
        L.ref_subtract_average.argtypes = [_P, _I, _I]
        L.ref_gradient.argtypes = [_P, _I, _I, _P, _P]
        L.ref_divergence.argtypes = [_P, _P, _I, _I, _P]
        L.ref_scale_minmax.argtypes = [_P, _I, _I, _D, _D, _P]
        L.ref_bilinear.argtypes = [_P, _I, _I, _D, _D]
        L.ref_bilinear.restype = _D
        L.ref_integrate_polygon.argtypes = [_P, _I, _I, _D, _P, _I]
        L.ref_integrate_polygon.restype = _D
        L.ref_design_create.argtypes = [_I, _I, _I, _I, _D, _D, _D, _D, _I]
        L.ref_design_create.restype = _P
        L.ref_design_destroy.argtypes = [_P]
        L.ref_design_set_threads.argtypes = [_P, _I]
        L.ref_design_init.argtypes = [_P, _P]
        L.ref_get_target_cells.argtypes = [_P, _I, _P]
        L.ref_get_target_cells.restype = C.c_long
        L.ref_vertex_normal.argtypes = [_P, _P, _I, _P]
        L.ref_transport_step.argtypes = [_P]
        L.ref_transport_step.restype = _D
        L.ref_height_step.argtypes = [_P]
        L.ref_get_target_points.argtypes = [_P, _P]
        L.ref_get_source_points.argtypes = [_P, _P]
        L.ref_set_target_points.argtypes = [_P, _P]
        L.ref_set_source_z.argtypes = [_P, _P]
        L.ref_get_grid.argtypes = [_P, _I, _P]
        L.ref_bvh_query.argtypes = [_P, _I, _I, _P, _I, _P]
        L.ref_set_grid.argtypes = [_P, _I, _P]
        L.ref_get_vector.argtypes = [_P, _I, _P]
        L.ref_inverted_map.argtypes = [_P, _P]
        L.ref_laplacian_smoothing.argtypes = [_P, _D]
        L.ref_save_solid_obj.argtypes = [C.c_char_p, _P, _P, _I, _I, _D, _D, _D]
        L.ref_export_grid_svg.argtypes = [C.c_char_p, _P, _I, _I, _I, _D, _D, _D]
        L.ref_export_triangles_svg.argtypes = [C.c_char_p, _P, _I, _I, _I, _D, _D, _D]

    def bvh_query(self, xy, mx, my, pts):
        """Triangle the reference's BVH returns for each point (-1: none)."""
        xy, pts = _f64(np.asarray(xy)[:, :2]), _f64(pts)
        face = np.empty(len(pts), dtype=np.int32)
        self.L.ref_bvh_query(_ptr(xy), mx, my, _ptr(pts), len(pts), face.ctypes.data)
        return face

    @staticmethod
    def available(path: str = REF_SO) -> bool:
        return os.path.exists(path)

    def save_solid_obj(self, path, front, back, res_x, res_y, thickness, width, height):
        f, b = _f64(front), _f64(back)
        self.L.ref_save_solid_obj(path.encode(), f.ctypes.data, b.ctypes.data, res_x, res_y,
                                  thickness, width, height)

    def export_svg(self, path, pts, res_x, res_y, width, height, stroke, triangles=False):
        p = _f64(pts)
        fn = self.L.ref_export_triangles_svg if triangles else self.L.ref_export_grid_svg
        fn(path.encode(), p.ctypes.data, p.shape[1], res_x, res_y, width, height, stroke)

    def poisson_solver(self, D, phi, max_iterations, threshold, threads=1):
        D = _f64(D)
        phi = _f64(phi).copy()
        h, w = D.shape
        self.L.ref_poisson_solver(_ptr(D), _ptr(phi), w, h, max_iterations, threshold, threads)
        return phi

    def subtract_average(self, a):
        a = _f64(a).copy()
        self.L.ref_subtract_average(_ptr(a), a.shape[1], a.shape[0])
        return a

    def gradient(self, a):
        a = _f64(a)
        gx, gy = np.empty_like(a), np.empty_like(a)
        self.L.ref_gradient(_ptr(a), a.shape[1], a.shape[0], _ptr(gx), _ptr(gy))
        return gx, gy

    def divergence(self, nx, ny):
        nx, ny = _f64(nx), _f64(ny)
        out = np.empty_like(nx)
        self.L.ref_divergence(_ptr(nx), _ptr(ny), nx.shape[1], nx.shape[0], _ptr(out))
        return out

    def scale_minmax(self, a, lo=0.0, hi=1.0):
        a = _f64(a)
        out = np.empty_like(a)
        self.L.ref_scale_minmax(_ptr(a), a.shape[1], a.shape[0], lo, hi, _ptr(out))
        return out

    def bilinear(self, img, x, y):
        img = _f64(img)
        return self.L.ref_bilinear(_ptr(img), img.shape[1], img.shape[0], x, y)

    def integrate_polygon(self, image, W, poly):
        image, poly = _f64(image), _f64(poly)
        return self.L.ref_integrate_polygon(_ptr(image), image.shape[1], image.shape[0], W,
                                            _ptr(poly), poly.shape[0])


class RefDesign:
    """One reference Caustic_design instance (caustic_design.h:7-66)."""

    def __init__(self, ref: RefLib, mx, my, nx, ny, W, H, focal_l, thickness=0.1, threads=1):
        self.L = ref.L
        self.mx, self.my, self.nx, self.ny = mx, my, nx, ny
        self.h = self.L.ref_design_create(mx, my, nx, ny, W, H, focal_l, thickness, threads)

    def __del__(self):
        try:
            self.L.ref_design_destroy(self.h)
        except Exception:
            pass

    def init(self, image):
        image = _f64(image)
        assert image.shape == (self.ny, self.nx)
        self.L.ref_design_init(self.h, _ptr(image))

    def set_threads(self, n):
        self.L.ref_design_set_threads(self.h, n)

    def transport_step(self):
        return self.L.ref_transport_step(self.h)

    def height_step(self):
        self.L.ref_height_step(self.h)

    def target_points(self):
        a = np.empty((self.mx * self.my, 3))
        self.L.ref_get_target_points(self.h, _ptr(a))
        return a

    def source_points(self):
        a = np.empty((self.mx * self.my, 3))
        self.L.ref_get_source_points(self.h, _ptr(a))
        return a

    def set_target_points(self, xyz):
        xyz = _f64(xyz)
        self.L.ref_set_target_points(self.h, _ptr(xyz))

    def set_source_z(self, z):
        z = _f64(z)
        self.L.ref_set_source_z(self.h, _ptr(z))

    def laplacian_smoothing(self, factor):
        self.L.ref_laplacian_smoothing(self.h, factor)

    def target_cells(self, initial: bool) -> np.ndarray:
        """first_target_cells (initial) or target_cells, flattened: per vertex
        its quad count, then the quads' 4 x {x, y, z}, in the reference's order."""
        buf = np.empty(self.mx * self.my * (1 + 6 * 12))
        n = self.L.ref_get_target_cells(self.h, int(initial), _ptr(buf))
        return buf[:n].copy()

    def vertex_normal(self, xyz, vertex: int) -> np.ndarray:
        xyz = _f64(xyz)
        out = np.empty(3)
        self.L.ref_vertex_normal(self.h, _ptr(xyz), vertex, _ptr(out))
        return out

    def grid(self, which):
        a = np.empty((self.ny, self.nx))
        rc = self.L.ref_get_grid(self.h, which, _ptr(a))
        if rc != 0:
            raise ValueError(f"grid {which} unavailable ({rc})")
        return a

    def set_grid(self, which, a):
        a = _f64(a)
        self.L.ref_set_grid(self.h, which, _ptr(a))

    def vector(self, which):
        a = np.empty(self.mx * self.my)
        n = self.L.ref_get_vector(self.h, which, _ptr(a))
        if n < 0:
            raise ValueError(f"vector {which} unavailable ({n})")
        return a[:n]

    def inverted_map(self):
        a = np.empty((self.mx * self.my, 2))
        n = self.L.ref_inverted_map(self.h, _ptr(a))
        return a[:n], n
