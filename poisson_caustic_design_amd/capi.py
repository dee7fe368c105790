"""ctypes binding of libcaustic_amd.so (the C ABI declared in include/cd_api.h).

Python mirror of the reference's C++ interfaces for this path:

* ``poisson_solver(D, phi, width, height, max_iterations, threshold, max_threads)``
  mirrors ``poisson_solver`` (src/solver.h:8): D and phi are [height][width]
  arrays, phi is updated in place, the stop rule is the reference's
  max-update threshold.
* ``CausticDesign`` mirrors ``class Caustic_design`` (src/caustic_design.h:7-66):
  the same setters, ``initialize_solvers``, ``perform_transport_iteration`` and
  ``perform_height_map_iteration``, and the same public state as properties.

There is no CPU fallback: loading fails loudly when the library is missing,
and every call fails with ``CDError`` when the HIP device is unavailable.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref
from typing import Optional

import numpy as np

LIB_PATH = os.environ.get(
    "CD_LIB_PATH",
    os.path.join(os.path.dirname(os.path.abspath(__file__)), "libcaustic_amd.so"))

_D, _I, _P, _SZ, _LL = C.c_double, C.c_int, C.c_void_p, C.c_size_t, C.c_longlong

OK, RASTER_MISS, FOLDED, NOT_CONVERGED, NAN_, BAD_ARG, HIP_ERROR, NO_DEVICE, OOM, STATE = range(10)
PREC_MIXED, PREC_FP64 = 0, 1
STOP_RESIDUAL, STOP_MAX_UPDATE = 0, 1
KERNEL_SOR_SWEEP, KERNEL_RESIDUAL, KERNEL_RASTER, KERNEL_VERTEX, KERNEL_LEARNED = 0, 1, 2, 3, 4

FIELD = {
    "target_points": 0, "source_points": 1, "phi": 2, "h": 3, "raster": 4, "errors": 5,
    "target_areas": 6, "divergence": 7, "norm_x": 8, "norm_y": 9, "normals": 10,
    "vertex_gradient": 11, "pixels": 12, "inverse_map": 13, "image": 14,
}


class SolveOpts(C.Structure):
    _fields_ = [("max_sweeps", _I), ("stop", _I), ("tol", _D), ("precision", _I),
                ("check_every", _I), ("learned_init", _I), ("omega", _D)]


class SolveStats(C.Structure):
    _fields_ = [("status", _I), ("sweeps", _I), ("checks", _I), ("learned_applied", _I),
                ("learned_rejected", _I), ("rel_residual", _D), ("initial_rel_residual", _D),
                ("last_max_update", _D), ("elapsed_ms", _D)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class DesignParams(C.Structure):
    _fields_ = [("mesh_res_x", _I), ("mesh_res_y", _I), ("res_x", _I), ("res_y", _I),
                ("width", _D), ("height", _D), ("focal_l", _D), ("thickness", _D),
                ("ior", _D), ("step_size", _D), ("phi_solve", SolveOpts), ("h_solve", SolveOpts)]


class RunStats(C.Structure):
    _fields_ = [("status", _I), ("transport_steps", _I), ("height_steps", _I),
                ("last_step", _D), ("phi_sweeps", _LL), ("h_sweeps", _LL), ("checks", _LL),
                ("learned_applied", _I), ("learned_rejected", _I),
                ("max_phi_residual", _D), ("max_h_residual", _D), ("elapsed_ms", _D),
                ("recoveries", _I), ("folded_triangles", _LL), ("folded_steps", _I),
                ("first_folded_step", _I), ("unconverged_solves", _I)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class NetArch(C.Structure):
    _fields_ = [("levels", _I), ("channels", _I * 6), ("coarsen", _I), ("smooth_sweeps", _I),
                ("smooth_omega", _D), ("linear", _I), ("out_smooth", _I), ("loss_smooth", _I),
                ("loss_energy", _I), ("mid_sweeps", _I)]


class CDError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(msg)
        self.status = status


_lib = None


def load(path: str = LIB_PATH):
    """Load libcaustic_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} not built: run `python -m poisson_caustic_design_amd.build`")
    L = C.CDLL(path)
    sig = {
        "cd_get_api_version": ([], _I),
        "cd_status_string": ([_I], C.c_char_p),
        "cd_solve_opts_default": ([C.POINTER(SolveOpts)], None),
        "cd_context_create": ([_I, C.POINTER(_P)], _I),
        "cd_context_destroy": ([_P], None),
        "cd_context_synchronize": ([_P], _I),
        "cd_context_last_error": ([_P], C.c_char_p),
        "cd_context_set_profiling": ([_P, _I], _I),
        "cd_context_kernel_stats": ([_P, _I, C.POINTER(_D), C.POINTER(_LL), C.POINTER(_D)], _I),
        "cd_context_reset_kernel_stats": ([_P], _I),
        "cd_poisson_solve": ([_P, _P, _P, _I, _I, C.POINTER(SolveOpts), C.POINTER(SolveStats)], _I),
        "cd_poisson_solve_device": ([_P, _P, _P, _I, _I, C.POINTER(SolveOpts),
                                     C.POINTER(SolveStats)], _I),
        "cd_poisson_solve_device_noalloc": ([_P, _P, _P, _I, _I, C.POINTER(SolveOpts),
                                             C.POINTER(SolveStats)], _I),
        "cd_design_params_default": ([C.POINTER(DesignParams)], None),
        "cd_design_create": ([_P, C.POINTER(DesignParams), C.POINTER(_P)], _I),
        "cd_design_destroy": ([_P], None),
        "cd_design_set_solve_opts": ([_P, C.POINTER(SolveOpts), C.POINTER(SolveOpts)], _I),
        "cd_design_init": ([_P, _P], _I),
        "cd_design_init_device": ([_P, _P], _I),
        "cd_design_transport_step": ([_P, C.POINTER(_D)], _I),
        "cd_design_height_step": ([_P, C.POINTER(_D)], _I),
        "cd_design_run": ([_P, _I, _D, _I, C.POINTER(RunStats)], _I),
        "cd_designs_run_batch": ([_P, _P, _I, _I, _D, _I, C.POINTER(RunStats)], _I),
        "cd_net_set_fused_forward": ([_P, _I], _I),
        "cd_net_set_tiled_forward": ([_P, _I], _I),
        "cd_design_field_size": ([_P, _I], _SZ),
        "cd_design_get": ([_P, _I, _P, _SZ], _I),
        "cd_design_set": ([_P, _I, _P, _SZ], _I),
        "cd_design_device_ptr": ([_P, _I], _P),
        "cd_design_last_solve_stats": ([_P, C.POINTER(SolveStats), C.POINTER(SolveStats)], _I),
        "cd_design_get_params": ([_P, C.POINTER(DesignParams)], _I),
        "cd_context_set_graphs": ([_P, _I], _I),
        "cd_design_laplacian_smoothing": ([_P, _D], _I),
        "cd_design_set_miss_recovery": ([_P, _I], _I),
        "cd_design_set_fold_check": ([_P, _I], _I),
        "cd_design_count_folds": ([_P, C.POINTER(_LL)], _I),
        "cd_design_set_first_hit": ([_P, _I], _I),
        "cd_bvh_first_hit_order": ([_P, _I, _I, _P], _I),
        "cd_png_read_rgba8": ([C.c_char_p, C.POINTER(_I), C.POINTER(_I), _P, _SZ], _I),
        "cd_image_gray_resize": ([_P, _P, _I, _I, _I, _I, _P], _I),
        "cd_image_gray_resize_device": ([_P, _P, _I, _I, _I, _I, _P], _I),
        "cd_obj_write_solid": ([C.c_char_p, _P, _P, _I, _I, _D, _D, _D], _I),
        "cd_json_write_heightmap": ([C.c_char_p, _P, _I, _I], _I),
        "cd_svg_write_grid": ([C.c_char_p, _P, _I, _I, _I, _D, _D, _D], _I),
        "cd_svg_write_triangles": ([C.c_char_p, _P, _I, _I, _I, _D, _D, _D], _I),
        "cd_design_save_solid_obj": ([_P, C.c_char_p, _I], _I),
        "cd_design_save_heightmap_json": ([_P, C.c_char_p], _I),
        "cd_design_export_svg": ([_P, C.c_char_p, _I, _D], _I),
        "cd_net_arch_default": ([C.POINTER(NetArch)], None),
        "cd_net_create": ([_P, C.POINTER(NetArch), C.POINTER(_P)], _I),
        "cd_net_destroy": ([_P], None),
        "cd_net_num_params": ([_P], _SZ),
        "cd_net_init_params": ([_P, C.c_ulonglong], _I),
        "cd_net_set_params": ([_P, _P, _SZ], _I),
        "cd_net_get_params": ([_P, _P, _SZ], _I),
        "cd_net_forward": ([_P, _P, _P, _I, _I, _I], _I),
        "cd_net_loss_grad": ([_P, _P, _P, _I, _I, _I, C.POINTER(_D), _P, _SZ], _I),
        "cd_net_train_step": ([_P, _P, _P, _I, _I, _I, C.c_float, C.POINTER(_D)], _I),
        "cd_net_loss_grad_operator": ([_P, _P, _I, _I, _I, C.POINTER(_D), _P, _SZ], _I),
        "cd_net_train_step_operator": ([_P, _P, _I, _I, _I, C.c_float, C.POINTER(_D)], _I),
        "cd_context_set_net": ([_P, _P], _I),
        "cd_context_set_capture": ([_P, _P, _P, _I, _I, _I, _I, _I], _I),
        "cd_context_capture_count": ([_P, C.POINTER(_I)], _I),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def exported_symbols():
    return [
        "cd_get_api_version", "cd_status_string", "cd_solve_opts_default", "cd_context_create",
        "cd_context_destroy", "cd_context_synchronize", "cd_context_last_error",
        "cd_context_set_profiling", "cd_context_kernel_stats", "cd_context_reset_kernel_stats",
        "cd_poisson_solve", "cd_poisson_solve_device", "cd_poisson_solve_device_noalloc",
        "cd_design_params_default",
        "cd_design_create", "cd_design_destroy", "cd_design_set_solve_opts", "cd_design_init",
        "cd_design_init_device", "cd_design_transport_step", "cd_design_height_step",
        "cd_design_run", "cd_design_field_size", "cd_design_get", "cd_design_set",
        "cd_design_device_ptr", "cd_design_last_solve_stats", "cd_net_arch_default",
        "cd_net_create", "cd_net_destroy", "cd_net_num_params", "cd_net_init_params",
        "cd_net_set_params", "cd_net_get_params", "cd_net_forward", "cd_net_loss_grad",
        "cd_net_train_step", "cd_net_loss_grad_operator", "cd_net_train_step_operator",
        "cd_context_set_net", "cd_context_set_capture",
        "cd_context_capture_count", "cd_design_get_params", "cd_png_read_rgba8",
        "cd_context_set_graphs",
        "cd_image_gray_resize", "cd_image_gray_resize_device", "cd_obj_write_solid",
        "cd_json_write_heightmap", "cd_svg_write_grid", "cd_svg_write_triangles",
        "cd_design_save_solid_obj", "cd_design_save_heightmap_json", "cd_design_export_svg",
        "cd_design_laplacian_smoothing", "cd_design_set_miss_recovery",
        "cd_designs_run_batch", "cd_net_set_fused_forward", "cd_net_set_tiled_forward",
        "cd_design_set_fold_check",
        "cd_design_count_folds", "cd_design_set_first_hit", "cd_bvh_first_hit_order",
    ]


def status_string(s: int) -> str:
    return load().cd_status_string(s).decode()


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def default_opts(**kw) -> SolveOpts:
    o = SolveOpts()
    load().cd_solve_opts_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


_live_contexts: "weakref.WeakSet[Context]" = weakref.WeakSet()


@atexit.register
def _close_all():
    # before the HIP runtime's own exit handlers run: nets and designs first,
    # then their contexts
    for c in list(_live_contexts):
        c.close()


class Context:
    """One HIP device + stream (cd_context).  Nets and designs created on it
    are closed before it (close() or interpreter exit)."""

    def __init__(self, device: int = 0):
        L = load()
        h = _P()
        rc = L.cd_context_create(device, C.byref(h))
        if rc != OK:
            raise CDError(rc, f"cd_context_create: {status_string(rc)}")
        self.L, self.h = L, h
        self.children = weakref.WeakSet()
        _live_contexts.add(self)

    def close(self):
        if getattr(self, "h", None):
            for ch in list(getattr(self, "children", ())):
                ch.close()
            self.L.cd_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc: int, what: str, allow=(OK,)):
        if rc not in allow:
            msg = self.L.cd_context_last_error(self.h).decode()
            raise CDError(rc, f"{what}: {status_string(rc)}: {msg}")
        return rc

    def synchronize(self):
        self.check(self.L.cd_context_synchronize(self.h), "synchronize")

    def set_profiling(self, on: bool):
        self.check(self.L.cd_context_set_profiling(self.h, int(on)), "set_profiling")

    def reset_kernel_stats(self):
        self.check(self.L.cd_context_reset_kernel_stats(self.h), "reset_kernel_stats")

    def kernel_stats(self, kernel: int):
        ms, n, b = _D(), _LL(), _D()
        self.check(self.L.cd_context_kernel_stats(self.h, kernel, C.byref(ms), C.byref(n),
                                                  C.byref(b)), "kernel_stats")
        return ms.value, n.value, b.value

    def poisson_solve(self, rhs, phi, opts: Optional[SolveOpts] = None, allow=(OK, NOT_CONVERGED)):
        """Solve in place on host arrays ([height][width]); returns (phi, stats)."""
        rhs = _f64(rhs)
        phi = _f64(phi).copy()
        h, w = rhs.shape
        if phi.shape != rhs.shape:
            raise ValueError("phi and rhs shapes differ")
        st = SolveStats()
        rc = self.L.cd_poisson_solve(self.h, rhs.ctypes.data, phi.ctypes.data, w, h,
                                     C.byref(opts or default_opts()), C.byref(st))
        self.check(rc, "cd_poisson_solve", allow)
        return phi, st

    def set_graphs(self, on: bool):
        """Replay learned FCG iterations as one captured HIP graph (default on)."""
        self.check(self.L.cd_context_set_graphs(self.h, int(on)), "set_graphs")

    def set_net(self, net: Optional["Net"]):
        self.check(self.L.cd_context_set_net(self.h, net.h if net else None), "set_net")

    def set_capture(self, d_x: int, d_t: int, capacity: int, h: int, w: int, coarsen: int,
                    every: int = 1):
        self.check(self.L.cd_context_set_capture(self.h, d_x, d_t, capacity, h, w, coarsen,
                                                  every), "set_capture")

    def clear_capture(self):
        self.check(self.L.cd_context_set_capture(self.h, None, None, 0, 0, 0, 0, 0), "set_capture")

    def capture_count(self) -> int:
        n = _I()
        self.check(self.L.cd_context_capture_count(self.h, C.byref(n)), "capture_count")
        return n.value

    def poisson_solve_device(self, d_rhs: int, d_phi: int, width: int, height: int,
                             opts: Optional[SolveOpts] = None, allow=(OK, NOT_CONVERGED)):
        st = SolveStats()
        rc = self.L.cd_poisson_solve_device(self.h, d_rhs, d_phi, width, height,
                                            C.byref(opts or default_opts()), C.byref(st))
        self.check(rc, "cd_poisson_solve_device", allow)
        return st

    def poisson_solve_device_noalloc(self, d_rhs: int, d_phi: int, width: int, height: int,
                                     opts: Optional[SolveOpts] = None,
                                     allow=(OK, NOT_CONVERGED)):
        """cd_poisson_solve_device_noalloc: the same solve with the library's
        static per-device scratch instead of allocated buffers (grids up to
        2048 x 2048; overlapping calls on one GPU race)."""
        st = SolveStats()
        rc = self.L.cd_poisson_solve_device_noalloc(self.h, d_rhs, d_phi, width, height,
                                                    C.byref(opts or default_opts()), C.byref(st))
        self.check(rc, "cd_poisson_solve_device_noalloc", allow)
        return st


_shared_ctx = None


def poisson_solver(D, phi, width, height, max_iterations, convergence_threshold, max_threads):
    """Mirror of poisson_solver (src/solver.h:8): phi (numpy [height][width]) updated in place."""
    global _shared_ctx
    if max_threads <= 0 or max_iterations <= 0:  # zero tiles: no sweep (solver.cpp:73-83)
        return None
    if _shared_ctx is None:
        _shared_ctx = Context(int(os.environ.get("CD_DEVICE", "0")))
    D = _f64(D)[:height, :width]
    o = default_opts(max_sweeps=max_iterations, stop=STOP_MAX_UPDATE, tol=convergence_threshold,
                     omega=-1.0)
    out, st = _shared_ctx.poisson_solve(D, np.asarray(phi)[:height, :width], o,
                                        allow=(OK, NOT_CONVERGED, NAN_))
    phi[:height, :width] = out
    return st


class CausticDesign:
    """Mirror of class Caustic_design (src/caustic_design.h:7-66) on the device."""

    def __init__(self, ctx: Optional[Context] = None):
        self.ctx = ctx or Context(int(os.environ.get("CD_DEVICE", "0")))
        self.L = self.ctx.L
        self.mesh_res_x = self.mesh_res_y = 0
        self.resolution_x = self.resolution_y = 0
        self.width = self.height = 0.0
        self.focal_l = 0.0
        self.thickness = 0.0
        self.nthreads = 0
        self.phi_opts: Optional[SolveOpts] = None
        self.h_opts: Optional[SolveOpts] = None
        self.step_size: Optional[float] = None  # None: the reference's (double)0.05f
        self.first_hit = 0
        self.d = None

    # setters of caustic_design.cpp:28-53
    def set_mesh_resolution(self, w, h): self.mesh_res_x, self.mesh_res_y = int(w), int(h)
    def set_domain_resolution(self, w, h): self.resolution_x, self.resolution_y = int(w), int(h)
    def set_mesh_size(self, w, h): self.width, self.height = float(w), float(h)
    def set_lens_focal_length(self, f): self.focal_l = float(f)
    def set_lens_thickness(self, t): self.thickness = float(t)
    def set_solver_max_threads(self, n): self.nthreads = int(n)

    def set_solve_opts(self, phi: Optional[SolveOpts] = None, h: Optional[SolveOpts] = None):
        self.phi_opts = phi or self.phi_opts
        self.h_opts = h or self.h_opts
        if self.d is not None:
            self.ctx.check(self.L.cd_design_set_solve_opts(
                self.d, C.byref(self.phi_opts) if self.phi_opts else None,
                C.byref(self.h_opts) if self.h_opts else None), "set_solve_opts")

    def _params(self) -> DesignParams:
        p = DesignParams()
        self.L.cd_design_params_default(C.byref(p))
        p.mesh_res_x, p.mesh_res_y = self.mesh_res_x, self.mesh_res_y
        p.res_x, p.res_y = self.resolution_x, self.resolution_y
        p.width, p.height = self.width, self.height
        p.focal_l, p.thickness = self.focal_l, self.thickness
        if self.phi_opts is not None:
            p.phi_solve = self.phi_opts
        if self.h_opts is not None:
            p.h_solve = self.h_opts
        if self.step_size is not None:
            p.step_size = self.step_size
        return p

    def create(self):
        if self.d is not None:
            self.L.cd_design_destroy(self.d)
        h = _P()
        p = self._params()
        self.ctx.check(self.L.cd_design_create(self.ctx.h, C.byref(p), C.byref(h)),
                       "cd_design_create")
        self.d = h
        if self.first_hit:
            self.ctx.check(self.L.cd_design_set_first_hit(self.d, self.first_hit), "set_first_hit")
        self.ctx.children.add(self)

    def close(self):
        if getattr(self, "d", None):
            self.L.cd_design_destroy(self.d)
            self.d = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def initialize_solvers(self, image):
        """initialize_solvers (caustic_design.cpp:334-364); image [res_y][res_x]."""
        image = _f64(image)
        if image.shape != (self.resolution_y, self.resolution_x):
            raise ValueError(f"image shape {image.shape} != {(self.resolution_y, self.resolution_x)}")
        self.create()
        self.ctx.check(self.L.cd_design_init(self.d, image.ctypes.data), "initialize_solvers")

    def initialize_solvers_device(self, d_image: int):
        self.create()
        self.ctx.check(self.L.cd_design_init_device(self.d, d_image), "initialize_solvers")

    def reinitialize_device(self, d_image: int):
        """Re-run initialisation on an existing design (device-resident image)."""
        self.ctx.check(self.L.cd_design_init_device(self.d, d_image), "initialize_solvers")

    def perform_transport_iteration(self) -> float:
        s = _D()
        self.ctx.check(self.L.cd_design_transport_step(self.d, C.byref(s)),
                       "perform_transport_iteration")
        return s.value

    def perform_height_map_iteration(self, itr: int = 0) -> float:
        u = _D()
        self.ctx.check(self.L.cd_design_height_step(self.d, C.byref(u)),
                       "perform_height_map_iteration")
        return u.value

    def run(self, max_transport_steps=50, conv_tres=float(np.float32(0.01)), height_steps=3,
            allow=(OK,)):
        rs = RunStats()
        rc = self.L.cd_design_run(self.d, max_transport_steps, conv_tres, height_steps,
                                  C.byref(rs))
        self.ctx.check(rc, "cd_design_run", allow)
        return rs

    def laplacian_smoothing(self, factor: float):
        """mesh->laplacian_smoothing(mesh->target_points, factor) (mesh.cpp:534-606)."""
        self.ctx.check(self.L.cd_design_laplacian_smoothing(self.d, factor), "laplacian_smoothing")

    def set_fold_check(self, fail: bool):
        """fail: a transport step that leaves inverted triangles raises CDError(FOLDED)."""
        self.ctx.check(self.L.cd_design_set_fold_check(self.d, int(fail)), "set_fold_check")

    def count_folds(self) -> int:
        """Inverted triangles of the current target mesh."""
        n = _LL()
        self.ctx.check(self.L.cd_design_count_folds(self.d, C.byref(n)), "count_folds")
        return n.value

    def set_miss_recovery(self, on: bool):
        self.ctx.check(self.L.cd_design_set_miss_recovery(self.d, int(on)), "set_miss_recovery")

    def set_first_hit(self, mode: int):
        """0: a sample held by several triangles takes the lowest index; 1: the
        one the reference's BVH query returns (bvh.cpp:196-266).  Kept for the
        design that initialize_solvers creates, applied at once to an existing one."""
        if int(mode) not in (0, 1):
            raise CDError(BAD_ARG, f"set_first_hit: mode {mode} is not 0 or 1")
        self.first_hit = int(mode)
        if self.d is not None:
            self.ctx.check(self.L.cd_design_set_first_hit(self.d, self.first_hit), "set_first_hit")

    # file outputs (caustic_design.cpp:20-26,55-61; main.cpp:269)
    def save_solid_obj_source(self, filename: str):
        self.ctx.check(self.L.cd_design_save_solid_obj(self.d, filename.encode(), 0),
                       "save_solid_obj_source")

    def save_solid_obj_target(self, filename: str):
        self.ctx.check(self.L.cd_design_save_solid_obj(self.d, filename.encode(), 1),
                       "save_solid_obj_target")

    def save_heightmap_json(self, filename: str):
        self.ctx.check(self.L.cd_design_save_heightmap_json(self.d, filename.encode()),
                       "save_heightmap_json")

    def export_paramererization_to_svg(self, filename: str, line_width: float):
        self.ctx.check(self.L.cd_design_export_svg(self.d, filename.encode(), 0, line_width),
                       "export_paramererization_to_svg")

    def export_inverted_transport_map(self, filename: str, stroke_width: float):
        self.ctx.check(self.L.cd_design_export_svg(self.d, filename.encode(), 1, stroke_width),
                       "export_inverted_transport_map")

    def export_to_svg(self, filename: str, stroke_width: float):
        self.ctx.check(self.L.cd_design_export_svg(self.d, filename.encode(), 2, stroke_width),
                       "export_to_svg")

    def get(self, name: str) -> np.ndarray:
        f = FIELD[name]
        n = self.L.cd_design_field_size(self.d, f)
        out = np.empty(n)
        self.ctx.check(self.L.cd_design_get(self.d, f, out.ctypes.data, n), f"get {name}")
        V = self.mesh_res_x * self.mesh_res_y
        if name in ("target_points", "source_points"):
            return out.reshape(V, 3)
        if name in ("vertex_gradient", "inverse_map"):
            return out.reshape(V, 2)
        if name == "normals":
            return out.reshape(2, V)
        if n == self.resolution_x * self.resolution_y and name not in ("errors", "target_areas"):
            return out.reshape(self.resolution_y, self.resolution_x)
        return out

    def set(self, name: str, values):
        f = FIELD[name]
        v = _f64(values).ravel()
        self.ctx.check(self.L.cd_design_set(self.d, f, v.ctypes.data, v.size), f"set {name}")

    def device_ptr(self, name: str) -> int:
        return self.L.cd_design_device_ptr(self.d, FIELD[name])

    def last_solve_stats(self):
        a, b = SolveStats(), SolveStats()
        self.ctx.check(self.L.cd_design_last_solve_stats(self.d, C.byref(a), C.byref(b)), "stats")
        return a, b

    # public members of the reference class
    phi = property(lambda self: self.get("phi"))
    h = property(lambda self: self.get("h"))
    raster = property(lambda self: self.get("raster"))
    errors = property(lambda self: self.get("errors"))
    target_areas = property(lambda self: self.get("target_areas"))
    divergence = property(lambda self: self.get("divergence"))
    norm_x = property(lambda self: self.get("norm_x"))
    norm_y = property(lambda self: self.get("norm_y"))
    pixels = property(lambda self: self.get("pixels"))
    normals = property(lambda self: self.get("normals"))
    vertex_gradient = property(lambda self: self.get("vertex_gradient").T.copy())


BATCH_MAX = 64


def run_designs_batch(designs, d_images=None, max_transport_steps=50,
                      conv_tres=float(np.float32(0.01)), height_steps=3, allow=(OK,)):
    """Run a batch of designs concurrently, one per context (each on its own
    HIP stream): optional re-initialisation from device images, then the
    driver loop of main.cpp:243-262 for each.  Returns one RunStats per
    design; raises on the first design whose status is not allowed."""
    designs = list(designs)
    n = len(designs)
    if n == 0:
        return []
    if any(d.d is None for d in designs):
        raise CDError(STATE, "cd_designs_run_batch: initialise every design first")
    L = designs[0].L
    D = (C.c_void_p * n)(*[d.d for d in designs])
    imgs = None
    if d_images is not None:
        if len(d_images) != n:
            raise ValueError("one device image (or None) per design")
        imgs = (C.c_void_p * n)(*[p if p else None for p in d_images])
    stats = (RunStats * n)()
    rc = L.cd_designs_run_batch(D, imgs, n, max_transport_steps, conv_tres, height_steps, stats)
    if rc == BAD_ARG:
        raise CDError(rc, "cd_designs_run_batch: every design needs its own context "
                          f"and n <= {BATCH_MAX}")
    for d, r in zip(designs, stats):
        d.ctx.check(r.status, "cd_designs_run_batch", allow)
    return list(stats)


def bvh_first_hit_order(xy, mx: int, my: int) -> np.ndarray:
    """Host: the triangle at each position of the reference BVH's build array
    (bvh.cpp:20-155) for mesh vertices xy [mx*my, 2]."""
    L = load()
    xy = np.ascontiguousarray(np.asarray(xy, dtype=np.float64)[:, :2])
    assert xy.shape == (mx * my, 2)
    out = np.empty(2 * (mx - 1) * (my - 1), dtype=np.int32)
    rc = L.cd_bvh_first_hit_order(xy.ctypes.data, mx, my, out.ctypes.data)
    if rc != OK:
        raise CDError(rc, f"cd_bvh_first_hit_order: {status_string(rc)}")
    return out


def png_read_rgba8(path: str) -> np.ndarray:
    """PNG -> [h][w][4] uint8 with the reference's libpng conversions."""
    L = load()
    w, h = _I(), _I()
    rc = L.cd_png_read_rgba8(path.encode(), C.byref(w), C.byref(h), None, 0)
    if rc != OK:
        raise CDError(rc, f"cd_png_read_rgba8({path}): {status_string(rc)}")
    out = np.empty((h.value, w.value, 4), dtype=np.uint8)
    rc = L.cd_png_read_rgba8(path.encode(), C.byref(w), C.byref(h), out.ctypes.data, out.size)
    if rc != OK:
        raise CDError(rc, f"cd_png_read_rgba8({path}): {status_string(rc)}")
    return out


def image_gray_resize(ctx: Context, rgba: np.ndarray, out_w: int, out_h: int) -> np.ndarray:
    """Gray + nearest resize on the device (main.cpp:13-27, 95-98)."""
    rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
    h, w = rgba.shape[:2]
    out = np.empty((out_h, out_w))
    ctx.check(ctx.L.cd_image_gray_resize(ctx.h, rgba.ctypes.data, w, h, out_w, out_h,
                                         out.ctypes.data), "cd_image_gray_resize")
    return out


def obj_write_solid(path: str, front, back, res_x: int, res_y: int, thickness: float,
                    width: float, height: float):
    f, b = _f64(front), _f64(back)
    rc = load().cd_obj_write_solid(path.encode(), f.ctypes.data, b.ctypes.data, res_x, res_y,
                                   thickness, width, height)
    if rc != OK:
        raise CDError(rc, f"cd_obj_write_solid({path})")


def json_write_heightmap(path: str, h):
    h = _f64(h)
    rc = load().cd_json_write_heightmap(path.encode(), h.ctypes.data, h.shape[1], h.shape[0])
    if rc != OK:
        raise CDError(rc, f"cd_json_write_heightmap({path})")


def svg_write(path: str, points, res_x: int, res_y: int, width: float, height: float,
              stroke_width: float, triangles: bool = False):
    p = _f64(points)
    fn = load().cd_svg_write_triangles if triangles else load().cd_svg_write_grid
    rc = fn(path.encode(), p.ctypes.data, p.shape[1], res_x, res_y, width, height, stroke_width)
    if rc != OK:
        raise CDError(rc, f"svg write ({path})")


def default_arch(**kw) -> NetArch:
    a = NetArch()
    load().cd_net_arch_default(C.byref(a))
    for k, v in kw.items():
        if k == "channels":
            for i, c in enumerate(v):
                a.channels[i] = c
        else:
            setattr(a, k, v)
    return a


class Net:
    """The learned-init network (cd_net) on one context."""

    def __init__(self, ctx: Context, arch: Optional[NetArch] = None):
        self.ctx, self.L = ctx, ctx.L
        self.arch = arch or default_arch()
        h = _P()
        ctx.check(self.L.cd_net_create(ctx.h, C.byref(self.arch), C.byref(h)), "cd_net_create")
        self.h = h
        ctx.children.add(self)
        self.n = self.L.cd_net_num_params(h)

    def close(self):
        if getattr(self, "h", None):
            self.L.cd_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_params(self, seed: int = 0):
        self.ctx.check(self.L.cd_net_init_params(self.h, seed), "init_params")

    def get_params(self) -> np.ndarray:
        p = np.empty(self.n, dtype=np.float32)
        self.ctx.check(self.L.cd_net_get_params(self.h, p.ctypes.data, self.n), "get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self.ctx.check(self.L.cd_net_set_params(self.h, p.ctypes.data, p.size), "set_params")

    def forward(self, d_x: int, d_y: int, batch: int, h: int, w: int):
        self.ctx.check(self.L.cd_net_forward(self.h, d_x, d_y, batch, h, w), "net forward")

    def set_fused_forward(self, on: bool):
        """Inference forwards fold the decoder upsampling into the convolutions
        (default) or materialise it as training does; outputs are identical."""
        self.ctx.check(self.L.cd_net_set_fused_forward(self.h, int(on)), "set_fused_forward")

    def set_tiled_forward(self, on: bool):
        """Forward convolutions on LDS-staged 8x8 tiles (default) or with
        per-chunk global gathers; bit-identical outputs."""
        self.ctx.check(self.L.cd_net_set_tiled_forward(self.h, int(on)), "set_tiled_forward")

    def loss_grad(self, d_x: int, d_t: int, batch: int, h: int, w: int):
        loss = _D()
        g = np.empty(self.n, dtype=np.float32)
        self.ctx.check(self.L.cd_net_loss_grad(self.h, d_x, d_t, batch, h, w, C.byref(loss),
                                               g.ctypes.data, self.n), "net loss_grad")
        return loss.value, g

    def train_step(self, d_x: int, d_t: int, batch: int, h: int, w: int, lr: float) -> float:
        loss = _D()
        self.ctx.check(self.L.cd_net_train_step(self.h, d_x, d_t, batch, h, w, lr, C.byref(loss)),
                       "net train_step")
        return loss.value

    def loss_grad_operator(self, d_x: int, batch: int, h: int, w: int):
        loss = _D()
        g = np.empty(self.n, dtype=np.float32)
        self.ctx.check(self.L.cd_net_loss_grad_operator(self.h, d_x, batch, h, w, C.byref(loss),
                                                        g.ctypes.data, self.n),
                       "net loss_grad_operator")
        return loss.value, g

    def train_step_operator(self, d_x: int, batch: int, h: int, w: int, lr: float) -> float:
        loss = _D()
        self.ctx.check(self.L.cd_net_train_step_operator(self.h, d_x, batch, h, w, lr,
                                                         C.byref(loss)), "net train_step_operator")
        return loss.value
