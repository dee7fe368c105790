"""Restatement of the reference's image ingest (test infrastructure).

* gray: image_to_grid (main.cpp:86-101): per RGBA8 pixel, r = R/255.0,
  g = G/255.0, b = B/255.0, gray = (0.299*r) + (0.587*g) + (0.114*b), in that
  operation order (numpy float64, no fused multiply-add).
* resize: resize_image (main.cpp:13-27): nearest, src_x = x*old_w // new_w,
  src_y = y*old_h // new_h in integer arithmetic.
* the driver's geometry (main.cpp:218-232): aspect = w/h, resized to
  (4*M, int(4*M/aspect)), mesh (M, int(M/aspect)), mesh height
  floor(M/aspect) * (width/M).
"""
from __future__ import annotations

import math

import numpy as np


def gray(rgba: np.ndarray) -> np.ndarray:
    c = rgba.astype(np.float64)
    r, g, b = c[..., 0] / 255.0, c[..., 1] / 255.0, c[..., 2] / 255.0
    return (0.299 * r) + (0.587 * g) + (0.114 * b)


def resize(img: np.ndarray, new_w: int, new_h: int) -> np.ndarray:
    old_h, old_w = img.shape[:2]
    sx = (np.arange(new_w, dtype=np.int64) * old_w) // new_w
    sy = (np.arange(new_h, dtype=np.int64) * old_h) // new_h
    return img[sy][:, sx]


def driver_geometry(img_w: int, img_h: int, mesh_res: int, lens_width: float):
    """(grid_w, grid_h, mesh_x, mesh_y, mesh_height) as main.cpp:218-232 sets them."""
    aspect = img_w / img_h
    grid_w, grid_h = 4 * mesh_res, int(4 * mesh_res / aspect)
    mesh_x, mesh_y = mesh_res, int(mesh_res / aspect)
    mesh_height = math.floor(mesh_res / aspect) * (lens_width / mesh_res)
    return grid_w, grid_h, mesh_x, mesh_y, mesh_height


def heightmap_json(h: np.ndarray) -> str:
    """save_heightmap_as_json (main.cpp:111-135); std::ostream's default
    double output is printf's %g."""
    rows = ["  [" + ", ".join("%g" % v for v in row) + "]" for row in h]
    return "[\n" + ",\n".join(rows) + "\n]\n"
