"""Shared test helpers: exact discrete Poisson solutions and small designs."""
import numpy as np
import scipy.fft


def dct_solve(b: np.ndarray) -> np.ndarray:
    """Exact zero-mean solution of  sum_nbr phi - cnt*phi = P b  for the
    cell-centred Neumann 5-point operator (no holes): the 2-D DCT-II
    diagonalises it with eigenvalues -4 sin^2(pi k/2w) - 4 sin^2(pi l/2h)."""
    h, w = b.shape
    b = b - b.mean()
    bh = scipy.fft.dctn(b, type=2, norm="ortho")
    k = np.arange(w)
    l = np.arange(h)
    lam = -4 * np.sin(np.pi * k / (2 * w))[None, :] ** 2 - 4 * np.sin(np.pi * l / (2 * h))[:, None] ** 2
    lam[0, 0] = 1.0
    ph = bh / lam
    ph[0, 0] = 0.0
    return scipy.fft.idctn(ph, type=2, norm="ortho")


def apply_operator(phi: np.ndarray) -> np.ndarray:
    """(A phi)_c = sum of in-domain neighbours - cnt * phi_c (solver.cpp:29-47)."""
    h, w = phi.shape
    out = np.zeros_like(phi)
    cnt = np.zeros_like(phi)
    out[:, 1:] += phi[:, :-1]; cnt[:, 1:] += 1
    out[:, :-1] += phi[:, 1:]; cnt[:, :-1] += 1
    out[1:, :] += phi[:-1, :]; cnt[1:, :] += 1
    out[:-1, :] += phi[1:, :]; cnt[:-1, :] += 1
    return out - cnt * phi


def eigvec(w, h, k, l):
    x = (np.arange(w) + 0.5) / w
    y = (np.arange(h) + 0.5) / h
    v = np.cos(np.pi * l * y)[:, None] * np.cos(np.pi * k * x)[None, :]
    lam = -4 * np.sin(np.pi * k / (2 * w)) ** 2 - 4 * np.sin(np.pi * l / (2 * h)) ** 2
    return v, lam


def smooth_rhs(rng, w, h, n_modes=6):
    """Random smooth zero-mean right-hand side."""
    b = np.zeros((h, w))
    for _ in range(n_modes):
        k, l = rng.integers(0, 6, 2)
        if k == 0 and l == 0:
            continue
        v, _ = eigvec(w, h, k, l)
        b += rng.standard_normal() * v
    b += 0.05 * rng.standard_normal((h, w))
    return b - b.mean()


def small_image(seed, n):
    from poisson_caustic_design_amd import synth
    return synth.resize_nearest(synth.gray_from_u8(synth.gauss5(seed, 64)), n, n)


def mesh_triangles(mx, my):
    """Vertex indices of the structured triangulation of mesh.cpp:57-63:
    cell (i, j) with base idx = i*mx + j holds {idx, idx+1, idx+mx} and
    {idx+mx, idx+1, idx+mx+1}."""
    i, j = np.meshgrid(np.arange(my - 1), np.arange(mx - 1), indexing="ij")
    idx = (i * mx + j).ravel()
    t = np.empty((2 * idx.size, 3), dtype=np.int64)
    t[0::2] = np.stack([idx, idx + 1, idx + mx], 1)
    t[1::2] = np.stack([idx + mx, idx + 1, idx + mx + 1], 1)
    return t


def fold_count(xy, mx, my):
    """Triangles whose signed shoelace area is not positive (inverted or
    collapsed); every triangle of the undeformed mesh has area +dx*dy/2."""
    xy = np.asarray(xy, dtype=np.float64).reshape(-1, xy.shape[-1])[:, :2]
    t = mesh_triangles(mx, my)
    x, y = xy[t, 0], xy[t, 1]
    area2 = (x[:, 0] * y[:, 1] - x[:, 1] * y[:, 0] + x[:, 1] * y[:, 2] - x[:, 2] * y[:, 1]
             + x[:, 2] * y[:, 0] - x[:, 0] * y[:, 2])
    return int(np.count_nonzero(~(area2 > 0)))
