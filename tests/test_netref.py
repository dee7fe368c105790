"""Pins the FP64 network restatement (oracle/netref.py) by central finite
differences; it is then the checker of the GPU network kernels."""
import numpy as np
import pytest

from oracle.netref import NetRef, bf16


def test_netref_gradient_finite_differences():
    net = NetRef(levels=3, channels=(8, 16, 16))
    p = net.init_params(3).astype(np.float64)
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 8, 8))
    t = rng.standard_normal((2, 8, 8)) * 0.1
    loss, g = net.loss_grad(p, x, t)
    nz = np.flatnonzero(p)
    idx = rng.choice(nz, 25, replace=False)
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (net.loss_grad(pp, x, t)[0] - net.loss_grad(pm, x, t)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_smoothing_step_closed_form():
    from tests.helpers import eigvec
    e, lam = eigvec(12, 10, 2, 3)
    # S = I + Lc/8 scales an eigenvector by 1 + lambda/8 (lambda <= 0)
    assert np.allclose(NetRef.smooth(e[None], 3)[0], (1 + lam / 8.0) ** 3 * e, atol=1e-13)
    c = np.full((1, 5, 7), 2.5)
    assert np.array_equal(NetRef.smooth(c, 4), c)  # constants pass unchanged


def test_linear_init_is_the_zero_operator():
    net = NetRef(levels=3, channels=(8, 16, 16), linear=True)
    p = net.init_params(1).astype(np.float64)
    x = np.random.default_rng(2).standard_normal((1, 16, 16))
    y, _ = net.forward(p, x)
    assert np.all(y == 0)
    loss, g = net.loss_grad_operator(p, x)
    assert loss == pytest.approx(np.mean(x ** 2))  # u = -x
    assert np.abs(g).max() > 0  # the head still receives a gradient


def test_netref_shapes_and_layout():
    net = NetRef()
    assert len(net.convs) == 4 * 4 - 1
    p = net.init_params(0)
    assert p.size == net.nparam
    y, _ = net.forward(p.astype(np.float64), np.zeros((1, 16, 16)))
    assert y.shape == (1, 16, 16) and np.all(y == 0)  # bias-free: zero in, zero out
    # positive homogeneity of a bias-free ReLU network
    x = np.random.default_rng(1).standard_normal((1, 16, 16))
    y1, _ = net.forward(p.astype(np.float64), x)
    y2, _ = net.forward(p.astype(np.float64), 3.0 * x)
    assert np.allclose(y2, 3.0 * y1)


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5])
    r = bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # exact tie -> even mantissa
    assert r[2] == 1.0 and r[3] == 1.0 + 2 ** -7 and r[4] == -2.5


@pytest.mark.parametrize("out_smooth,loss_smooth,energy", [(0, 0, False), (2, 3, False),
                                                           (2, 0, True)])
def test_netref_operator_gradient_finite_differences(out_smooth, loss_smooth, energy):
    for linear in (True, False):
        net = NetRef(levels=3, channels=(8, 16, 16), linear=linear, out_smooth=out_smooth,
                     loss_smooth=loss_smooth, loss_energy=energy)
        p = net.random_params(4).astype(np.float64)
        rng = np.random.default_rng(2)
        x = rng.standard_normal((2, 8, 8))
        loss, g = net.loss_grad_operator(p, x)
        idx = rng.choice(np.flatnonzero(p), 20, replace=False)
        eps = 1e-6
        for i in idx:
            pp, pm = p.copy(), p.copy()
            pp[i] += eps
            pm[i] -= eps
            fd = (net.loss_grad_operator(pp, x)[0] - net.loss_grad_operator(pm, x)[0]) / (2 * eps)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (linear, i, fd, g[i])


def test_netref_operator_matches_independent_stencil_and_exact_inverse():
    from tests.helpers import apply_operator, dct_solve, eigvec
    rng = np.random.default_rng(3)
    v = rng.standard_normal((12, 20))
    assert np.allclose(NetRef.lap(v[None])[0], apply_operator(v), atol=1e-12)
    e, lam = eigvec(20, 12, 3, 2)
    assert np.allclose(NetRef.lap(e[None])[0], lam * e, atol=1e-12)
    # the loss vanishes exactly at the inverse operator: y = Lc^{-1} x / cs
    x = rng.standard_normal((12, 20))
    x -= x.mean()
    cs = (20 / np.pi) ** 2
    y = dct_solve(x) / cs
    u = cs * NetRef.lap(y[None])[0] - x
    assert np.abs(u).max() < 1e-9


def test_energy_loss_is_minimal_at_the_inverse_operator():
    """J(y) = mean(y (x - u) / 2), u = cs Lc y - x, is the energy
    y'(-cs Lc)y/2 + x'y per pixel: J(y* + d) - J(y*) = mean(d (-cs Lc) d) / 2
    >= 0 at y* = Lc^{-1} x / cs (exact DCT inverse), for any d."""
    from tests.helpers import dct_solve
    rng = np.random.default_rng(6)
    x = rng.standard_normal((12, 20))
    x -= x.mean()
    cs = (20 / np.pi) ** 2

    def J(y):
        u = cs * NetRef.lap(y[None])[0] - x
        return np.mean(0.5 * y * (x - u))

    ys = dct_solve(x) / cs
    for k in range(5):
        d = rng.standard_normal((12, 20)) * 10.0 ** -k
        expect = np.mean(0.5 * d * (-cs * NetRef.lap(d[None])[0]))
        assert expect > 0
        assert J(ys + d) - J(ys) == pytest.approx(expect, rel=1e-6, abs=1e-12)


def test_linear_netref_is_linear():
    net = NetRef(levels=3, channels=(8, 16, 16), linear=True)
    p = net.random_params(5).astype(np.float64)
    rng = np.random.default_rng(4)
    x1, x2 = rng.standard_normal((2, 1, 16, 16))
    y1, _ = net.forward(p, x1)
    y2, _ = net.forward(p, x2)
    y3, _ = net.forward(p, 2.0 * x1 - 0.5 * x2)
    assert np.allclose(y3, 2.0 * y1 - 0.5 * y2, atol=1e-12)
