"""Pins the FP64 network restatement (oracle/netref.py) by central finite
differences and by closed forms of its layers and wiring; it is then the
checker of the GPU network kernels."""
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


# ---- closed-form checks of the layers and of the layer wiring -------------

def _set_tap(p, cv, co, ci, tap, val):
    """Write one weight of the flat [coutpad][kpad] layout, k = tap*Cin + ci."""
    p[cv.woff + co * cv.kpad + tap * (cv.c1 + cv.c2) + ci] = val


def test_conv_is_a_zero_padded_cross_correlation():
    from scipy.signal import correlate2d
    rng = np.random.default_rng(5)
    x = rng.standard_normal((1, 7, 9, 3))
    wt = rng.standard_normal((9, 3, 2))
    y = NetRef.conv(x, wt)
    for co in range(2):
        want = sum(correlate2d(x[0, :, :, ci], wt[:, ci, co].reshape(3, 3), mode="same",
                               boundary="fill") for ci in range(3))
        assert np.allclose(y[0, :, :, co], want, atol=1e-12)
    # an impulse reproduces the taps: tap (dy, dx) lands at offset (1 - dy, 1 - dx)
    d = np.zeros((1, 5, 5, 3))
    d[0, 2, 2, 1] = 1.0
    y = NetRef.conv(d, wt)
    for tap in range(9):
        dy, dx = divmod(tap, 3)
        assert np.array_equal(y[0, 3 - dy, 3 - dx], wt[tap, 1])


def test_pool_and_upsampling_closed_forms():
    c = np.full((2, 6, 8, 3), -1.75)
    assert np.array_equal(NetRef.pool(c), c[:, :3, :4])  # constants survive the pool
    chk = np.indices((6, 8)).sum(axis=0) % 2 * 2.0 - 1.0
    assert np.all(NetRef.pool(chk[None, :, :, None]) == 0)  # a checkerboard averages out
    rng = np.random.default_rng(6)
    x = rng.standard_normal((1, 4, 6, 2))
    assert np.array_equal(NetRef.pool(NetRef.up(x)), x)
    assert np.array_equal(NetRef.up(x)[0, 3, 5], x[0, 1, 2])  # nearest neighbour
    # the backward passes are the adjoints
    X, G = rng.standard_normal((1, 8, 12, 2)), rng.standard_normal((1, 4, 6, 2))
    assert np.isclose(np.sum(NetRef.pool(X) * G), np.sum(X * NetRef.pool_bwd(G)))
    assert np.isclose(np.sum(NetRef.up(G) * X), np.sum(G * NetRef.up_bwd(X)))


def _identity_net(levels, through_skip):
    """Linear network whose every convolution copies channel 0 (centre tap) and
    whose decoder convolutions take either the upsampled path or the skip."""
    net = NetRef(levels=levels, channels=(8, 16, 16, 16)[:levels], linear=True)
    p = np.zeros(net.nparam)
    for cv in net.convs:
        ci = cv.c1 if (cv.c2 > 0 and through_skip) else 0
        _set_tap(p, cv, 0, ci, 4, 1.0)
    return net, p


def test_identity_wiring_reproduces_block_means_and_skips():
    x = np.random.default_rng(7).standard_normal((2, 16, 16))
    for levels in (2, 3):
        net, p = _identity_net(levels, through_skip=False)
        y, _ = net.forward(p, x)
        f = 2 ** (levels - 1)  # the deepest level sees f x f block means
        means = x.reshape(2, 16 // f, f, 16 // f, f).mean(axis=(2, 4))
        assert np.allclose(y, np.repeat(np.repeat(means, f, axis=1), f, axis=2), atol=1e-13)
        net, p = _identity_net(levels, through_skip=True)
        y, _ = net.forward(p, x)
        assert np.array_equal(y, x)  # the first skip carries the input unchanged


def test_head_stencil_and_relu_closed_forms():
    from scipy.ndimage import correlate
    x = np.random.default_rng(8).standard_normal((1, 10, 12))
    net = NetRef(levels=1, channels=(8,), linear=True)
    p = np.zeros(net.nparam)
    _set_tap(p, net.convs[0], 0, 0, 4, 1.0)
    _set_tap(p, net.convs[1], 0, 0, 4, 1.0)
    st = np.array([[0, 1, 0], [1, -4, 1], [0, 1, 0]], float)
    for tap in range(9):
        _set_tap(p, net.convs[2], 0, 0, tap, st.flat[tap])
    y, _ = net.forward(p, x)
    assert np.allclose(y[0], correlate(x[0], st, mode="constant"), atol=1e-13)
    relu_net = NetRef(levels=1, channels=(8,))
    y, _ = relu_net.forward(p, x)  # same weights, ReLU after the two inner layers
    assert np.allclose(y[0], correlate(np.maximum(x[0], 0), st, mode="constant"), atol=1e-13)
