"""GPU tests of the learned coarse solve: MFMA network kernels against the
FP64 restatement (oracle/netref.py) for both losses, training, capture of
solver-generated inputs, the fallback to plain SOR, and the gain of a trained
network in the preconditioned solver."""
import numpy as np
import pytest

from oracle.netref import NetRef
from poisson_caustic_design_amd import capi, learned, synth
from tests.helpers import dct_solve, smooth_rhs

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def small_arch(linear=1, **kw):
    return capi.default_arch(levels=3, channels=[8, 16, 16], linear=linear, **kw)


def dev(a):
    return torch.tensor(np.ascontiguousarray(a, dtype=np.float32), device="cuda:0")


@pytest.mark.parametrize("linear", [1, 0])
def test_net_forward_matches_netref(ctx, linear):
    net = capi.Net(ctx, small_arch(linear))
    ref = NetRef(3, (8, 16, 16), linear=bool(linear))
    net.set_params(ref.random_params(5))
    p = net.get_params()
    assert ref.nparam == net.n
    x = np.random.default_rng(0).standard_normal((2, 16, 24))
    dx, dy = dev(x), torch.zeros((2, 16, 24), device="cuda:0")
    torch.cuda.synchronize()
    net.forward(dx.data_ptr(), dy.data_ptr(), 2, 16, 24)
    y = dy.cpu().numpy()
    yr, _ = ref.forward(p.astype(np.float64), x, emulate_bf16=True)
    assert np.abs(y - yr).max() <= 2e-2 * np.abs(yr).max()


@pytest.mark.parametrize("linear,out_smooth", [(1, 0), (0, 0), (1, 6), (0, 6)])
def test_fused_forward_is_bit_identical(ctx, linear, out_smooth):
    """Upsampled decoder inputs and pooled encoder inputs read through the
    convolutions, the two convolutions of a level in one launch, and (with 6
    output smoothing steps, the shipped setting) the head convolution and the
    smoothing in one launch (inference) give exactly the output of the
    materialised tensors and separate launches (training), on a batch of two
    32 x 40 fields (partial tiles at every level)."""
    arch = capi.default_arch(linear=linear, channels=(16, 32, 32, 32, 0, 0),
                             out_smooth=out_smooth)
    net = capi.Net(ctx, arch)
    net.set_params(NetRef(4, (16, 32, 32, 32), linear=bool(linear)).random_params(3))
    x = np.random.default_rng(2).standard_normal((2, 32, 40))
    dx = dev(x)
    ys = []
    for fused in (True, False):
        net.set_fused_forward(fused)
        dy = torch.zeros((2, 32, 40), device="cuda:0")
        torch.cuda.synchronize()
        net.forward(dx.data_ptr(), dy.data_ptr(), 2, 32, 40)
        ys.append(dy.cpu().numpy())
    assert np.abs(ys[0]).max() > 0
    assert np.array_equal(ys[0], ys[1])


@pytest.mark.parametrize("channels,shape,linear", [((16, 32, 32, 32), (2, 32, 40), 1),
                                                   ((16, 32, 64, 64), (1, 128, 128), 1),
                                                   ((16, 32, 64, 64), (3, 24, 56), 0)])
def test_tiled_forward_is_bit_identical(ctx, channels, shape, linear):
    """Forward convolutions on LDS-staged 8x8 tiles (the default) give exactly
    the output of the per-chunk gather kernel, and with the fused inference
    forward (pairs of convolutions in one launch, conv b's weights staged with
    conv a's or after it, and the 64+64-channel pair that does not fit the LDS
    left as two launches) exactly that of the materialised tensors: the
    architectures 16/32/32/32 (the shipped weights) and 16/32/64/64 (so
    64+64-channel decoder inputs and 64-channel outputs split over two
    workgroups), ReLU, batches, and image sides that leave partial tiles at the
    coarsest level."""
    arch = capi.default_arch(linear=linear, channels=(*channels, 0, 0))
    net = capi.Net(ctx, arch)
    net.set_params(NetRef(4, channels, linear=bool(linear)).random_params(5))
    b, h, w = shape
    dx = dev(np.random.default_rng(6).standard_normal(shape))
    ys = []
    for tiled in (True, False):
        net.set_tiled_forward(tiled)
        for fused in (True, False):
            net.set_fused_forward(fused)
            dy = torch.zeros(shape, device="cuda:0")
            torch.cuda.synchronize()
            net.forward(dx.data_ptr(), dy.data_ptr(), b, h, w)
            ys.append(dy.cpu().numpy())
    assert np.abs(ys[0]).max() > 0
    for y in ys[1:]:
        assert np.array_equal(ys[0], y)


@pytest.mark.parametrize("linear", [1, 0])
def test_net_gradient_matches_netref(ctx, linear):
    net = capi.Net(ctx, small_arch(linear))
    ref = NetRef(3, (8, 16, 16), linear=bool(linear))
    net.set_params(ref.random_params(7))
    p = net.get_params()
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 16, 16))
    t = 0.3 * rng.standard_normal((2, 16, 16))
    dx, dt = dev(x), dev(t)
    torch.cuda.synchronize()
    loss, g = net.loss_grad(dx.data_ptr(), dt.data_ptr(), 2, 16, 16)
    lr, gr = ref.loss_grad(p.astype(np.float64), x, t, emulate_bf16=True)
    assert loss == pytest.approx(lr, rel=2e-2)
    assert np.linalg.norm(g - gr) <= 5e-2 * np.linalg.norm(gr)
    cos = np.dot(g, gr) / (np.linalg.norm(g) * np.linalg.norm(gr))
    assert cos > 0.998


@pytest.mark.parametrize("linear,out_smooth,loss_smooth,energy",
                         [(1, 0, 0, 0), (0, 0, 0, 0), (1, 2, 3, 0), (1, 2, 0, 1), (0, 0, 0, 1)])
def test_net_operator_gradient_matches_netref(ctx, linear, out_smooth, loss_smooth, energy):
    net = capi.Net(ctx, small_arch(linear, out_smooth=out_smooth, loss_smooth=loss_smooth,
                                   loss_energy=energy))
    ref = NetRef(3, (8, 16, 16), linear=bool(linear), out_smooth=out_smooth,
                 loss_smooth=loss_smooth, loss_energy=bool(energy))
    net.set_params(ref.random_params(9))
    p = net.get_params()
    rng = np.random.default_rng(5)
    x = rng.standard_normal((2, 16, 24))
    dx = dev(x)
    torch.cuda.synchronize()
    loss, g = net.loss_grad_operator(dx.data_ptr(), 2, 16, 24)
    lr, gr = ref.loss_grad_operator(p.astype(np.float64), x, emulate_bf16=True)
    assert loss == pytest.approx(lr, rel=2e-2)
    assert np.linalg.norm(g - gr) <= 5e-2 * np.linalg.norm(gr)
    cos = np.dot(g, gr) / (np.linalg.norm(g) * np.linalg.norm(gr))
    assert cos > 0.998


def test_operator_training_learns_coarse_inverse(ctx):
    """Training on the operator loss alone (random fields, no targets) makes the
    network an approximate inverse of the coarse Neumann operator: held-out
    smooth problems are solved to within tens of per cent (exact solution by
    DCT), which is what a coarse solve inside the FCG iteration needs."""
    x = learned.random_fields(512, 32, 32, seed=1)
    dx = dev(x)
    net = capi.Net(ctx, small_arch())
    net.init_params(2)
    torch.cuda.synchronize()
    losses = learned.train_operator(net, dx, steps=3000, batch=16, lr=2e-3, log=None)
    # the zero operator (the initial network) has loss mean(x^2) = 1
    assert losses[0] == pytest.approx(1.0, rel=1e-3)
    assert np.mean(losses[-50:]) < 0.1
    rng = np.random.default_rng(7)
    xt = smooth_rhs(rng, 32, 32)
    xt /= np.sqrt(np.mean(xt ** 2))
    y = torch.zeros((1, 32, 32), device="cuda:0")
    dxt = dev(xt[None])
    torch.cuda.synchronize()
    net.forward(dxt.data_ptr(), y.data_ptr(), 1, 32, 32)
    yv = y.cpu().numpy()[0] * (32 / np.pi) ** 2
    ex = dct_solve(xt)
    rel = np.linalg.norm((yv - yv.mean()) - ex) / np.linalg.norm(ex)
    print(f"operator loss {np.mean(losses[:10]):.3e} -> {np.mean(losses[-50:]):.3e}, "
          f"held-out relative error {rel:.3f}")
    # a coarse solver inside FCG needs the right shape, not a tight fit
    assert rel <= 0.5


def test_default_arch_forward_and_step(ctx):
    net = capi.Net(ctx)  # 4 levels, 16-64 channels
    net.set_params(NetRef(linear=True).random_params(0))
    x = dev(np.random.default_rng(2).standard_normal((2, 64, 64)))
    t = dev(np.zeros((2, 64, 64)))
    torch.cuda.synchronize()
    l0 = net.train_step(x.data_ptr(), t.data_ptr(), 2, 64, 64, 1e-3)
    l1 = net.train_step(x.data_ptr(), t.data_ptr(), 2, 64, 64, 1e-3)
    assert np.isfinite(l0) and np.isfinite(l1) and l1 < l0


def test_training_reduces_loss(ctx):
    """Fit the normalised coarse Poisson solve on a fixed small batch."""
    rng = np.random.default_rng(3)
    xs, ts = [], []
    for _ in range(4):
        r = smooth_rhs(rng, 16, 16)
        e = dct_solve(r)
        s = np.sqrt(np.mean(r ** 2))
        xs.append(r / s)
        ts.append(e / (s * (16 / np.pi) ** 2))
    x, t = dev(np.array(xs)), dev(np.array(ts))
    net = capi.Net(ctx, small_arch())
    net.init_params(1)
    torch.cuda.synchronize()
    losses = [net.train_step(x.data_ptr(), t.data_ptr(), 4, 16, 16, 3e-3) for _ in range(300)]
    assert np.mean(losses[-20:]) < 0.2 * np.mean(losses[:5])


def test_learned_solve_with_untrained_net_still_converges(ctx):
    """A prediction that raises the residual is rejected and plain SOR finishes."""
    net = capi.Net(ctx)
    net.set_params(NetRef(linear=True).random_params(11, head_scale=3.0))  # garbage
    ctx.set_net(net)
    try:
        b = smooth_rhs(np.random.default_rng(4), 256, 256)
        phi, st = ctx.poisson_solve(b, np.zeros_like(b), capi.default_opts(tol=1e-6, learned_init=1))
    finally:
        ctx.set_net(None)
    assert st.status == capi.OK and st.rel_residual <= 1e-6
    assert st.learned_applied + st.learned_rejected >= 1
    ex = dct_solve(b)
    assert np.abs((phi - phi.mean()) - ex).max() <= 1e-4 * np.abs(ex).max()


def test_learned_solve_that_blows_up_restarts_from_the_warm_start(ctx):
    """A network whose output overflows makes the learned iterations' residual
    non-finite: they are abandoned, the solve goes back to its warm start and
    plain SOR finishes to the tolerance, as many sweeps as plain SOR alone
    needs from there (plus the abandoned cycles' smoothing sweeps)."""
    net = capi.Net(ctx)
    net.set_params(NetRef(linear=True).random_params(11, head_scale=3.0) * 1e20)
    b = smooth_rhs(np.random.default_rng(5), 256, 256)
    p0 = dct_solve(b) * 0.9  # a warm start the blow-up would otherwise destroy
    plain, sp = ctx.poisson_solve(b, p0, capi.default_opts(tol=1e-6))
    ctx.set_net(net)
    try:
        phi, st = ctx.poisson_solve(b, p0, capi.default_opts(tol=1e-6, learned_init=1))
    finally:
        ctx.set_net(None)
    print(f"plain {sp.sweeps} sweeps; learned {st.learned_applied} iterations, "
          f"{st.learned_rejected} abandoned, {st.sweeps} sweeps, rel {st.rel_residual:.2e}")
    assert st.status == capi.OK and st.rel_residual <= 1e-6
    assert st.learned_rejected == 1
    assert st.sweeps <= sp.sweeps + 16 * (st.learned_applied + 1)
    ex = dct_solve(b)
    assert np.abs((phi - phi.mean()) - ex).max() <= 1e-4 * np.abs(ex).max()


def test_intermediate_grid_preconditioner_converges(ctx):
    """mid_sweeps > 0: coarse correction -> intermediate grid (coarsen/4 cells
    per cell) smoothed there -> fine grid.  A non-square grid, an untrained
    small net: the solve still reaches 1e-6 and the DCT solution, and the
    captured graph replays it bit-identically."""
    arch = capi.default_arch(levels=3, channels=[8, 16, 16], linear=1, coarsen=8,
                             smooth_sweeps=4, smooth_omega=1.5, mid_sweeps=4)
    net = capi.Net(ctx, arch)
    net.set_params(NetRef(3, (8, 16, 16), linear=True).random_params(21, head_scale=0.05))
    b = smooth_rhs(np.random.default_rng(8), 256, 192)
    o = capi.default_opts(tol=1e-6, learned_init=1)
    ctx.set_net(net)
    try:
        ctx.set_graphs(False)
        phi_e, st_e = ctx.poisson_solve(b, np.zeros_like(b), o)
        ctx.set_graphs(True)
        phi_g, st_g = ctx.poisson_solve(b, np.zeros_like(b), o)
    finally:
        ctx.set_net(None)
    assert st_e.status == capi.OK and st_e.rel_residual <= 1e-6
    assert st_e.learned_applied > 2
    assert np.array_equal(phi_e, phi_g) and st_e.sweeps == st_g.sweeps
    ex = dct_solve(b)
    assert np.abs((phi_e - phi_e.mean()) - ex).max() <= 1e-4 * np.abs(ex).max()


def test_capture_pairs(ctx):
    x, t = learned.generate_pairs(ctx, [5, 6], mesh_res=32, capacity=256, log=None)
    assert x.shape[0] > 4
    xa, ta = x.cpu().numpy(), t.cpu().numpy()
    assert np.all(np.isfinite(xa)) and np.all(np.isfinite(ta))
    rms = np.sqrt((xa ** 2).mean(axis=(1, 2)))
    assert np.allclose(rms, 1.0, atol=1e-3)
    # the first pair of a cold solve: target = normalised converged solution,
    # which solves the coarse problem approximately
    assert np.abs(ta).max() > 0


def test_trained_network_gains_over_plain_sor(ctx):
    """Train on solver residuals of a few designs plus random fields, then
    held-out designs reach the same 1e-6 residual in every Poisson solve with
    a small fraction of the SOR sweeps."""
    arch = small_arch()
    net, losses, _ = learned.fit(ctx, arch, 32, range(100, 108), random_count=512,
                                 capacity=2048, steps=2000, batch=16, lr=2e-3, rounds=1,
                                 log=None)
    res = {}
    for use in (0, 1):
        ctx.set_net(net if use else None)
        sweeps, iters, worst = 0, 0, 0.0
        for seed in (900, 901, 902):
            img, prm = synth.target_for("mixture", seed, 32)
            o = capi.default_opts(tol=1e-6, learned_init=use)
            cd = learned.design_for(ctx, img, prm, o)
            rs = cd.run(50, float(np.float32(0.01)), 3)
            sweeps += rs.phi_sweeps + rs.h_sweeps
            iters += rs.learned_applied
            worst = max(worst, rs.max_phi_residual, rs.max_h_residual)
            cd.close()
        res[use] = (sweeps, iters, worst)
    ctx.set_net(None)
    print("plain vs learned (sweeps, learned iterations, worst residual):", res)
    assert res[0][2] <= 1e-6 and res[1][2] <= 1e-6
    assert res[1][0] < 0.25 * res[0][0]


def test_shipped_weights_speed_up_grid_2048(ctx):
    """The committed network (weights/learned_grid2048.npz, trained on other
    seeds) on a held-out 2048^2 problem: same 1e-6 residual, no fallback,
    a fraction of the sweeps and of the time of plain SOR."""
    import os
    w = os.path.join(os.path.dirname(capi.LIB_PATH), "weights", "learned_grid2048.npz")
    if not os.path.exists(w):
        pytest.skip("no shipped weights")
    net = learned.load_net(ctx, w)
    img, _ = synth.target_for("mixture", 4242, 512)
    b = img / img.mean() - 1.0
    _, st0 = ctx.poisson_solve(b, np.zeros_like(b), capi.default_opts(tol=1e-6))
    ctx.set_net(net)
    try:
        for _ in range(2):  # the second run is timed warm
            phi, st1 = ctx.poisson_solve(b, np.zeros_like(b),
                                         capi.default_opts(tol=1e-6, learned_init=1))
    finally:
        ctx.set_net(None)
    print(f"plain {st0.sweeps} sweeps {st0.elapsed_ms:.1f} ms; learned {st1.sweeps} sweeps "
          f"{st1.learned_applied} iterations {st1.elapsed_ms:.1f} ms")
    assert st0.rel_residual <= 1e-6 and st1.rel_residual <= 1e-6
    assert st1.learned_rejected == 0
    assert st1.sweeps < 0.1 * st0.sweeps
    assert st1.elapsed_ms < 0.5 * st0.elapsed_ms


def test_graph_replay_matches_eager_launches(ctx):
    """The captured FCG iteration, replayed across iterations and solves with
    eager work and synchronisations in between, gives bit-identical iterates
    and statistics to launching the same kernels one by one."""
    import os
    w = os.path.join(os.path.dirname(capi.LIB_PATH), "weights", "learned_grid2048.npz")
    if not os.path.exists(w):
        pytest.skip("no shipped weights")
    net = learned.load_net(ctx, w)
    rhs = []
    for seed in (11, 12):
        img, _ = synth.target_for("mixture", seed, 512)
        rhs.append(img / img.mean() - 1.0)
    o = capi.default_opts(tol=1e-6, learned_init=1)
    ctx.set_net(net)
    try:
        ctx.set_graphs(False)
        eager = [ctx.poisson_solve(b, np.zeros_like(b), o) for b in rhs]
        ctx.set_graphs(True)
        graph = [ctx.poisson_solve(b, np.zeros_like(b), o) for b in rhs + rhs[:1]]
    finally:
        ctx.set_graphs(True)
        ctx.set_net(None)
    for (pe, se), (pg, sg) in zip(eager + eager[:1], graph):
        assert se.learned_applied > 2 and se.rel_residual <= 1e-6
        assert np.array_equal(pe, pg)
        assert (se.sweeps, se.learned_applied, se.rel_residual) == \
            (sg.sweeps, sg.learned_applied, sg.rel_residual)
