"""Training and use of the learned coarse solve (cd_net).

The network approximates the inverse of the coarse Neumann operator; solves
with learned_init use it as the coarse solve of a multigrid-like
preconditioner (optionally through an intermediate grid, mid_sweeps) of
flexible conjugate gradients.  It is trained on the operator loss
mean((cs * Lc net(x) - x)^2), which needs inputs only:

* coarse residuals captured from the solver itself: designs over seeded
  synthetic densities run while the context records, at every convergence
  check of every Poisson solve, the normalised coarse residual (with plain SOR
  first, then again with the network in use, so that the inputs it meets in
  its own iterations are covered);
* random fields of several smoothness scales and white noise, so that the
  operator is learned on every frequency.

Training runs on the GPU through the library's own BF16-MFMA forward/backward
kernels and Adam (cd_net_train_step_operator).  Device buffers for the inputs
are torch tensors (memory plumbing only).  Import torch before the first
capi.Context of the process: torch ships its own HIP runtime with the same
soname, and it must be the one loaded first or torch finds no device.
Loading and using trained weights (load_net) needs no torch.

    python -m poisson_caustic_design_amd.learned --mesh-res 512 \
        --out poisson_caustic_design_amd/weights/learned_grid2048.npz
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import capi, synth

WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")


def design_for(ctx: capi.Context, img: np.ndarray, prm: dict, opts: capi.SolveOpts):
    cd = capi.CausticDesign(ctx)
    cd.set_mesh_resolution(prm["mesh_res"], prm["mesh_res"])
    cd.set_domain_resolution(prm["res"], prm["res"])
    cd.set_mesh_size(prm["width"], prm["height"])
    cd.set_lens_focal_length(prm["focal_l"])
    cd.set_lens_thickness(0.1)
    cd.set_solve_opts(opts, opts)
    cd.initialize_solvers(img)
    return cd


def generate_pairs(ctx: capi.Context, seeds, mesh_res: int, config: str = "mixture",
                   capacity: int = 4096, coarsen: int = 4, every: int = 1, tol: float = 1e-6,
                   max_ot: int = 50, log=print, learned_init: int = 0):
    """Capture (x, t) pairs from plain-solver designs; returns torch tensors."""
    import torch

    n = 4 * mesh_res
    hc = wc = n // coarsen
    x = torch.zeros((capacity, hc, wc), dtype=torch.float32, device="cuda:0")
    t = torch.zeros_like(x)
    torch.cuda.synchronize()
    ctx.set_capture(x.data_ptr(), t.data_ptr(), capacity, hc, wc, coarsen, every)
    opts = capi.default_opts(tol=tol, learned_init=learned_init)
    t0 = time.time()
    try:
        for i, seed in enumerate(seeds):
            img, prm = synth.target_for(config, int(seed), mesh_res)
            cd = design_for(ctx, img, prm, opts)
            rs = cd.run(max_ot, float(np.float32(0.01)), 3, allow=(capi.OK, capi.RASTER_MISS))
            cd.close()
            if log and i % 20 == 0:
                log(f"pairs: design {i} seed {seed}: {rs.transport_steps} OT steps, "
                    f"{ctx.capture_count()} pairs, {time.time() - t0:.1f}s")
            if ctx.capture_count() >= capacity:
                break
    finally:
        count = ctx.capture_count()
        ctx.clear_capture()
    return x[:count], t[:count]


def random_fields(count: int, h: int, w: int, seed: int = 0) -> np.ndarray:
    """Mean-free unit-RMS random fields: Gaussian-filtered noise over scales
    0.5 .. 32 coarse cells, and white noise."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    scales = [0.0, 0.5, 1.0, 2.0, 4.0, 8.0, 16.0, 32.0]
    out = np.empty((count, h, w), dtype=np.float32)
    for i in range(count):
        g = rng.standard_normal((h, w))
        sc = scales[i % len(scales)]
        if sc > 0:
            g = gaussian_filter(g, sc, mode="reflect")
        g -= g.mean()
        out[i] = g / max(np.sqrt(np.mean(g * g)), 1e-30)
    return out


def capture_inputs(ctx: capi.Context, seeds, mesh_res: int, config: str = "mixture",
                   capacity: int = 4096, coarsen: int = 4, net: capi.Net = None,
                   tol: float = 1e-6, max_ot: int = 50, log=print, every: int = 1):
    """Coarse residuals met by solves of seeded designs (with `net` in use when
    given), one per `every` residual checks; returns a torch tensor
    [count][h][w] on the device."""
    if net is not None:
        ctx.set_net(net)
    try:
        x, _ = generate_pairs(ctx, seeds, mesh_res, config, capacity, coarsen, every, tol,
                              max_ot, log=log, learned_init=1 if net is not None else 0)
    finally:
        if net is not None:
            ctx.set_net(None)
    return x


def train_operator(net: capi.Net, x, steps: int = 3000, batch: int = 16, lr: float = 2e-3,
                   seed: int = 0, log=print, log_every: int = 500):
    """Adam on the operator loss over random mini-batches of the inputs x
    (torch tensor [n][h][w] on the device); returns the per-step losses."""
    import torch

    g = torch.Generator(device="cpu").manual_seed(seed)
    n, hc, wc = x.shape
    losses = []
    t0 = time.time()
    for step in range(steps):
        idx = torch.randint(0, n, (batch,), generator=g).to(x.device)
        xb = x.index_select(0, idx)
        # flips keep the operator's symmetries in the data
        if step % 2:
            xb = xb.flip(2)
        if step % 4 >= 2:
            xb = xb.flip(1)
        xb = xb.contiguous()
        # the batch is written on torch's stream, read on the library's: wait
        torch.cuda.synchronize()
        rate = lr * 0.5 * (1 + np.cos(np.pi * step / steps))  # cosine schedule
        losses.append(net.train_step_operator(xb.data_ptr(), batch, hc, wc, rate))
        if log and (step % log_every == 0 or step == steps - 1):
            log(f"train: step {step} loss {np.mean(losses[-log_every:]):.4e} "
                f"lr {rate:.2e} {time.time() - t0:.1f}s")
    return np.array(losses)


def train(net: capi.Net, x, t, steps: int = 2000, batch: int = 8, lr: float = 1e-3,
          seed: int = 0, log=print, log_every: int = 200):
    """Adam on random mini-batches; returns the per-step losses."""
    import torch

    g = torch.Generator(device="cpu").manual_seed(seed)
    n, hc, wc = x.shape
    losses = []
    t0 = time.time()
    for step in range(steps):
        idx = torch.randint(0, n, (batch,), generator=g).to(x.device)
        xb = x.index_select(0, idx).contiguous()
        tb = t.index_select(0, idx).contiguous()
        # random flips keep the operator's symmetries in the data
        if step % 2:
            xb, tb = xb.flip(2).contiguous(), tb.flip(2).contiguous()
        if step % 4 >= 2:
            xb, tb = xb.flip(1).contiguous(), tb.flip(1).contiguous()
        torch.cuda.synchronize()  # written on torch's stream, read on the library's
        rate = lr * 0.5 * (1 + np.cos(np.pi * step / steps))  # cosine schedule
        losses.append(net.train_step(xb.data_ptr(), tb.data_ptr(), batch, hc, wc, rate))
        if log and (step % log_every == 0 or step == steps - 1):
            log(f"train: step {step} loss {np.mean(losses[-log_every:]):.4e} "
                f"lr {rate:.2e} {time.time() - t0:.1f}s")
    return np.array(losses)


def fit(ctx: capi.Context, arch: capi.NetArch, mesh_res: int, seeds, config: str = "mixture",
        random_count: int = 1000, capacity: int = 4096, steps: int = 3000, batch: int = 16,
        lr: float = 2e-3, rounds: int = 1, seed: int = 0, log=print, every: int = 1):
    """Train a network for solves of (4*mesh_res)^2 grids: solver residuals of
    plain-SOR designs plus random fields, then `rounds` more passes that add
    the residuals met with the network in use.  Returns (net, losses, inputs)."""
    import torch

    n = 4 * mesh_res
    hc = wc = n // arch.coarsen
    xs = [capture_inputs(ctx, seeds, mesh_res, config, capacity, arch.coarsen, log=log,
                         every=every)]
    xs.append(torch.from_numpy(random_fields(random_count, hc, wc, seed)).to(xs[0].device))
    pool = torch.cat(xs)
    if log:
        log(f"inputs: {pool.shape[0]} ({xs[0].shape[0]} from solves)")
    net = capi.Net(ctx, arch)
    net.init_params(seed)
    losses = [train_operator(net, pool, steps, batch, lr, seed, log=log)]
    for r in range(rounds):
        more = capture_inputs(ctx, seeds, mesh_res, config, capacity, arch.coarsen, net=net,
                              log=log, every=every)
        pool = torch.cat([pool, more])
        if log:
            log(f"round {r + 1}: +{more.shape[0]} inputs met with the network in use")
        losses.append(train_operator(net, pool, steps // 2, batch, lr / 2, seed + 1 + r, log=log))
    return net, np.concatenate(losses), pool


def design_errors(f, ref, width: float) -> dict:
    """Final-design differences of f from ref: vertex positions / width,
    vertex heights / max height, phi and h mean-free (relative max)."""
    mf = lambda g: g - g.mean()
    return {"xy": float(np.abs(f["tp"][:, :2] - ref["tp"][:, :2]).max() / width),
            "z": float(np.abs(f["sp"][:, 2] - ref["sp"][:, 2]).max() /
                       max(np.abs(ref["sp"][:, 2]).max(), 1e-300)),
            "phi": float(np.abs(mf(f["phi"]) - mf(ref["phi"])).max() /
                         max(np.abs(mf(ref["phi"])).max(), 1e-300)),
            "h": float(np.abs(mf(f["h"]) - mf(ref["h"])).max() /
                       max(np.abs(mf(ref["h"])).max(), 1e-300))}


def run_design(ctx: capi.Context, img, prm, opts: capi.SolveOpts, net=None, max_ot: int = 50):
    """One complete design (main.cpp:243-262 loop); its run statistics and
    final fields."""
    ctx.set_net(net)
    try:
        cd = design_for(ctx, img, prm, opts)
        rs = cd.run(max_ot, float(np.float32(0.01)), 3,
                    allow=(capi.OK, capi.RASTER_MISS, capi.NOT_CONVERGED))
        out = {"tp": cd.get("target_points"), "sp": cd.get("source_points"),
               "phi": cd.get("phi"), "h": cd.get("h")}
        cd.close()
    finally:
        ctx.set_net(None)
    return rs, out


def evaluate(ctx: capi.Context, net: capi.Net, config: str, seeds, mesh_res: int,
             tol: float = 1e-6, log=print, control: bool = False) -> dict:
    """Full designs on held-out seeds, plain SOR and learned form on the same
    GPU at the same residual tolerance: per-design times, sweeps, iterations,
    residuals, the inverted triangles each form's final mesh holds, and the
    final-design difference between the two forms."""
    rows = []
    conv = float(np.float32(0.01))
    for i, seed in enumerate(seeds):
        img, prm = synth.target_for(config, int(seed), mesh_res)
        rec = {"seed": int(seed)}
        final = {}
        for use in (0, 1, 2):
            if use == 2 and not control:
                continue
            ctx.set_net(net if use == 1 else None)
            # use 2: plain SOR at a 2x tighter tolerance, the sensitivity of the
            # design loop itself to solves that differ within the tolerance
            cd = design_for(ctx, img, prm, capi.default_opts(tol=tol / 2 if use == 2 else tol,
                                                             learned_init=1 if use == 1 else 0))
            rs = cd.run(50, conv, 3, allow=(capi.OK, capi.RASTER_MISS))
            key = ("plain", "learned", "plain_tight")[use]
            rec[key] = {"ms": rs.elapsed_ms, "status": rs.status, "ot_steps": rs.transport_steps,
                        "sweeps": rs.phi_sweeps + rs.h_sweeps, "iterations": rs.learned_applied,
                        "fallbacks": rs.learned_rejected,
                        "max_residual": max(rs.max_phi_residual, rs.max_h_residual),
                        "unconverged_solves": int(rs.unconverged_solves),
                        "folded_triangles": int(rs.folded_triangles),
                        "folded_steps": int(rs.folded_steps),
                        "first_folded_step": int(rs.first_folded_step)}
            final[use] = (cd.get("target_points"), cd.get("source_points"))
            cd.close()
        ctx.set_net(None)
        W = prm["width"]
        zr = max(np.abs(final[0][1][:, 2]).max(), 1e-300)
        for k, lbl in ((1, ""), (2, "control_")):
            if k in final:
                rec[lbl + "max_xy_diff_over_width"] = float(
                    np.abs(final[k][0][:, :2] - final[0][0][:, :2]).max() / W)
                rec[lbl + "max_height_diff_rel"] = float(
                    np.abs(final[k][1][:, 2] - final[0][1][:, 2]).max() / zr)
        rows.append(rec)
        if log and i % 25 == 0:
            log(f"eval {i}: seed {seed} plain {rec['plain']['ms']:.1f} ms learned "
                f"{rec['learned']['ms']:.1f} ms ({rec['learned']['iterations']} it)")
    ok = [r for r in rows if r["plain"]["status"] == capi.OK and r["learned"]["status"] == capi.OK]
    tp = sum(r["plain"]["ms"] for r in ok)
    tl = sum(r["learned"]["ms"] for r in ok)
    return {
        "designs": len(rows), "both_ok": len(ok),
        "plain_ms_mean": tp / max(len(ok), 1), "learned_ms_mean": tl / max(len(ok), 1),
        "speedup_total_time": tp / max(tl, 1e-30),
        "speedup_median": float(np.median([r["plain"]["ms"] / r["learned"]["ms"] for r in ok])) if ok else None,
        "learned_iterations_mean": float(np.mean([r["learned"]["iterations"] for r in ok])) if ok else None,
        "fallbacks": int(sum(r["learned"]["fallbacks"] for r in ok)),
        "unconverged_solves": {"plain": int(sum(r["plain"]["unconverged_solves"] for r in ok)),
                               "learned": int(sum(r["learned"]["unconverged_solves"] for r in ok))},
        "worst_residual_plain": max((r["plain"]["max_residual"] for r in ok), default=None),
        "worst_residual_learned": max((r["learned"]["max_residual"] for r in ok), default=None),
        "max_xy_diff_over_width": max((r["max_xy_diff_over_width"] for r in ok), default=None),
        "max_height_diff_rel": max((r["max_height_diff_rel"] for r in ok), default=None),
        "median_xy_diff_over_width": float(np.median([r["max_xy_diff_over_width"] for r in ok]))
        if ok else None,
        "outliers": outlier_summary(ok),
        "rows": rows,
    }


def outlier_summary(rows, threshold: float = 1e-3) -> dict:
    """Designs whose two forms end more than `threshold` widths apart, and
    which of them hold inverted triangles in the plain form, the learned form
    or both (a folded mesh is a degenerate design: the reference steps interior
    vertices unclamped, mesh.cpp:485-532)."""
    out = [r for r in rows if r["max_xy_diff_over_width"] > threshold]
    fp = lambda r: r["plain"]["folded_triangles"] > 0 or r["plain"]["folded_steps"] > 0
    fl = lambda r: r["learned"]["folded_triangles"] > 0 or r["learned"]["folded_steps"] > 0
    res = {"threshold_width": threshold, "count": len(out),
           "folded_both": sum(1 for r in out if fp(r) and fl(r)),
           "folded_plain_only": sum(1 for r in out if fp(r) and not fl(r)),
           "folded_learned_only": sum(1 for r in out if fl(r) and not fp(r)),
           "unfolded": sum(1 for r in out if not fp(r) and not fl(r)),
           "folded_designs_all": {"plain": sum(1 for r in rows if fp(r)),
                                  "learned": sum(1 for r in rows if fl(r))},
           "seeds": [r["seed"] for r in out]}
    if "control_max_xy_diff_over_width" in (out[0] if out else {}):
        res["within_10x_control"] = sum(
            1 for r in out
            if r["max_xy_diff_over_width"] <= 10 * r["control_max_xy_diff_over_width"])
    return res


def equal_error(ctx: capi.Context, net: capi.Net, config: str, seeds, mesh_res: int,
                ref_tol: float, tols=(1e-6, 1e-7, 1e-8, 1e-9), max_ot: int = 50,
                log=print) -> dict:
    """Plain SOR at 1e-6 and the learned form at each of `tols`, against the
    FP64 reference design (FP64 sweeps to relative residual ref_tol): per
    design the errors, times and the tightest-needed learned tolerance at which
    every error (positions, heights, phi, h) is at most plain's."""
    rows = []
    for seed in seeds:
        img, prm = synth.target_for(config, int(seed), mesh_res)
        W = prm["width"]
        rr, ref = run_design(ctx, img, prm, capi.default_opts(tol=ref_tol,
                                                              precision=capi.PREC_FP64),
                             max_ot=max_ot)
        rp, plain = run_design(ctx, img, prm, capi.default_opts(tol=1e-6), max_ot=max_ot)
        ep = design_errors(plain, ref, W)
        rec = {"seed": int(seed), "reference": {"ms": rr.elapsed_ms, "ot_steps": rr.transport_steps,
                                                "status": rr.status,
                                                "max_residual": max(rr.max_phi_residual,
                                                                    rr.max_h_residual)},
               "plain": {"ms": rp.elapsed_ms, "ot_steps": rp.transport_steps, "errors": ep},
               "learned": {}}
        chosen = None
        for t in tols:
            rl, lf = run_design(ctx, img, prm, capi.default_opts(tol=t, learned_init=1), net,
                                max_ot=max_ot)
            el = design_errors(lf, ref, W)
            rec["learned"][f"{t:g}"] = {"ms": rl.elapsed_ms, "ot_steps": rl.transport_steps,
                                        "iterations": rl.learned_applied,
                                        "fallbacks": rl.learned_rejected, "errors": el}
            if chosen is None and all(el[k] <= ep[k] for k in ep):
                chosen = t
        rec["equal_error_tol"] = chosen
        if chosen is not None:
            rec["speedup_at_equal_error"] = rp.elapsed_ms / rec["learned"][f"{chosen:g}"]["ms"]
        rows.append(rec)
        if log:
            log(f"equal error seed {seed}: plain {ep}, chosen learned tol {chosen}")
    return {"config": config, "mesh_res": mesh_res, "reference_tol": ref_tol,
            "max_ot_steps": max_ot, "rows": rows}


def save_params(net: capi.Net, path: str, meta: dict):
    a = net.arch
    np.savez_compressed(path, params=net.get_params(), levels=a.levels,
                        channels=np.array(list(a.channels)), coarsen=a.coarsen,
                        smooth_sweeps=a.smooth_sweeps, smooth_omega=a.smooth_omega,
                        linear=a.linear, out_smooth=a.out_smooth, loss_smooth=a.loss_smooth,
                        loss_energy=a.loss_energy, mid_sweeps=a.mid_sweeps,
                        meta=json.dumps(meta))


def load_net(ctx: capi.Context, path: str) -> capi.Net:
    d = np.load(path, allow_pickle=False)
    arch = capi.default_arch(levels=int(d["levels"]), channels=[int(c) for c in d["channels"]],
                             coarsen=int(d["coarsen"]), smooth_sweeps=int(d["smooth_sweeps"]),
                             smooth_omega=float(d["smooth_omega"]), linear=int(d["linear"]),
                             out_smooth=int(d["out_smooth"]), loss_smooth=int(d["loss_smooth"]),
                             loss_energy=int(d["loss_energy"]) if "loss_energy" in d.files else 0,
                             mid_sweeps=int(d["mid_sweeps"]) if "mid_sweeps" in d.files else 0)
    net = capi.Net(ctx, arch)
    net.set_params(d["params"])
    return net


def main(argv=None):
    import torch  # noqa: F401  (before the first Context: see the module docstring)

    ap = argparse.ArgumentParser()
    ap.add_argument("--mesh-res", type=int, default=512)
    ap.add_argument("--designs", type=int, default=16)
    ap.add_argument("--seed0", type=int, default=1000)
    ap.add_argument("--config", default="mixture")
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--channels", default="16,32,32,32")
    ap.add_argument("--relu", action="store_true")
    ap.add_argument("--out-smooth", type=int, default=6)
    ap.add_argument("--loss-smooth", type=int, default=0)
    ap.add_argument("--loss", choices=["operator", "energy"], default="operator",
                    help="operator: mean((cs Lc y - x)^2); energy: y'(-cs Lc)y/2 + x'y")
    ap.add_argument("--coarsen", type=int, default=16)
    ap.add_argument("--sweeps", type=int, default=8, help="smoothing sweeps per FCG iteration")
    ap.add_argument("--omega", type=float, default=1.5)
    ap.add_argument("--mid-sweeps", type=int, default=0,
                    help="smoothing sweeps on the intermediate grid (coarsen/4 cells per cell)")
    ap.add_argument("--random", type=int, default=1200)
    ap.add_argument("--capacity", type=int, default=4000)
    ap.add_argument("--capture-every", type=int, default=1,
                    help="keep the coarse residual of one in this many checks")
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--rounds", type=int, default=2)
    ap.add_argument("--out", default=None)
    ap.add_argument("--evaluate", default=None, metavar="FIRST:COUNT",
                    help="evaluate --weights (or the trained net) on COUNT held-out seeds")
    ap.add_argument("--weights", default=None)
    ap.add_argument("--eval-out", default=None)
    ap.add_argument("--control", action="store_true",
                    help="also run plain SOR at half the tolerance (design-loop sensitivity)")
    ap.add_argument("--equal-error", default=None, metavar="FIRST:COUNT",
                    help="with --weights: plain at 1e-6 and learned at 1e-6..1e-9 against the "
                         "FP64 reference design (--ref-tol) on COUNT seeds")
    ap.add_argument("--ref-tol", type=float, default=2e-9)
    ap.add_argument("--max-ot", type=int, default=50,
                    help="transport steps of the equal-error designs (the reference loop: 50)")
    a = ap.parse_args(argv)

    def log(*m):
        print(*m, file=sys.stderr, flush=True)

    ctx = capi.Context(0)
    if a.weights and a.equal_error:
        first, count = (int(v) for v in a.equal_error.split(":"))
        net = load_net(ctx, a.weights)
        res = equal_error(ctx, net, a.config, range(first, first + count), a.mesh_res,
                          a.ref_tol, max_ot=a.max_ot, log=log)
        res.update(weights=a.weights)
        if a.eval_out:
            with open(a.eval_out, "w") as f:
                json.dump(res, f, indent=1)
        print(json.dumps(res), flush=True)
        return
    if a.weights and a.evaluate:
        first, count = (int(v) for v in a.evaluate.split(":"))
        net = load_net(ctx, a.weights)
        res = evaluate(ctx, net, a.config, range(first, first + count), a.mesh_res, log=log,
                       control=a.control)
        res.update(config=a.config, mesh_res=a.mesh_res, weights=a.weights,
                   seeds=f"{first}..{first + count - 1}")
        if a.eval_out:
            with open(a.eval_out, "w") as f:
                json.dump(res, f, indent=1)
        print(json.dumps({k: v for k, v in res.items() if k != "rows"}), flush=True)
        return
    arch = capi.default_arch(levels=a.levels, channels=[int(c) for c in a.channels.split(",")],
                             linear=0 if a.relu else 1, out_smooth=a.out_smooth,
                             loss_smooth=a.loss_smooth, coarsen=a.coarsen,
                             loss_energy=int(a.loss == "energy"), mid_sweeps=a.mid_sweeps,
                             smooth_sweeps=a.sweeps, smooth_omega=a.omega)
    t0 = time.time()
    net, losses, pool = fit(ctx, arch, a.mesh_res, range(a.seed0, a.seed0 + a.designs), a.config,
                            a.random, a.capacity, a.steps, a.batch, a.lr, a.rounds, a.seed0,
                            log=log, every=a.capture_every)
    out = a.out or os.path.join(WEIGHTS_DIR, f"learned_grid{4 * a.mesh_res}.npz")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    meta = {"mesh_res": a.mesh_res, "designs": a.designs, "seed0": a.seed0, "config": a.config,
            "inputs": int(pool.shape[0]), "capture_every": a.capture_every,
            "steps": a.steps, "rounds": a.rounds,
            "loss": a.loss, "final_loss": float(np.mean(losses[-100:])),
            "train_seconds": round(time.time() - t0, 1)}
    save_params(net, out, meta)
    print(json.dumps({"out": out, **meta, "loss_first": float(np.mean(losses[:50]))}), flush=True)


if __name__ == "__main__":
    main()
