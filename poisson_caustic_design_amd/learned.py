"""Training and use of the learned initial correction (cd_net).

Pairs come from the plain solver itself: designs over seeded synthetic
densities run with the FP32/FP64 red-black SOR solver while the context
captures, at every convergence check of every Poisson solve (transport
potential and height), the normalised coarse residual and, once the solve has
finished, the normalised coarse difference between the converged solution and
the iterate at that check.  The first check of a cold solve gives exactly the
(density mismatch, converged solution) pair.

Training runs on the GPU through the library's own BF16-MFMA forward/backward
kernels and Adam (cd_net_train_step).  Device buffers for the pairs are torch
tensors (memory plumbing only).

    python -m poisson_caustic_design_amd.learned --mesh-res 512 --designs 200 \
        --out poisson_caustic_design_amd/weights/learned_2048.npz
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
                   max_ot: int = 50, log=print):
    """Capture (x, t) pairs from plain-solver designs; returns torch tensors."""
    import torch

    n = 4 * mesh_res
    hc = wc = n // coarsen
    x = torch.zeros((capacity, hc, wc), dtype=torch.float32, device="cuda:0")
    t = torch.zeros_like(x)
    torch.cuda.synchronize()
    ctx.set_capture(x.data_ptr(), t.data_ptr(), capacity, hc, wc, coarsen, every)
    opts = capi.default_opts(tol=tol)
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
        rate = lr * 0.5 * (1 + np.cos(np.pi * step / steps))  # cosine schedule
        losses.append(net.train_step(xb.data_ptr(), tb.data_ptr(), batch, hc, wc, rate))
        if log and (step % log_every == 0 or step == steps - 1):
            log(f"train: step {step} loss {np.mean(losses[-log_every:]):.4e} "
                f"lr {rate:.2e} {time.time() - t0:.1f}s")
    return np.array(losses)


def save_params(net: capi.Net, path: str, meta: dict):
    a = net.arch
    np.savez_compressed(path, params=net.get_params(), levels=a.levels,
                        channels=np.array(list(a.channels)), coarsen=a.coarsen,
                        smooth_sweeps=a.smooth_sweeps, smooth_omega=a.smooth_omega,
                        meta=json.dumps(meta))


def load_net(ctx: capi.Context, path: str) -> capi.Net:
    d = np.load(path, allow_pickle=False)
    arch = capi.default_arch(levels=int(d["levels"]), channels=[int(c) for c in d["channels"]],
                             coarsen=int(d["coarsen"]), smooth_sweeps=int(d["smooth_sweeps"]),
                             smooth_omega=float(d["smooth_omega"]))
    net = capi.Net(ctx, arch)
    net.set_params(d["params"])
    return net


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--mesh-res", type=int, default=256)
    ap.add_argument("--designs", type=int, default=200)
    ap.add_argument("--seed0", type=int, default=1000)
    ap.add_argument("--config", default="mixture")
    ap.add_argument("--capacity", type=int, default=6000)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)

    def log(*m):
        print(*m, file=sys.stderr, flush=True)

    ctx = capi.Context(0)
    x, t = generate_pairs(ctx, range(a.seed0, a.seed0 + a.designs), a.mesh_res, a.config,
                          a.capacity, log=log)
    log(f"pairs: {x.shape[0]}")
    net = capi.Net(ctx)
    net.init_params(a.seed0)
    losses = train(net, x, t, a.steps, a.batch, a.lr, log=log)
    out = a.out or os.path.join(WEIGHTS_DIR, f"learned_grid{4 * a.mesh_res}.npz")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    save_params(net, out, {"mesh_res": a.mesh_res, "designs": a.designs, "seed0": a.seed0,
                           "config": a.config, "pairs": int(x.shape[0]), "steps": a.steps,
                           "final_loss": float(np.mean(losses[-100:]))})
    print(json.dumps({"out": out, "pairs": int(x.shape[0]),
                      "loss_first": float(np.mean(losses[:50])),
                      "loss_last": float(np.mean(losses[-100:]))}), flush=True)


if __name__ == "__main__":
    main()
