"""Benchmark: full caustic designs (OT + height) per second on one MI355X.

One step = one complete design from the device-resident target image:
initialize_solvers (min-max, mesh, target-area clipping), up to 50 transport
(OT) steps with the reference's stop rule on the step size (main.cpp:243-256),
then 3 height steps (main.cpp:260-262).  Every inner Poisson solve stops at
FP64 relative residual ||P b - A phi||_2 / ||P b||_2 <= 1e-6.

Workload (default): 512x512 mesh, 2048x2048 Poisson grid, seeded 3-octave
value-noise target (1024x1024 f32, nearest-resized to 2048x2048), W = 0.5,
focal length 1.5.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1]

Prints one JSON line.  The CPU baseline (reference code on the host cores,
bounded sample, extrapolated with the GPU run's own step and sweep counts) runs
in a subprocess after the GPU measurement.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "OT+height solves/sec at 512^2 @1e-6 residual; learned-init vs plain, same GPU"
HBM_PEAK_GBS = 8000.0
SWEEP_BYTES_PER_CELL = 12.0  # one red+black FP32 sweep: e read + e write + rhs read
# Compulsory HBM bytes of one temporally blocked launch (several sweeps fused):
# the FP32 correction read once and written once, the BF16 right-hand side read once.
LAUNCH_BYTES_PER_CELL = 10.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(config: str, seed: int, per_design: dict, timeout: int) -> dict | None:
    threads = min(os.cpu_count() or 1, 16)
    cmd = [sys.executable, "-m", "oracle.cpu_baseline", "--config", config, "--seed", str(seed),
           "--threads", str(threads)]
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, cwd=ROOT)
    except subprocess.TimeoutExpired:
        log("cpu baseline: timed out")
        return None
    if r.returncode != 0:
        log("cpu baseline failed:", r.stderr[-2000:])
        return None
    m = json.loads(r.stdout.strip().splitlines()[-1])
    t_design = (m["t_init"] + per_design["ot_steps"] * m["t_ot_ns"]
                + per_design["h_steps"] * m["t_h_ns"] + per_design["sweeps"] * m["t_sweep"])
    sample = (f"{m['kind']} code built {m['flags']}, {m['tiles']} SOR tile threads on "
              f"{m['nproc']} host cpus ({m['cpu_model']}): measured initialize_solvers ({m['t_init']:.2f}s), one transport "
              f"step without its solve ({m['t_ot_ns']:.2f}s), one height step without its solve "
              f"({m['t_h_ns']:.2f}s) and {m['sweeps_timed']} SOR sweeps ({m['t_sweep']*1e3:.2f} ms/"
              f"sweep) at {m['grid']}^2; extrapolated to one design with the GPU run's per-design "
              f"counts ({per_design['ot_steps']:.1f} transport steps, {per_design['h_steps']:.0f} "
              f"height steps, {per_design['sweeps']:.0f} sweeps; the reference's own omega needs "
              f"more sweeps, so this favours the CPU)")
    return {"value": 1.0 / t_design, "unit": "solves/s", "cores": m["tiles"],
            "kind": m["kind"], "sample": sample, "seconds_per_design": t_design,
            "components": m}


def c1_end_to_end(timeout: int) -> dict | None:
    """C1 end to end on both sides, the reference CLI's flow and defaults
    (main.cpp:224-262: max-update stops 1e-7 / 1e-8, the reference's omega,
    <= 50 transport steps with the break at conv_tres 0.01f, 3 height steps):
    the reference's own code on the host cores, measured in full (not
    extrapolated), and the GPU library with the same settings
    (cd_design_params_default), median of 3 designs."""
    import numpy as np

    from poisson_caustic_design_amd import capi, synth

    threads = min(os.cpu_count() or 1, 16)
    cmd = [sys.executable, "-m", "oracle.cpu_baseline", "--config", "gauss5", "--seed", "0",
           "--threads", str(threads), "--full"]
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, cwd=ROOT)
    except subprocess.TimeoutExpired:
        log("C1 end to end: CPU run timed out")
        return None
    if r.returncode != 0:
        log("C1 end to end: CPU run failed:", r.stderr[-1000:])
        return None
    m = json.loads(r.stdout.strip().splitlines()[-1])
    img, prm = synth.target_for("gauss5", 0)
    ctx = capi.Context(0)
    cd = capi.CausticDesign(ctx)     # the drop-in defaults: the reference's solve settings
    cd.set_mesh_resolution(prm["mesh_res"], prm["mesh_res"])
    cd.set_domain_resolution(prm["res"], prm["res"])
    cd.set_mesh_size(prm["width"], prm["height"])
    cd.set_lens_focal_length(prm["focal_l"])
    cd.set_lens_thickness(float(np.float32(0.1)))
    cd.initialize_solvers(img)
    d_img = cd.device_ptr("image")
    conv = float(np.float32(0.01))
    times, rs = [], None
    for _ in range(4):
        ctx.synchronize()
        t0 = time.perf_counter()
        cd.reinitialize_device(d_img)
        rs = cd.run(50, conv, 3)
        ctx.synchronize()
        times.append(time.perf_counter() - t0)
    cd.close()
    t_gpu = float(np.median(times[1:]))
    log(f"C1 end to end: CPU {m['t_design']:.1f} s ({m['ot_steps']} OT steps), GPU "
        f"{t_gpu * 1e3:.1f} ms ({rs.transport_steps} OT steps)")
    return {"workload": "C1: 100x100 mesh, 400x400 grid, gauss5 seed 0, the reference CLI's "
                        "solver defaults (max-update 1e-7 / 1e-8, reference omega)",
            "cpu_seconds": m["t_design"], "cpu_ot_steps": m["ot_steps"],
            "cpu": f"reference code built {m['flags']}, {m['tiles']} SOR tile threads on "
                   f"{m['nproc']} host cpus ({m['cpu_model']}), measured end to end",
            "gpu_seconds": t_gpu, "gpu_ot_steps": rs.transport_steps,
            "gpu_sweeps": rs.phi_sweeps + rs.h_sweeps, "speedup": m["t_design"] / t_gpu}


PROFILE_SUMMARY = os.path.join(ROOT, "profiles", "k_rb_blocked_design_512mesh_2048grid.md")


def profiled_traffic(grid: int):
    """HBM bytes per k_rb_blocked launch at 2048^2 from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (counter runs cannot happen inside the timed
    benchmark); None for other grids or when the summary is absent."""
    if grid != 2048 or not os.path.exists(PROFILE_SUMMARY):
        return None
    for line in open(PROFILE_SUMMARY):
        if line.startswith("* traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def fixed_sweeps(a) -> int:
    """Exactly a.sweeps red-black sweeps (mixed precision) on a seeded smooth
    right-hand side of the workload grid size, device-resident."""
    import numpy as np
    import torch

    from poisson_caustic_design_amd import capi, synth

    _, prm = synth.target_for(a.config, a.seed, a.mesh_res)
    n = prm["res"]
    rng = np.random.default_rng(a.seed)
    yy, xx = np.mgrid[0:n, 0:n] / n
    b = np.cos(3 * np.pi * xx) * np.cos(2 * np.pi * yy) + 0.1 * rng.standard_normal((n, n))
    b -= b.mean()
    ctx = capi.Context(0)
    d_b = torch.tensor(b, dtype=torch.float64, device="cuda:0")
    d_phi = torch.zeros_like(d_b)
    torch.cuda.synchronize()
    # residual rule with tol 0 never stops early: exactly a.sweeps sweeps of the
    # production (non-tracking) kernel, one FP64 check at the end; the first
    # solve warms up, the last `steps` are timed
    o = capi.default_opts(max_sweeps=a.sweeps, stop=capi.STOP_RESIDUAL, tol=0.0,
                          check_every=a.sweeps)
    best = None
    for i in range(1 + max(a.steps, 1)):
        d_phi.zero_()
        torch.cuda.synchronize()
        ctx.reset_kernel_stats()
        ctx.set_profiling(i > 0)
        st = ctx.poisson_solve_device(d_b.data_ptr(), d_phi.data_ptr(), n, n, o)
        if i == 0:
            continue
        ms, launches, nbytes = ctx.kernel_stats(capi.KERNEL_SOR_SWEEP)
        if best is None or ms < best[0]:
            best = (ms, launches, nbytes)
    ms, launches, nbytes = best
    print(json.dumps({"mode": "sweeps", "grid": n,
                      "sweeps": st.sweeps, "sor_ms": ms, "us_per_sweep": 1e3 * ms / st.sweeps,
                      "launches": launches, "avg_launch_us": 1e3 * ms / max(launches, 1),
                      "bytes_per_launch": nbytes / max(launches, 1),
                      "achieved_GBs": nbytes / max(ms, 1e-9) / 1e6}), flush=True)
    return 0


def device_copy(a) -> int:
    """A plain device-to-device copy kernel of a.mib MiB (read + write)."""
    import torch

    n = a.mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    reps = max(a.steps, 1)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        y.copy_(x)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    print(json.dumps({"mode": "copy", "bytes_read": 4 * n, "bytes_written": 4 * n,
                      "ms_per_copy": ms, "GBs": 8 * n / (ms * 1e-3) / 1e9}), flush=True)
    return 0


def learned_leg(ctx, cd, one_design, wpath, a, plain_value) -> dict:
    """The same designs with the learned coarse solve in every Poisson solve
    (same stop rule and tolerance); timed like the plain leg."""
    from poisson_caustic_design_amd import capi, learned

    net = learned.load_net(ctx, wpath)
    ctx.set_net(net)
    o = capi.default_opts(tol=a.tol, learned_init=1)
    cd.set_solve_opts(o, o)
    try:
        for i in range(a.warmup):
            one_design()
        ctx.synchronize()
        runs = []
        t0 = time.perf_counter()
        for i in range(a.steps):
            rs = one_design()
            runs.append(rs.as_dict())
            log(f"learned step {i}: {rs.elapsed_ms:.1f} ms, {rs.transport_steps} OT steps, "
                f"{rs.phi_sweeps}+{rs.h_sweeps} sweeps, {rs.learned_applied} learned "
                f"iterations, {rs.learned_rejected} fallbacks, max residual "
                f"{max(rs.max_phi_residual, rs.max_h_residual):.2e}")
        ctx.synchronize()
        elapsed = time.perf_counter() - t0
    finally:
        ctx.set_net(None)
        o = capi.default_opts(tol=a.tol)
        cd.set_solve_opts(o, o)
        net.close()
    for r in runs:
        if r["status"] != capi.OK:
            raise SystemExit(f"learned design failed: status {r['status']}")
    value = a.steps / elapsed
    n = len(runs)
    return {
        "value": value,
        "ms_per_step": 1e3 * elapsed / a.steps,
        "speedup_vs_plain": value / plain_value if plain_value else None,
        "weights": os.path.relpath(wpath, ROOT),
        "per_design": {
            "ot_steps": sum(r["transport_steps"] for r in runs) / n,
            "sweeps": sum(r["phi_sweeps"] + r["h_sweeps"] for r in runs) / n,
            "learned_iterations": sum(r["learned_applied"] for r in runs) / n,
            "fallbacks": sum(r["learned_rejected"] for r in runs) / n,
            "max_residual": max(max(r["max_phi_residual"], r["max_h_residual"]) for r in runs),
        },
    }


def final_design(ctx, img, prm, prec, tol, net=None):
    """One complete design from a fresh initialisation; returns its final
    fields, run statistics and wall time."""
    import numpy as np

    from poisson_caustic_design_amd import capi

    M, N = prm["mesh_res"], prm["res"]
    ctx.set_net(net)
    try:
        cd = capi.CausticDesign(ctx)
        cd.set_mesh_resolution(M, M)
        cd.set_domain_resolution(N, N)
        cd.set_mesh_size(prm["width"], prm["height"])
        cd.set_lens_focal_length(prm["focal_l"])
        cd.set_lens_thickness(0.1)
        o = capi.default_opts(tol=tol, precision=prec, learned_init=int(net is not None))
        cd.set_solve_opts(o, o)
        cd.initialize_solvers(img)
        ctx.synchronize()
        t0 = time.perf_counter()
        rs = cd.run(50, float(np.float32(0.01)), 3, allow=(capi.OK, capi.NOT_CONVERGED))
        ctx.synchronize()
        el = time.perf_counter() - t0
        out = {k: cd.get(k) for k in ("phi", "h", "target_points", "source_points")}
        cd.close()
    finally:
        ctx.set_net(None)
    return out, rs, el


def design_error(f, ref, width):
    """Final-design differences of f from ref: phi and h mean-free (relative
    max), vertex positions / width, vertex heights / max height."""
    import numpy as np

    mf = lambda g: g - g.mean()
    return {
        "phi": float(np.abs(mf(f["phi"]) - mf(ref["phi"])).max() / np.abs(mf(ref["phi"])).max()),
        "h": float(np.abs(mf(f["h"]) - mf(ref["h"])).max() / np.abs(mf(ref["h"])).max()),
        "target_xy_over_width": float(np.abs(f["target_points"][:, :2] -
                                             ref["target_points"][:, :2]).max() / width),
        "height_over_max_height": float(np.abs(f["source_points"][:, 2] -
                                               ref["source_points"][:, 2]).max() /
                                        np.abs(ref["source_points"][:, 2]).max()),
    }


# FP64 reference of the precision comparisons: FP64 sweeps to this relative
# residual (the FP64 residual of a 2048^2 Neumann problem has a rounding floor
# near (4 N^2 / pi^2) * 2^-53 ~ 2e-10, so 1e-12 is not attainable there)
REF_TOL = 1e-9


def precision_report(a) -> int:
    """Precision form on the benchmark workload: designs with FP32 sweeps
    (FP64 residual, zero-mean projection, defect correction), FP64 sweeps and
    the learned coarse solve against the same design with FP64 sweeps solved
    to REF_TOL -- the FP64 reference: the same red-black iteration whose FP64
    iterates equal the oracle's at test sizes."""
    from poisson_caustic_design_amd import capi, learned, synth

    img, prm = synth.target_for(a.config, a.seed, a.mesh_res)
    M, N, W = prm["mesh_res"], prm["res"], prm["width"]
    ctx = capi.Context(0)
    ref, rref, tref = final_design(ctx, img, prm, capi.PREC_FP64, REF_TOL)
    forms = [("mixed_1e-6", capi.PREC_MIXED, 1e-6, False),
             ("fp64_1e-6", capi.PREC_FP64, 1e-6, False),
             ("mixed_1e-9", capi.PREC_MIXED, REF_TOL, False)]
    wpath = a.weights or os.path.join(ROOT, "poisson_caustic_design_amd", "weights",
                                      f"learned_grid{N}.npz")
    net = learned.load_net(ctx, wpath) if os.path.exists(wpath) else None
    if net is not None:
        forms += [(f"learned_{t:g}", capi.PREC_MIXED, t, True) for t in (1e-6, 1e-7, 1e-8, 1e-9)]
    rows = {}
    for name, prec, tol, use in forms:
        f, rs, el = final_design(ctx, img, prm, prec, tol, net if use else None)
        rows[name] = {"ms": el * 1e3, "status": rs.status, "sweeps": rs.phi_sweeps + rs.h_sweeps,
                      "learned_iterations": rs.learned_applied,
                      "max_residual": max(rs.max_phi_residual, rs.max_h_residual),
                      "vs_reference": design_error(f, ref, W)}
        log(f"precision {name}: {el * 1e3:.1f} ms, {rows[name]['vs_reference']}")
    if net is not None:
        net.close()
    print(json.dumps({"mode": "precision", "mesh": M, "grid": N, "config": a.config,
                      "reference": {"form": f"FP64 sweeps, relative residual {REF_TOL:g}",
                                    "ms": tref * 1e3, "status": rref.status,
                                    "sweeps": rref.phi_sweeps + rref.h_sweeps,
                                    "max_residual": max(rref.max_phi_residual,
                                                        rref.max_h_residual)},
                      "forms": rows}), flush=True)
    return 0


def equal_error_leg(ctx, cd, one_design, wpath, a, plain_value, plain_final, img, prm) -> dict:
    """The learned form at equal final error: both forms' final designs are
    compared with the FP64 reference (FP64 sweeps to REF_TOL); the learned
    form's tolerance is tightened (1e-6, 1e-7, ...) until every one of its
    errors (phi, h, vertex positions, vertex heights) is at most the plain
    form's at 1e-6, then timed like the other legs."""
    from poisson_caustic_design_amd import capi, learned

    W = prm["width"]
    ref, rref, _ = final_design(ctx, img, prm, capi.PREC_FP64, REF_TOL)
    plain_err = design_error(plain_final, ref, W)
    net = learned.load_net(ctx, wpath)
    ctx.set_net(net)
    chosen = err = None
    try:
        for tol in (1e-6, 1e-7, 1e-8, 1e-9):
            o = capi.default_opts(tol=tol, learned_init=1)
            cd.set_solve_opts(o, o)
            rs = one_design()
            f = {k: cd.get(k) for k in ("phi", "h", "target_points", "source_points")}
            err = design_error(f, ref, W)
            log(f"equal error: learned at {tol:g}: {err} (plain at 1e-6: {plain_err})")
            if rs.status == capi.OK and all(err[k] <= plain_err[k] for k in err):
                chosen = tol
                break
        if chosen is None:
            return {"tol": None, "plain_error_vs_fp64": plain_err, "learned_error_vs_fp64": err}
        for i in range(a.warmup):
            one_design()
        ctx.synchronize()
        runs = []
        t0 = time.perf_counter()
        for i in range(a.steps):
            runs.append(one_design())
        ctx.synchronize()
        elapsed = time.perf_counter() - t0
    finally:
        ctx.set_net(None)
        o = capi.default_opts(tol=a.tol)
        cd.set_solve_opts(o, o)
        net.close()
    value = a.steps / elapsed
    return {
        "tol": chosen,
        "value": value,
        "ms_per_step": 1e3 * elapsed / a.steps,
        "speedup_vs_plain": value / plain_value,
        "learned_iterations": sum(r.learned_applied for r in runs) / len(runs),
        "reference": f"FP64 sweeps to relative residual {REF_TOL:g} "
                     f"(max residual {max(rref.max_phi_residual, rref.max_h_residual):.2e})",
        "plain_error_vs_fp64": plain_err,
        "learned_error_vs_fp64": err,
    }


def learned_only(a) -> int:
    """The learned leg alone on the benchmark workload (for kernel traces)."""
    import numpy as np

    from poisson_caustic_design_amd import capi, synth

    img, prm = synth.target_for(a.config, a.seed, a.mesh_res)
    M, N = prm["mesh_res"], prm["res"]
    wpath = a.weights or os.path.join(ROOT, "poisson_caustic_design_amd", "weights",
                                      f"learned_grid{N}.npz")
    ctx = capi.Context(0)
    # graphs off: each launch of the learned iteration goes to the stream on
    # its own (a kernel tracer then times every dispatch)
    ctx.set_graphs(a.graphs == "on")
    cd = capi.CausticDesign(ctx)
    cd.set_mesh_resolution(M, M)
    cd.set_domain_resolution(N, N)
    cd.set_mesh_size(prm["width"], prm["height"])
    cd.set_lens_focal_length(prm["focal_l"])
    cd.set_lens_thickness(0.1)
    cd.initialize_solvers(img)
    d_img = cd.device_ptr("image")
    conv = float(np.float32(0.01))

    def one_design():
        cd.reinitialize_device(d_img)
        return cd.run(50, conv, 3)

    res = learned_leg(ctx, cd, one_design, wpath, a, None)
    print(json.dumps({"mode": "learned", "mesh": M, "grid": N, **res}), flush=True)
    return 0


def batched_leg(img, prm, a, wpath) -> dict:
    """a.batch independent designs of the same target, one context (HIP
    stream) each, run together through cd_designs_run_batch; timed per batch,
    plain and (with weights) learned."""
    from poisson_caustic_design_amd import capi, learned

    M, N = prm["mesh_res"], prm["res"]
    S = a.batch
    conv = float(_np().float32(0.01))
    res = {"designs_per_batch": S,
           "note": (f"{S} designs of the same target in flight at once, one context and HIP "
                    "stream each (cd_designs_run_batch); value = designs/s")}
    for form in ("plain", "learned"):
        if form == "learned" and not (a.learned == "auto" and os.path.exists(wpath)):
            continue
        ctxs = [capi.Context(0) for _ in range(S)]
        nets, designs = [], []
        try:
            for c in ctxs:
                if form == "learned":
                    nets.append(learned.load_net(c, wpath))
                    c.set_net(nets[-1])
                cd = capi.CausticDesign(c)
                cd.set_mesh_resolution(M, M)
                cd.set_domain_resolution(N, N)
                cd.set_mesh_size(prm["width"], prm["height"])
                cd.set_lens_focal_length(prm["focal_l"])
                cd.set_lens_thickness(0.1)
                o = capi.default_opts(tol=a.tol, learned_init=int(form == "learned"))
                cd.set_solve_opts(o, o)
                cd.initialize_solvers(img)
                designs.append(cd)
            imgs = [d.device_ptr("image") for d in designs]
            for _ in range(a.warmup):
                capi.run_designs_batch(designs, imgs, 50, conv, 3)
            for c in ctxs:
                c.synchronize()
            worst = 0.0
            t0 = time.perf_counter()
            for i in range(a.steps):
                stats = capi.run_designs_batch(designs, imgs, 50, conv, 3)
                worst = max([worst] + [max(r.max_phi_residual, r.max_h_residual) for r in stats])
                log(f"batched {form} step {i}: {S} designs, max residual {worst:.2e}")
            for c in ctxs:
                c.synchronize()
            elapsed = time.perf_counter() - t0
        finally:
            for d in designs:
                d.close()
            for n in nets:
                n.close()
            for c in ctxs:
                c.close()
        res[form] = {"value": S * a.steps / elapsed, "ms_per_batch": 1e3 * elapsed / a.steps,
                     "max_residual": worst}
    return res


def _np():
    import numpy as np
    return np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="noise3")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--mesh-res", type=int, default=None)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-timeout", type=int, default=240)
    ap.add_argument("--no-c1", action="store_true",
                    help="skip the C1 end-to-end comparison (reference code on the host "
                         "against the GPU, the reference CLI's settings)")
    ap.add_argument("--mode", choices=["design", "sweeps", "copy", "learned", "precision"],
                    default="design",
                    help="design: the benchmark; sweeps: a fixed number of SOR sweeps at the "
                         "workload grid (profiling); copy: a device copy of known size "
                         "(calibrates rocprofv3 byte counters); learned: only the learned-"
                         "init designs (profiling); precision: the mixed-precision design "
                         "against FP64 sweeps to 1e-9")
    ap.add_argument("--graphs", choices=["on", "off"], default="on",
                    help="--mode learned: replay the learned iteration as a HIP graph")
    ap.add_argument("--learned", choices=["auto", "off"], default="auto",
                    help="auto: when weights for the grid exist, also time the same designs "
                         "with the learned coarse solve and report them beside the plain ones")
    ap.add_argument("--weights", default=None)
    ap.add_argument("--batch", type=int, default=4,
                    help="also time batches of this many concurrent designs (0: skip)")
    ap.add_argument("--sweeps", type=int, default=64)
    ap.add_argument("--mib", type=int, default=256)
    a = ap.parse_args(argv)
    if a.mode == "sweeps":
        return fixed_sweeps(a)
    if a.mode == "copy":
        return device_copy(a)

    if a.mode == "learned":
        return learned_only(a)
    if a.mode == "precision":
        return precision_report(a)

    if a.gpus != 1:
        print(f"bench.py: measures one GPU only (asked for --gpus {a.gpus}); "
              "independent designs shard one per GPU with no communication, so N GPUs give "
              "N x the one-GPU value.", file=sys.stderr)
        return 2

    import numpy as np

    from poisson_caustic_design_amd import capi, synth

    img, prm = synth.target_for(a.config, a.seed, a.mesh_res)
    M, N, W, H, f = prm["mesh_res"], prm["res"], prm["width"], prm["height"], prm["focal_l"]
    log(f"bench: config {a.config} seed {a.seed} mesh {M} grid {N}")

    ctx = capi.Context(0)
    cd = capi.CausticDesign(ctx)
    cd.set_mesh_resolution(M, M)
    cd.set_domain_resolution(N, N)
    cd.set_mesh_size(W, H)
    cd.set_lens_focal_length(f)
    cd.set_lens_thickness(0.1)
    prec = capi.PREC_MIXED if a.precision == "mixed" else capi.PREC_FP64
    o = capi.default_opts(tol=a.tol, precision=prec)
    cd.set_solve_opts(o, o)
    cd.initialize_solvers(img)          # uploads the image; it stays resident
    d_img = cd.device_ptr("image")
    conv = float(np.float32(0.01))      # --conv_tres default 0.01 as float (main.cpp:151,183)

    def one_design():
        cd.reinitialize_device(d_img)
        return cd.run(50, conv, 3)

    for i in range(a.warmup):
        rs = one_design()
        log(f"warmup {i}: {rs.elapsed_ms:.1f} ms, {rs.transport_steps} OT steps, "
            f"{rs.phi_sweeps}+{rs.h_sweeps} sweeps")

    ctx.synchronize()
    ctx.reset_kernel_stats()
    ctx.set_profiling(True)
    runs = []
    t0 = time.perf_counter()
    for i in range(a.steps):
        rs = one_design()
        runs.append(rs.as_dict())
        log(f"step {i}: {rs.elapsed_ms:.1f} ms, {rs.transport_steps} OT steps, "
            f"{rs.phi_sweeps}+{rs.h_sweeps} sweeps, max residual "
            f"{max(rs.max_phi_residual, rs.max_h_residual):.2e}")
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    ctx.set_profiling(False)
    sor_ms, sor_n, sor_bytes = ctx.kernel_stats(capi.KERNEL_SOR_SWEEP)
    res_ms, res_n, _ = ctx.kernel_stats(capi.KERNEL_RESIDUAL)
    ras_ms, ras_n, _ = ctx.kernel_stats(capi.KERNEL_RASTER)
    vtx_ms, vtx_n, _ = ctx.kernel_stats(capi.KERNEL_VERTEX)

    for r in runs:
        if r["status"] != capi.OK:
            raise SystemExit(f"design failed: status {r['status']}")
    value = a.steps / elapsed
    per_design = {
        "ot_steps": sum(r["transport_steps"] for r in runs) / len(runs),
        "h_steps": sum(r["height_steps"] for r in runs) / len(runs),
        "sweeps": sum(r["phi_sweeps"] + r["h_sweeps"] for r in runs) / len(runs),
    }
    avg_launch_ms = sor_ms / max(sor_n, 1)
    # the library's timer counts SWEEP_BYTES_PER_CELL per cell per fused sweep
    sweep_bytes_per_launch = sor_bytes / max(sor_n, 1)
    sweeps_per_launch = sweep_bytes_per_launch / (SWEEP_BYTES_PER_CELL * N * N)
    launch_bytes = LAUNCH_BYTES_PER_CELL * N * N
    achieved = launch_bytes / (avg_launch_ms * 1e-3) / 1e9 if sor_n else 0.0
    sweep_rate = sweep_bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9 if sor_n else 0.0

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * elapsed / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 sweeps, bf16 rhs, f64 residual" if prec == capi.PREC_MIXED else "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"full caustic design (init + <=50 OT steps + 3 height steps), "
                         f"{M}x{M} mesh, {N}x{N} Poisson grid, target '{a.config}' seed {a.seed}"),
            "mesh": M, "grid": N, "width": W, "focal_l": f,
            "stop": f"FP64 relative residual <= {a.tol:g} for every Poisson solve",
            "precision": ("red-black SOR sweeps computed in FP32 on an FP32 correction; the "
                          "sweeps read a BF16 copy of the right-hand side (the FP64 residual, "
                          "rounded at each FP64 check to <= 2^-9 relative); FP64 base, residual "
                          "norm, zero-mean projection, defect correction and stop test; FP64 "
                          "geometry")
            if prec == capi.PREC_MIXED else "FP64 red-black SOR sweeps, FP64 geometry",
            "learned_init": False,
            "per_design": per_design,
        },
        "roofline": {
            "kernel": "k_rb_blocked (temporally blocked red-black SOR, "
                      f"{sweeps_per_launch:.2f} sweeps per launch on average)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": profiled_traffic(N),
            "traffic_source": os.path.relpath(PROFILE_SUMMARY, ROOT) if profiled_traffic(N) else None,
            "bytes_per_launch": launch_bytes,
            "bytes_model": (f"{LAUNCH_BYTES_PER_CELL:g} B per cell per launch: FP32 correction "
                            "read + written once, BF16 right-hand side read once"),
            "avg_launch_us": avg_launch_ms * 1e3,
            "launches": sor_n,
            "sweeps_per_launch": sweeps_per_launch,
            "sweep_rate_GBs": sweep_rate,
            "sweep_rate_model": (f"{SWEEP_BYTES_PER_CELL:g} B per cell per red+black sweep x "
                                 "sweeps per launch / launch time: the bytes an unfused sweep "
                                 "would move, not HBM traffic"),
        },
        "time_split_ms": {"sor": sor_ms, "residual_checks": res_ms, "raster": ras_ms,
                          "vertex": vtx_ms, "total": elapsed * 1e3},
    }
    wpath = a.weights or os.path.join(ROOT, "poisson_caustic_design_amd", "weights",
                                      f"learned_grid{N}.npz")
    if a.learned == "auto" and os.path.exists(wpath):
        plain_final = {k: cd.get(k) for k in ("phi", "h", "target_points", "source_points")}
        out["learned"] = learned_leg(ctx, cd, one_design, wpath, a, value)
        # equal final error: the two forms' final designs against each other
        tp, sp = cd.get("target_points"), cd.get("source_points")
        zr = np.abs(plain_final["source_points"][:, 2]).max()
        out["learned"]["vs_plain_final_design"] = {
            "max_abs_target_xy_over_width": float(
                np.abs(tp[:, :2] - plain_final["target_points"][:, :2]).max() / W),
            "max_abs_height_over_max_height": float(
                np.abs(sp[:, 2] - plain_final["source_points"][:, 2]).max() / max(zr, 1e-300)),
        }
        # the FP64 reference needs an attainable 1e-9 (rounding floor ~ 4 N^2 / pi^2 * 2^-53)
        if a.tol == 1e-6 and prec == capi.PREC_MIXED and N <= 2048:
            out["learned"]["equal_error"] = equal_error_leg(ctx, cd, one_design, wpath, a, value,
                                                            plain_final, img, prm)
    if a.batch > 1:
        out["batched"] = batched_leg(img, prm, a, wpath)
    cb = None
    if not a.no_cpu_baseline:
        cb = cpu_baseline(a.config, a.seed, per_design, a.cpu_timeout)
    out["cpu_baseline"] = ({k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
                           if cb else None)
    if not a.no_cpu_baseline and not a.no_c1:
        out["c1_end_to_end"] = c1_end_to_end(a.cpu_timeout)
    if cb:
        out["speedup_vs_cpu"] = value / cb["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
