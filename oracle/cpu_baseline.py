"""CPU baseline of the design path, measured on the host cores (test/bench
infrastructure; the GPU product never imports it).

Runs the reference's own code (oracle/_ref/libcdref.so, kind "reference") or,
when that build is absent, the FP64 restatement (liboracle.so, kind "port"),
on a bounded sample of the benchmark workload and prints one JSON line of
measured component times:

  t_init      initialize_solvers at full size                (caustic_design.cpp:334-364)
  t_ot_ns     one perform_transport_iteration without its Poisson solve
              (the reference's nthreads = 0 case runs zero SOR tiles,
              solver.cpp:73-83), i.e. dual cells, BVH rasterization, gradient,
              vertex step
  t_h_ns      one perform_height_map_iteration without its Poisson solve
  t_sweep     one SOR sweep of poisson_solver at full grid size with the
              reference's own tile threading (solver.cpp:70-147)

A full design then costs  t_init + n_ot*t_ot_ns + n_h*t_h_ns + sweeps*t_sweep,
with n_ot, n_h and the sweep count taken from the GPU run it is compared with.

With --full the design is instead run end to end, as the reference's CLI
runs it (main.cpp:237-262: initialize_solvers, <= 50 transport steps with the
break at conv_tres, 3 height steps, the reference's own max-update stops and
omega), and the wall time of the whole design is printed (t_design); the
benchmark does this at the C1 size (100x100 mesh, 400x400 grid).

The libraries are built by oracle/Makefile with -O3 -march=x86-64-v3
-ffp-contract=off (COMPILE_FLAGS below).

    python -m oracle.cpu_baseline --config noise3 --seed 0 --threads 16
    python -m oracle.cpu_baseline --config gauss5 --seed 0 --threads 16 --full
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


COMPILE_FLAGS = "-O3 -march=x86-64-v3 -ffp-contract=off"


def full_design(config: str, seed: int, threads: int) -> dict:
    """One complete design through the reference's own Caustic_design with the
    CLI's defaults, timed end to end (reference build only)."""
    from oracle import bind
    from poisson_caustic_design_amd import synth

    if not bind.RefLib.available():
        raise SystemExit("full-design baseline needs oracle/_ref/libcdref.so")
    img, prm = synth.target_for(config, seed)
    M, N, W, H, f = prm["mesh_res"], prm["res"], prm["width"], prm["height"], prm["focal_l"]
    d = bind.RefDesign(bind.RefLib(), M, M, N, N, W, H, f, float(np.float32(0.1)), threads)
    conv = float(np.float32(0.01))
    t0 = time.perf_counter()
    d.init(img)
    steps = 0
    for _ in range(50):
        steps += 1
        if d.transport_step() < conv:
            break
    for _ in range(3):
        d.height_step()
    t = time.perf_counter() - t0
    return dict(config=config, seed=seed, mesh_res=M, grid=N, threads=threads,
                tiles=int(math.floor(math.sqrt(threads))) ** 2, nproc=os.cpu_count(),
                cpu_model=cpu_model(), kind="reference", flags=COMPILE_FLAGS,
                t_design=t, ot_steps=steps, h_steps=3)


def measure(config: str, seed: int, threads: int, sweep_seconds: float = 4.0) -> dict:
    from oracle import bind
    from poisson_caustic_design_amd import synth

    img, prm = synth.target_for(config, seed)
    M, N, W, H, f = prm["mesh_res"], prm["res"], prm["width"], prm["height"], prm["focal_l"]
    out = dict(config=config, seed=seed, mesh_res=M, grid=N, threads=threads,
               tiles=int(math.floor(math.sqrt(threads))) ** 2, nproc=os.cpu_count(),
               cpu_model=cpu_model(), flags=COMPILE_FLAGS)
    if bind.RefLib.available():
        ref = bind.RefLib()
        out["kind"] = "reference"
        d = bind.RefDesign(ref, M, M, N, N, W, H, f, 0.1, threads)
        t = time.perf_counter(); d.init(img); out["t_init"] = time.perf_counter() - t
        d.set_threads(0)
        t = time.perf_counter(); d.transport_step(); out["t_ot_ns"] = time.perf_counter() - t
        t = time.perf_counter(); d.height_step(); out["t_h_ns"] = time.perf_counter() - t
        rng = np.random.default_rng(seed)
        b = rng.standard_normal((N, N)); b -= b.mean()
        p = np.zeros_like(b)
        t = time.perf_counter(); ref.poisson_solver(b, p, 0, 0.0, threads)
        t_copy = time.perf_counter() - t
        k = 4
        t = time.perf_counter(); ref.poisson_solver(b, p, k, 0.0, threads)
        per = max((time.perf_counter() - t - t_copy) / k, 1e-9)
        k = max(4, int(sweep_seconds / per))
        t = time.perf_counter(); ref.poisson_solver(b, p, k, 0.0, threads)
        out["t_sweep"] = max((time.perf_counter() - t - t_copy) / k, 1e-9)
        out["sweeps_timed"] = k
    else:
        o = bind.Oracle()
        out["kind"] = "port"
        out["threads"] = 1
        out["tiles"] = 1
        d = bind.OracleDesign(o, M, M, N, N, W, H, f)
        t = time.perf_counter(); d.init(img); out["t_init"] = time.perf_counter() - t
        d.set_solver(bind.STOP_REFERENCE, 1e-7, 1e-8, 0)
        t = time.perf_counter(); d.transport_step(); out["t_ot_ns"] = time.perf_counter() - t
        t = time.perf_counter(); d.height_step(); out["t_h_ns"] = time.perf_counter() - t
        rng = np.random.default_rng(seed)
        b = rng.standard_normal((N, N)); b -= b.mean()
        k = 8
        t = time.perf_counter(); o.poisson_sor(b, np.zeros_like(b), k, max_update_tol=1e-300)
        out["t_sweep"] = (time.perf_counter() - t) / k
        out["sweeps_timed"] = k
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="noise3")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--threads", type=int, default=min(os.cpu_count() or 1, 16))
    ap.add_argument("--sweep-seconds", type=float, default=4.0)
    ap.add_argument("--full", action="store_true", help="time one complete design end to end")
    a = ap.parse_args(argv)
    if a.full:
        print(json.dumps(full_design(a.config, a.seed, a.threads)), flush=True)
    else:
        print(json.dumps(measure(a.config, a.seed, a.threads, a.sweep_seconds)), flush=True)


if __name__ == "__main__":
    main()
