"""Build libcaustic_amd.so (gfx950) in-tree with hipcc.

    python -m poisson_caustic_design_amd.build

Every HIP translation unit is compiled with
  --offload-arch=gfx950 -O3 -ffp-contract=off
  -Xclang -target-feature -Xclang -packed-fp32-ops
FP contraction is off so the FP64 geometry kernels round exactly like the
FP64 CPU restatement used by the tests.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libcaustic_amd.so")

HIP_SOURCES = ["cd_context.hip", "cd_solver.hip", "cd_design.hip", "cd_learned.hip"]
CXX_SOURCES = ["dropin_solver.cpp", "dropin_caustic_design.cpp", "cd_io.cpp", "cd_bvh_order.cpp"]

ARCH = os.environ.get("CD_OFFLOAD_ARCH", "gfx950")


def _hipcc() -> str:
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(verbose: bool = False, defines=(), out: str = LIB) -> str:
    """Compile and link; `defines` are extra -D flags (kernel-geometry variants),
    `out` the library path."""
    build_dir = BUILD if out == LIB else out + ".objs"
    os.makedirs(build_dir, exist_ok=True)
    hipcc = _hipcc()
    common = ["-O3", "-fPIC", "-std=c++17", "-I" + INCLUDE, "-I" + CSRC, "-Wall",
              "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result"]
    hip_flags = ["-x", "hip", "--offload-arch=" + ARCH, "-ffp-contract=off",
                 "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                 "-munsafe-fp-atomics", *defines]
    jobs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(build_dir, src + ".o")
        jobs.append([hipcc, *common, *hip_flags, "-c", path, "-o", obj])
    for src in CXX_SOURCES:
        obj = os.path.join(build_dir, src + ".o")
        jobs.append(["g++", *common, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        results = list(ex.map(_run, jobs))
    if verbose:
        for r in results:
            sys.stdout.write(r.stdout + r.stderr)
    objs = [j[-1] for j in jobs]
    tmp = out + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=" + ARCH, *objs, "-o", tmp,
          "-Wl,-soname,libcaustic_amd.so", "-lz"])
    os.replace(tmp, out)
    if out == LIB:
        # the command-line drop-in, linked against the library next to it
        cli = os.path.join(os.path.dirname(LIB), "caustic_design_amd")
        _run(["g++", "-O2", "-std=c++17", "-I" + INCLUDE, os.path.join(CSRC, "cli_main.cpp"),
              "-o", cli + ".tmp", "-L" + os.path.dirname(LIB), "-l:libcaustic_amd.so",
              "-Wl,-rpath,$ORIGIN"])
        os.replace(cli + ".tmp", cli)
        # a caller compiled unchanged against the drop-in C++ headers (tests)
        src = os.path.join(ROOT, "tests", "dropin", "dropin_main.cpp")
        if os.path.exists(src):
            exe = os.path.join(ROOT, "tests", "dropin", "dropin_main")
            _run(["g++", "-O2", "-std=c++17", "-I" + INCLUDE, src, "-o", exe + ".tmp",
                  "-L" + os.path.dirname(LIB), "-l:libcaustic_amd.so",
                  "-Wl,-rpath,$ORIGIN/../../poisson_caustic_design_amd"])
            os.replace(exe + ".tmp", exe)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
