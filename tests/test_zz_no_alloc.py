"""cd_poisson_solve_device_noalloc, the allocation-free variant of
cd_poisson_solve_device: called through the C ABI on device buffers, it gives
the same solution and statistics as the allocating entry point (the same
kernels on the same inputs: bit-identical), on two calls in a row with the
second smaller than the first, with and without NaN holes and under both stop
rules; a grid beyond its scratch is refused before any work."""
import ctypes as C

import numpy as np
import pytest
import torch

from poisson_caustic_design_amd import capi
from tests.helpers import smooth_rhs

pytestmark = pytest.mark.gpu


def solve(ctx, fn, b, phi0, opts):
    h, w = b.shape
    d_b = torch.tensor(b, dtype=torch.float64, device="cuda:0")
    d_p = torch.tensor(phi0, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    st = capi.SolveStats()
    rc = fn(ctx.h, C.c_void_p(d_b.data_ptr()), C.c_void_p(d_p.data_ptr()), w, h, C.byref(opts),
            C.byref(st))
    ctx.synchronize()
    return rc, d_p.cpu().numpy(), st


def compare(ctx, b, phi0, opts):
    L = capi.load()
    rc_a, pa, sa = solve(ctx, L.cd_poisson_solve_device, b, phi0, opts)
    rc_n, pn, sn = solve(ctx, L.cd_poisson_solve_device_noalloc, b, phi0, opts)
    assert rc_a == rc_n == capi.OK, (rc_a, rc_n)
    assert np.array_equal(pa, pn, equal_nan=True)
    assert (sa.sweeps, sa.checks) == (sn.sweeps, sn.checks)
    assert sa.rel_residual == sn.rel_residual
    return sn


@pytest.mark.parametrize("stop", ["residual", "max_update"])
def test_noalloc_matches_allocating_solve_twice_smaller_second(ctx, stop):
    rng = np.random.default_rng(3)
    kw = (dict(tol=1e-8) if stop == "residual"
          else dict(stop=capi.STOP_MAX_UPDATE, tol=1e-7, omega=-1.0))
    for (w, h) in ((512, 384), (200, 136)):
        b = smooth_rhs(rng, w, h)
        st = compare(ctx, b, np.zeros_like(b), capi.default_opts(**kw))
        print(f"{stop} {w}x{h}: {st.sweeps} sweeps, rel {st.rel_residual:.2e}")


def test_noalloc_with_holes_and_warm_start(ctx):
    rng = np.random.default_rng(4)
    for (w, h) in ((256, 256), (96, 130)):
        b = smooth_rhs(rng, w, h)
        b[h // 3: h // 3 + 7, w // 4: w // 4 + 9] = np.nan
        phi0 = 1e-3 * rng.standard_normal(b.shape)
        compare(ctx, b, phi0, capi.default_opts(tol=1e-8))


def test_noalloc_refuses_grids_beyond_its_scratch(ctx):
    L = capi.load()
    st = capi.SolveStats()
    # no memory is touched: the size check comes first (null-free dummy pointers)
    rc = L.cd_poisson_solve_device_noalloc(ctx.h, C.c_void_p(16), C.c_void_p(16), 2049, 2048,
                                           C.byref(capi.default_opts()), C.byref(st))
    assert rc == capi.BAD_ARG
    assert "scratch" in capi.load().cd_context_last_error(ctx.h).decode()
    rc = L.cd_poisson_solve_device_noalloc(ctx.h, C.c_void_p(16), C.c_void_p(16), 512, 512,
                                           C.byref(capi.default_opts(learned_init=1)),
                                           C.byref(st))
    assert rc == capi.BAD_ARG
