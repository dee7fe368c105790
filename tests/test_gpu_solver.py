"""GPU parity of the Poisson solver (cd_poisson_solve through the C ABI)
against the FP64 oracle and the exact DCT solution."""
import numpy as np
import pytest

from oracle.bind import LEXICOGRAPHIC, RED_BLACK
from poisson_caustic_design_amd import capi
from tests.helpers import apply_operator, dct_solve, smooth_rhs

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    a = a - a.mean()
    b = b - b.mean()
    return np.abs(a - b).max() / np.abs(b).max()


def test_fp64_red_black_iterates_match_oracle(ctx, oracle, rng):
    """Fixed sweep count, FP64 sweeps, reference omega: the GPU red-black
    iterate equals the oracle's red-black iterate (tolerance 1e-12 rel: the
    GPU additionally projects the ~1e-17 mean of the right-hand side)."""
    for (w, h, k) in [(64, 64, 37), (33, 20, 11), (128, 96, 50)]:
        b = rng.standard_normal((h, w)); b -= b.mean()
        p0 = 0.1 * rng.standard_normal((h, w))
        o = capi.default_opts(max_sweeps=k, stop=capi.STOP_MAX_UPDATE, tol=1e-300,
                              precision=capi.PREC_FP64, omega=-1.0)
        g, st = ctx.poisson_solve(b, p0, o)
        ref, _ = oracle.poisson_sor(b, p0, k, max_update_tol=1e-300, ordering=RED_BLACK)
        assert st.sweeps == k
        assert np.abs(g - ref).max() <= 1e-12 * np.abs(ref).max(), (w, h)


@pytest.mark.parametrize("w,h,k,every", [(64, 64, 13, 5), (300, 130, 9, 4), (129, 77, 22, 7),
                                         (1000, 50, 6, 6)])
def test_fp32_blocked_iterates_match_oracle(ctx, oracle, rng, w, h, k, every):
    """Mixed precision (temporally blocked FP32 sweeps on a BF16 right-hand side,
    FP64 flushes every `every` sweeps): after k sweeps the iterate equals k
    red-black sweeps of the oracle to the BF16 rounding of the right-hand side
    (2^-9 relative), across launch and tile boundaries; an indexing or
    ordering error would be O(1)."""
    b = rng.standard_normal((h, w)); b -= b.mean()
    p0 = 0.1 * rng.standard_normal((h, w))
    o = capi.default_opts(max_sweeps=k, stop=capi.STOP_MAX_UPDATE, tol=1e-300,
                          precision=capi.PREC_MIXED, omega=-1.0, check_every=every)
    g, st = ctx.poisson_solve(b, p0, o)
    ref, _ = oracle.poisson_sor(b, p0, k, max_update_tol=1e-300, ordering=RED_BLACK)
    assert st.sweeps == k
    scale = np.abs(ref - p0).max()
    assert np.abs(g - ref).max() <= 1e-2 * scale, (w, h, np.abs(g - ref).max(), scale)


@pytest.mark.parametrize("w,h,k,every", [(64, 64, 13, 5), (300, 130, 9, 4), (129, 77, 22, 7),
                                         (1000, 50, 6, 6), (520, 390, 31, 31)])
def test_fp32_residual_rule_iterates_match_oracle(ctx, oracle, rng, w, h, k, every):
    """The residual rule's sweeps (BF16 right-hand side, non-tracking blocked
    kernel, `every` sweeps between FP64 checks): after k sweeps the iterate
    equals k red-black sweeps of the oracle to the BF16 rounding of the
    right-hand side, across launch and tile boundaries (an indexing error is
    O(1))."""
    b = rng.standard_normal((h, w)); b -= b.mean()
    p0 = 0.1 * rng.standard_normal((h, w))
    o = capi.default_opts(max_sweeps=k, stop=capi.STOP_RESIDUAL, tol=0.0,
                          precision=capi.PREC_MIXED, omega=-1.0, check_every=every)
    g, st = ctx.poisson_solve(b, p0, o)
    ref, _ = oracle.poisson_sor(b, p0, k, max_update_tol=1e-300, ordering=RED_BLACK)
    assert st.sweeps == k
    scale = np.abs(ref - p0).max()
    assert np.abs(g - ref).max() <= 1e-2 * scale, (w, h, np.abs(g - ref).max(), scale)


@pytest.mark.parametrize("precision", [capi.PREC_MIXED, capi.PREC_FP64])
@pytest.mark.parametrize("w,h", [(256, 256), (129, 77), (64, 200)])
def test_solve_to_tolerance_matches_dct(ctx, rng, precision, w, h):
    b = smooth_rhs(rng, w, h)
    o = capi.default_opts(tol=1e-9, precision=precision)
    phi, st = ctx.poisson_solve(b, np.zeros_like(b), o)
    assert st.status == capi.OK
    assert st.rel_residual <= 1e-9
    # independent FP64 residual on the host
    r = (b - b.mean()) - apply_operator(phi)
    assert np.linalg.norm(r) <= 1.01e-9 * np.linalg.norm(b - b.mean())
    assert rel_err(phi, dct_solve(b)) < 1e-7


def test_mixed_reaches_1e6_at_2048(ctx, rng):
    n = 2048
    b = smooth_rhs(rng, n, n)
    phi, st = ctx.poisson_solve(b, np.zeros_like(b), capi.default_opts(tol=1e-6))
    assert st.status == capi.OK and st.rel_residual <= 1e-6
    assert rel_err(phi, dct_solve(b)) < 1e-4
    print("2048 cold solve sweeps", st.sweeps, "ms", st.elapsed_ms)


def test_warm_start_needs_no_sweep(ctx, rng):
    b = smooth_rhs(rng, 96, 96)
    ex = dct_solve(b)
    phi, st = ctx.poisson_solve(b, ex, capi.default_opts(tol=1e-6))
    assert st.sweeps == 0 and st.status == capi.OK


def test_not_converged_status(ctx, rng):
    b = smooth_rhs(rng, 128, 128)
    phi, st = ctx.poisson_solve(b, np.zeros_like(b), capi.default_opts(max_sweeps=3, tol=1e-12))
    assert st.status == capi.NOT_CONVERGED and st.sweeps == 3


def test_max_update_rule_like_reference(ctx, oracle, rng):
    """Reference stop rule and omega: the GPU stops once a sweep's max update
    is below the threshold, at a solution as accurate as the reference's."""
    b = smooth_rhs(rng, 64, 64)
    o = capi.default_opts(stop=capi.STOP_MAX_UPDATE, tol=1e-7, omega=-1.0)
    g, st = ctx.poisson_solve(b, np.zeros_like(b), o)
    ref, rst = oracle.poisson_sor(b, np.zeros_like(b), 100000, max_update_tol=1e-7)
    assert st.status == capi.OK
    assert st.last_max_update < 1e-7
    assert 0.5 * rst.sweeps <= st.sweeps <= 2 * rst.sweeps
    ex = dct_solve(b)
    assert rel_err(g, ex) < 10 * max(rel_err(ref, ex), 1e-9)


def test_nan_holes(ctx, oracle, rng):
    b = rng.standard_normal((48, 40))
    b[10:14, 12:20] = np.nan
    b[0, 5] = np.nan
    m = ~np.isnan(b)
    b[m] -= b[m].mean()
    g, st = ctx.poisson_solve(b, np.zeros_like(b), capi.default_opts(tol=1e-11))
    ref, _ = oracle.poisson_sor(b, np.zeros_like(b), 200000, rel_residual_tol=1e-13,
                                check_every=20)
    assert np.isnan(g[~m]).all()
    d = (g[m] - g[m].mean()) - (ref[m] - ref[m].mean())
    assert np.abs(d).max() < 1e-8 * np.abs(ref[m]).max()


def test_dropin_poisson_solver_semantics(rng):
    """capi.poisson_solver mirrors solver.h:8 (in-place, max_threads <= 0 no-op)."""
    b = smooth_rhs(rng, 32, 32)
    phi = np.zeros_like(b)
    capi.poisson_solver(b, phi, 32, 32, 100000, 1e-7, 0)
    assert (phi == 0).all()
    capi.poisson_solver(b, phi, 32, 32, 100000, 1e-7, 4)
    assert rel_err(phi, dct_solve(b)) < 1e-4
