"""Pins the oracle's Poisson SOR (oracle.c: orc_poisson_sor, orc_residual)
against the reference build and against the mathematics of the operator."""
import numpy as np
import pytest

from oracle.bind import LEXICOGRAPHIC, RED_BLACK
from tests.helpers import apply_operator, dct_solve, eigvec, smooth_rhs


@pytest.mark.parametrize("w,h,iters", [(16, 16, 37), (33, 20, 200), (7, 41, 5)])
def test_sor_bitexact_vs_reference(oracle, reflib, rng, w, h, iters):
    D = rng.standard_normal((h, w)); D -= D.mean()
    p0 = 0.1 * rng.standard_normal((h, w))
    a = reflib.poisson_solver(D, p0, iters, 1e-300)
    b, st = oracle.poisson_sor(D, p0, iters, max_update_tol=1e-300)
    assert st.sweeps == iters
    assert np.array_equal(a, b)


def test_sor_stop_rule_vs_reference(oracle, reflib, rng):
    D = smooth_rhs(rng, 48, 40)
    p0 = np.zeros_like(D)
    a = reflib.poisson_solver(D, p0, 100000, 1e-7)
    b, st = oracle.poisson_sor(D, p0, 100000, max_update_tol=1e-7)
    assert np.array_equal(a, b)
    assert st.last_max_update < 1e-7


def test_sor_nan_holes_vs_reference(oracle, reflib, rng):
    D = rng.standard_normal((24, 30))
    D[5:9, 10:14] = np.nan
    D[0, 0] = np.nan
    m = ~np.isnan(D)
    D[m] -= D[m].mean()
    p0 = np.zeros_like(D)
    a = reflib.poisson_solver(D, p0, 300, 1e-300)
    b, _ = oracle.poisson_sor(D, p0, 300, max_update_tol=1e-300)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    assert np.array_equal(a[m], b[m])
    assert np.isnan(b[~m]).all()


def test_residual_on_eigenvector(oracle):
    """A v = lambda v for the DCT eigenvectors: residual of (b=lambda v, phi=v) is 0."""
    for (w, h, k, l) in [(20, 12, 3, 2), (17, 17, 1, 0), (9, 31, 4, 7)]:
        v, lam = eigvec(w, h, k, l)
        r, rn, bn = oracle.residual(lam * v, v)
        assert rn <= 1e-12 * bn
        # plausible-mistake check: a sign flip of the operator is caught
        r2, rn2, _ = oracle.residual(-lam * v, v)
        assert rn2 > 1.0 * bn


@pytest.mark.parametrize("ordering", [LEXICOGRAPHIC, RED_BLACK])
def test_sor_converges_to_dct_solution(oracle, rng, ordering):
    w, h = 40, 28
    b = smooth_rhs(rng, w, h)
    phi, st = oracle.poisson_sor(b, np.zeros_like(b), 100000, rel_residual_tol=1e-12,
                                 check_every=10, ordering=ordering)
    ref = dct_solve(b)
    err = (phi - phi.mean()) - ref
    assert st.rel_residual <= 1e-12
    assert np.abs(err).max() <= 1e-9 * np.abs(ref).max()
    # operator identity of the DCT solution itself
    assert np.abs(apply_operator(ref) - b).max() < 1e-10 * np.abs(b).max()


def test_sor_eigenvector_known_answer(oracle):
    v, lam = eigvec(24, 24, 2, 1)
    phi, st = oracle.poisson_sor(lam * v, np.zeros((24, 24)), 100000, rel_residual_tol=1e-13,
                                 check_every=5)
    assert np.abs((phi - phi.mean()) - v).max() < 1e-9


def test_omega_literal(oracle, rng):
    """The reference omega is 2/(1+3.14159265/width) (solver.cpp:71): passing it
    explicitly reproduces the default iterate exactly, another value does not."""
    D = rng.standard_normal((10, 14)); D -= D.mean()
    a, _ = oracle.poisson_sor(D, np.zeros_like(D), 20, max_update_tol=1e-300)
    b, _ = oracle.poisson_sor(D, np.zeros_like(D), 20, max_update_tol=1e-300,
                              omega=2.0 / (1.0 + 3.14159265 / 14))
    c, _ = oracle.poisson_sor(D, np.zeros_like(D), 20, max_update_tol=1e-300,
                              omega=2.0 / (1.0 + np.pi / 14))
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)
