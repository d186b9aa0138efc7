"""CPU: the Galerkin statement's pieces against the oracle (the assembled fine operator against the
oracle's residual, P against the oracle's prolongation) and R A P's defining properties."""
import numpy as np

import galerkin_ref as GR
import refbind as R


def smooth_def(n, amp, seed):
    d = np.random.default_rng(seed).standard_normal(n ** 3)
    for _ in range(4):
        R.or_nsmooth(d, n)
    return d * amp


def test_assembled_operator_matches_oracle_residual():
    n = 16
    rho = R.or_lognormal(n, 1.5, 3)
    d = smooth_def(n, 2.0, 4)
    G = R.or_metric(d, rho, 1, n, 1.0)
    A = GR.op19(G, n, 1.0)
    phi = np.random.default_rng(1).standard_normal(n ** 3)
    r, _ = R.or_residual(phi, np.zeros(n ** 3), G, n, 1.0)
    assert np.abs(-(A @ phi) - r).max() <= 1e-13 * np.abs(r).max()
    # symmetric by construction (gauss1.f90:586-588)
    assert abs(A - A.T).max() <= 1e-13 * abs(A).max()


def test_prolongation_matches_oracle():
    nc = 8
    x = np.random.default_rng(2).standard_normal(nc ** 3)
    # same weights; the oracle sums eight terms and divides by 8, the matrix weights each term
    y = R.or_prolong(x, 2 * nc)
    assert np.abs(GR.prolongation(nc) @ x - y).max() <= 1e-15 * np.abs(y).max()


def test_galerkin_operator_properties():
    n = 16
    rho = R.or_lognormal(n, 2.0, 5)
    (A1,) = GR.galerkin_levels(np.zeros(n ** 3), rho, 1, n, 1.0, 1)
    # symmetric, annihilates constants (the fine operator does), 27-point
    assert abs(A1 - A1.T).max() <= 1e-13 * abs(A1).max()
    assert np.abs(A1 @ np.ones(A1.shape[0])).max() <= 1e-12 * abs(A1).max()
    assert np.diff(A1.tocsr().indptr).max() == 27
