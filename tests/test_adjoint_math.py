"""The builder-defined loss gradient (no reference counterpart) pinned on the CPU: the adjoint
restatement in adjoint_ref.py against central finite differences of the forward map, and the
binned power estimator against a direct sum over modes.  The forward map's stages are the
oracle's, which the reference builds pin (test_oracle_ref.py)."""
import numpy as np

import adjoint_ref as A
import refbind as R


def test_gradient_matches_central_differences_16():
    n = 16
    rho = R.or_lognormal(n, 1.5, 99)
    w = np.random.default_rng(5).standard_normal(16)
    loss, g = A.gradient(rho, n, w)
    cells = np.random.default_rng(1).choice(n ** 3, 8, replace=False)
    fd = A.fd_central(rho, n, w, cells, 1e-6)
    assert np.abs(g[cells] - fd).max() <= 1e-6 * np.abs(fd).max()
    # a clipped cell (rho > 3) keeps only the metric path and still matches
    clipped = np.flatnonzero(rho > 3.5)[:2]
    fd2 = A.fd_central(rho, n, w, clipped, 1e-6)
    assert np.abs(g[clipped] - fd2).max() <= 1e-6 * np.abs(fd).max()


def test_power_estimator_direct_sum():
    n, nb = 8, 6
    x = np.random.default_rng(3).standard_normal(n ** 3)
    pk, nm = A.power(x, n, nb)
    e = A.log_edges(n, nb)
    X = x.reshape((n, n, n), order="F")
    num, cnt = np.zeros(nb), np.zeros(nb)
    for m1 in range(n):
        for m2 in range(n):
            for m3 in range(n):
                s = [m if m <= n // 2 else m - n for m in (m1, m2, m3)]
                kk = 2 * np.pi * np.sqrt(sum(v * v for v in s)) / n
                if kk < e[0] * (1 - 1e-12) or kk > e[-1] * (1 + 1e-12):
                    continue
                b = min(nb - 1, int(np.searchsorted(e[1:nb], kk, side="right")))
                ii, jj, kx = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
                F = (X * np.exp(-2j * np.pi * (m1 * ii + m2 * jj + m3 * kx) / n)).sum()
                num[b] += abs(F) ** 2
                cnt[b] += 1
    assert np.array_equal(cnt, nm)
    assert np.allclose(pk, np.where(cnt > 0, num / np.maximum(cnt, 1) / n ** 3, 0.0), rtol=1e-12, atol=0)
    # Parseval: all modes in [2 pi/n, pi] with the mean excluded
    assert cnt.sum() < n ** 3
