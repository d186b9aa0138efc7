"""Downsampling (from1024to512.f90) and column density (column_density.f90): the oracle against
direct numpy statements of the same formulas (CPU)."""
import numpy as np

import refbind as R


def test_coarsen_is_block_sum_with_mean_one():
    n = 16
    x = (np.random.default_rng(1).random(n ** 3) + 0.5).astype(np.float32)
    out = R.or_coarsen(x, n, 2)
    X = x.reshape((n, n, n), order="F").astype(np.float64)
    blk = X.reshape(4, 4, 4, 4, 4, 4, order="F").sum(axis=(0, 2, 4)).ravel(order="F")
    ref = blk / blk.sum() * 4 ** 3
    assert np.allclose(out, ref, rtol=3e-7, atol=0)
    assert abs(out.astype(np.float64).mean() - 1) < 1e-6


def test_column_density_flat_mesh_and_shift():
    ng, nplot = 16, 4
    den = (np.random.default_rng(2).random(ng ** 3) + 0.5).astype(np.float32)
    d = den.reshape((ng, ng, ng), order="F").astype(np.float64)
    dn = d / d.sum() * ng ** 3
    # undeformed mesh: the column is cells ng/2 .. ng/2 + nplot - 1 exactly
    proj = R.or_column_density(den, np.zeros(ng ** 3, np.float32), ng, nplot).reshape((ng, ng), order="F")
    k1 = ng // 2
    ref = dn[:, :, k1 - 1:k1 - 1 + nplot].sum(axis=2) / nplot
    assert np.allclose(proj, ref, rtol=1e-6, atol=0)
    # def = a z^2 / 2 moves cell faces by dz = a (z + 1/2)-ish; check against the formula directly
    z = np.arange(1, ng + 1, dtype=np.float64)
    a = 0.05
    defn = np.broadcast_to(0.5 * a * (z - ng / 2) ** 2, (ng, ng, ng)).astype(np.float32).ravel(order="F")
    proj2 = R.or_column_density(den, defn, ng, nplot).reshape((ng, ng), order="F")
    D = defn.reshape((ng, ng, ng), order="F").astype(np.float64)
    dspz = (np.roll(D, -1, 2) - D)  # 4-point means of a z-only field are the field itself
    kk1 = k1 - 1 + dspz[0, 0, k1 - 2]
    kk2 = k1 + nplot - 1 + dspz[0, 0, k1 + nplot - 2]
    nk1, nk2 = int(np.floor(kk1)) + 1, int(np.floor(kk2)) + 1
    r1, r2 = nk1 - kk1, kk2 - nk2 + 1
    ref2 = (r1 * dn[:, :, nk1 - 1] + dn[:, :, nk1:nk2 - 1].sum(axis=2) + r2 * dn[:, :, nk2 - 1]) / nplot
    assert np.allclose(proj2, ref2, rtol=1e-5, atol=0)
