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


# ---------------------------------------------------------------------------------------------
# gridfigure.f90:41-131 (the mesh picture).  The reference program cannot be built here: it is a
# program with fixed /scratch2 input paths, nf = 128, ng = 1024 and about 36 GB of static arrays.
# The oracle is pinned by the closed forms below.

def _gf_inputs(nf, ng, seed=3):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal(nf ** 3).astype(np.float32)
    for _ in range(3):
        R.or_nsmooth(d, nf)
    den = (rng.random(ng ** 3) + 0.5).astype(np.float32)
    return (d * np.float32(2.0)).astype(np.float32), den


def _slab(slab, nf):
    # dsp_slab(c, a, b, d'), c = 0..1, a, b = 0..nf, d' = 0..3 (plane nf/2 + d')
    return slab.reshape((2, nf + 1, nf + 1, 4), order="F")


def test_gridfigure_corner_displacements_closed_form():
    """def = A cos(2 pi x / nf): every cell's four x-differences are equal, so the x displacement of
    corner (i, j, k) is ratio * (def(i+1) - def(i)) and the y displacement is exactly 0."""
    nf, ng = 8, 32
    x = np.arange(1, nf + 1)
    d1 = (0.7 * np.cos(2 * np.pi * x / nf)).astype(np.float32)
    defn = np.broadcast_to(d1[:, None, None], (nf, nf, nf)).ravel(order="F").copy()
    pos, proj, slab = R.or_gridfigure(defn, nf, np.ones(ng ** 3, np.float32), ng)
    S = _slab(slab, nf)
    ratio = ng // nf
    want = ratio * (np.roll(d1, -1).astype(np.float64) - d1)
    for a in range(1, nf + 1):
        assert np.allclose(S[0, a, 1:, :], want[a - 1], rtol=1e-6, atol=1e-6)
    assert np.all(S[1, 1:, 1:, :] == 0)
    assert np.all(S[:, 0, :, :] == 0) and np.all(S[:, :, 0, :] == 0)  # ghosts before the fill


def test_gridfigure_positions_at_coarse_nodes():
    """A fine point on a coarse node has weights (0, ..., 0, 1): its position is index + 0.5 plus
    that corner's displacement, bit for bit.  Interior nodes read the corner (mi+1, mj+1, nf/2+1);
    nodes of the last column (mi = nf) read the next row's ghost, which the reference fills from the
    corner (nf, mj+1, nf/2), one plane down (its three-axis shift)."""
    nf, ng = 8, 32
    ratio = ng // nf
    defn, den = _gf_inputs(nf, ng)
    pos, proj, slab = R.or_gridfigure(defn, nf, den, ng)
    P = pos.reshape((2, nf, nf), order="F")
    S = _slab(slab, nf)
    for mj in range(1, nf):
        for mi in range(1, nf):
            assert P[0, mi - 1, mj - 1] == np.float32(np.float32(ratio * mi) + np.float32(0.5)) + S[0, mi + 1, mj + 1, 1]
            assert P[1, mi - 1, mj - 1] == np.float32(np.float32(ratio * mj) + np.float32(0.5)) + S[1, mi + 1, mj + 1, 1]
        # last column: corner (nf + 1, mj + 1) is storage-wise ghost (0, mj + 2), = corner (nf, mj + 1, plane - 1)
        if mj + 2 <= nf:
            assert P[0, nf - 1, mj - 1] == np.float32(np.float32(ng) + np.float32(0.5)) + S[0, nf, mj + 1, 0]


def test_gridfigure_column_density_flat_mesh():
    """def = 0 leaves the mesh Cartesian: the column of cells ng/2 .. ng/2 + nplot - 1 is their mean
    density (after scaling den to mean 1), and every position is index + 0.5."""
    nf, ng = 8, 32
    _, den = _gf_inputs(nf, ng)
    for nplot in (1, 3):
        pos, proj, _ = R.or_gridfigure(np.zeros(nf ** 3, np.float32), nf, den, ng, nplot)
        P = pos.reshape((2, nf, nf), order="F")
        r = (np.arange(1, nf + 1) * (ng // nf) + 0.5).astype(np.float32)
        assert np.array_equal(P[0], np.broadcast_to(r[:, None], (nf, nf)))
        assert np.array_equal(P[1], np.broadcast_to(r[None, :], (nf, nf)))
        d = den.reshape((ng, ng, ng), order="F").astype(np.float64)
        dn = d / d.sum() * ng ** 3
        k1 = ng // 2
        ref = dn[:, :, k1 - 1:k1 - 1 + nplot].sum(axis=2) / nplot
        assert np.allclose(proj.reshape((ng, ng), order="F"), ref, rtol=1e-6, atol=0)


def test_gridfigure_interpolation_is_trilinear():
    """Between nodes the displacement is the trilinear blend of the eight corners (weights summing to
    1): for a def linear in z every interior corner moves by the same dz, so the interpolated
    displacement is that constant and the column [kk1, kk2] is the undeformed one shifted by dz."""
    nf, ng = 8, 32
    z = np.arange(1, nf + 1)
    defn = np.broadcast_to(0.3 * z[None, None, :], (nf, nf, nf)).astype(np.float32).ravel(order="F")
    # def linear in z with a periodic wrap: interior cells have dz = ratio * 0.3, the last wraps
    _, den = _gf_inputs(nf, ng)
    pos, proj, slab = R.or_gridfigure(defn, nf, den, ng, 1)
    ratio = ng // nf
    # columns shift by dz = 0.3 ratio everywhere (planes ng/2 +- 1 are interior): kk1 = k1 - 1 + dz,
    # kk2 = k1 + dz
    dz = np.float32(np.float32(ratio) * np.float32(1.0) / np.float32(4.0)) * np.float32(4 * 0.3)
    d = den.reshape((ng, ng, ng), order="F").astype(np.float64)
    dn = d / d.sum() * ng ** 3
    k1 = ng // 2
    kk1, kk2 = k1 - 1 + float(dz), k1 + float(dz)
    nk1, nk2 = int(np.floor(kk1)) + 1, int(np.floor(kk2)) + 1
    r1, r2 = nk1 - kk1, kk2 - nk2 + 1
    ref = r1 * dn[:, :, nk1 - 1] + dn[:, :, nk1:nk2 - 1].sum(axis=2) + r2 * dn[:, :, nk2 - 1]
    # columns with i or j below ratio + 1 read the ghost layer (corner index 0), and columns i or
    # j = ng read corner nf + 1 (storage-wise the next row's ghost); those ghosts hold the shifted or
    # zero values the reference fills them with.  Every other column sees the constant dz.
    got = proj.reshape((ng, ng), order="F")
    inner = (slice(ratio, ng - 1), slice(ratio, ng - 1))
    assert np.allclose(got[inner], ref[inner], rtol=2e-5, atol=0)
    assert not np.allclose(got[-1, ratio:-1], ref[-1, ratio:-1], rtol=2e-5, atol=0)
