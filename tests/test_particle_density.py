"""Particle to grid density assignment (haloden.f90:49-195): the oracle's linked-list restatement
against an independent brute-force statement of the same rules on small inputs, and conservation
properties.  CPU only."""
import numpy as np

import refbind as R


def particles(npart, ng, seed, clustered=True):
    rng = np.random.default_rng(seed)
    if not clustered:
        return (rng.random((npart, 3)) * ng).astype(np.float32)
    centres = rng.random((8, 3)) * ng
    x = centres[rng.integers(0, 8, npart)] + rng.standard_normal((npart, 3)) * ng / 10
    return np.mod(x, ng).astype(np.float32)


def brute_force(xp, ng, nf):
    """The rules of haloden.f90 written directly: NGP cells; every empty cell takes the particle
    of smallest periodic squared distance to its centre among the particles of the smallest cube
    (half-width r = 1, 2, ...) around it that holds any, ties going to the first in the
    reference's scan order (cube cells in (k, j, i) order, a cell's particles by decreasing
    index); mass 1/wp per cell of a particle's chain, summed per cell in particle order."""
    x = (xp - np.floor(xp / np.float32(nf)) * np.float32(nf)).astype(np.float32)
    x = (x * np.float32(ng) / np.float32(nf)).astype(np.float32)
    cell = np.clip(np.floor(x).astype(int), 0, ng - 1)
    owner = {}
    for p in range(len(x)):
        owner.setdefault(tuple(cell[p]), []).append(p)
    wp = np.ones(len(x), int)
    fill = {}
    half, fng = np.float32(ng // 2), np.float32(ng)
    for k in range(ng):
        for j in range(ng):
            for i in range(ng):
                if (i, j, k) in owner:
                    continue
                g = np.array([i + 1, j + 1, k + 1], np.float32) - np.float32(0.5)
                for r in range(1, ng):
                    cands = []
                    for ks in range(k - r, k + r + 1):
                        for js in range(j - r, j + r + 1):
                            for is_ in range(i - r, i + r + 1):
                                c = (is_ % ng, js % ng, ks % ng)
                                for p in reversed(owner.get(c, [])):
                                    t = (x[p] - g) + half
                                    d = (t - np.floor(t / fng) * fng) - half
                                    r2 = np.float32(d[0] * d[0]) + np.float32(d[1] * d[1])
                                    r2 = np.float32(r2 + np.float32(d[2] * d[2]))
                                    cands.append((r2, len(cands), p))
                    if cands:
                        best = min(cands)[2]
                        fill[(i, j, k)] = best
                        wp[best] += 1
                        break
    den = np.zeros((ng, ng, ng), np.float32)  # [i, j, k]
    for (i, j, k), p in fill.items():
        den[i, j, k] = np.float32(1.0) / np.float32(wp[p])
    for p in range(len(x)):
        c = tuple(cell[p])
        den[c] = np.float32(den[c] + np.float32(1.0) / np.float32(wp[p]))
    den = den.ravel(order="F")
    s = den.astype(np.float64).sum()
    return (den.astype(np.float64) / s * float(np.float32(ng) ** 3)).astype(np.float32), wp


def test_oracle_matches_brute_force():
    for ng, npart, seed, cl in [(6, 40, 1, False), (8, 60, 2, True), (8, 200, 3, True)]:
        xp = particles(npart, ng, seed, cl) * np.float32(4.0)  # positions in a box of 4 ng
        den, wp = R.or_particle_density(xp, ng, 4.0 * ng)
        bden, bwp = brute_force(xp, ng, 4.0 * ng)
        assert np.array_equal(wp, bwp)
        assert np.allclose(den, bden, rtol=2.5e-7, atol=0)


def test_conservation():
    ng, npart = 16, 3000
    xp = particles(npart, ng, 7)
    den, wp = R.or_particle_density(xp, ng, float(ng))
    n_empty = ng ** 3 - len({tuple(c) for c in np.floor(xp).astype(int)})
    assert wp.sum() == npart + n_empty          # every cell belongs to exactly one chain or is occupied
    assert abs(den.astype(np.float64).mean() - 1.0) < 1e-6
    assert den.min() > 0
