"""CPU statement of the Galerkin coarse operators (test infrastructure).

The fine operator L phi = (t1 + t2 + tdiag phi) / dx^2 of stencil19 (mmh/gauss1.f90:954-978) is
assembled as a sparse matrix from the oracle's metric; P is the V-cycle's trilinear prolongation
(the oracle's prolong, prolong.f90) and R = P^T / 8 its full weighting.  A_l = R A_(l-1) P.
Both the assembled operator and P are pinned by the CPU tests against the oracle's own residual
and prolongation, so the GPU operators are checked against independent arithmetic."""
import numpy as np
import scipy.sparse as sp

import refbind as R


def op19(G, n, dx):
    n3 = n ** 3
    G1, G2, G3, G4, G5, G6 = [G[c * n3:(c + 1) * n3].reshape((n, n, n), order="F") for c in range(6)]

    def sh(a, di, dj, dk):  # value at (i + di, j + dj, k + dk)
        return np.roll(a, (-di, -dj, -dk), axis=(0, 1, 2))

    b11p, b11m = (sh(G1, 1, 0, 0) + G1) / 2, (sh(G1, -1, 0, 0) + G1) / 2
    b22p, b22m = (sh(G4, 0, 1, 0) + G4) / 2, (sh(G4, 0, -1, 0) + G4) / 2
    b33p, b33m = (sh(G6, 0, 0, 1) + G6) / 2, (sh(G6, 0, 0, -1) + G6) / 2
    co = {(0, 0, 0): -(b11p + b11m + b22p + b22m + b33p + b33m),
          (1, 0, 0): b11p, (-1, 0, 0): b11m, (0, 1, 0): b22p, (0, -1, 0): b22m, (0, 0, 1): b33p, (0, 0, -1): b33m}
    # cross terms of t2: each difference product spread onto its four corner points
    for (a, b), Gc in (((0, 1), G2), ((1, 2), G5), ((0, 2), G3)):
        def off(sa, sb):
            o = [0, 0, 0]
            o[a], o[b] = sa, sb
            return tuple(o)

        def at(sa, sb):
            o = off(sa, sb)
            return sh(Gc, *o)
        # (+a,+b): G(b+) + G(a+); (-a,+b): -(G(b+) + G(a-)); (+a,-b): -(G(b-) + G(a+)); (-a,-b): G(b-) + G(a-)
        co[off(1, 1)] = (at(0, 1) + at(1, 0)) / 4
        co[off(-1, 1)] = -(at(0, 1) + at(-1, 0)) / 4
        co[off(1, -1)] = -(at(0, -1) + at(1, 0)) / 4
        co[off(-1, -1)] = (at(0, -1) + at(-1, 0)) / 4
    I, J, K = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    p = (I + n * (J + n * K)).ravel(order="F")
    rows, cols, vals = [], [], []
    for (di, dj, dk), c in co.items():
        q = (((I + di) % n) + n * (((J + dj) % n) + n * ((K + dk) % n))).ravel(order="F")
        rows.append(p)
        cols.append(q)
        vals.append(c.ravel(order="F") / dx ** 2)
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n ** 3, n ** 3))


def prolongation(nc):
    """Trilinear prolongation from nc^3 to (2 nc)^3 (Fortran order, periodic)."""
    nf = 2 * nc
    r, c, v = [], [], []
    for f in range(nf):
        C, a = f // 2, f % 2
        if a == 0:
            r += [f]; c += [C]; v += [1.0]
        else:
            r += [f, f]; c += [C, (C + 1) % nc]; v += [0.5, 0.5]
    p1 = sp.csr_matrix((v, (r, c)), shape=(nf, nc))
    return sp.kron(p1, sp.kron(p1, p1)).tocsr()


def galerkin_levels(defn, rho_u, nu, n, dx, upto):
    """[A_1, ..., A_upto] as sparse matrices."""
    A = op19(R.or_metric(defn, rho_u, nu, n, dx), n, dx)
    out = []
    m = n
    for _ in range(upto):
        P = prolongation(m // 2)
        A = (P.T @ A @ P / 8.0).tocsr()
        out.append(A)
        m //= 2
    return out


def stencil_fields(A, m):
    """The 27 coefficient fields (Fortran order each) of a periodic 27-point operator on m^3."""
    I, J, K = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    p = (I + m * (J + m * K)).ravel(order="F")
    out = np.zeros((27, m ** 3))
    A = A.tocsr()
    for e in range(27):
        ex, ey, ez = e % 3 - 1, (e // 3) % 3 - 1, e // 9 - 1
        q = (((I + ex) % m) + m * (((J + ey) % m) + m * ((K + ez) % m))).ravel(order="F")
        out[e] = np.asarray(A[p, q]).ravel()
    return out
