"""Writes the committed golden vectors under tests/golden/ from the reference Fortran built by
oracle/build_ref.sh (oracle/_ref/).  Run in a container that has the reference sources:

    python tests/golden/make_golden.py

Each .npz holds inputs and the reference's outputs (no code):
  mg_<ng>_<prec>.npz     3 multigrid V-cycles (mmh/multigrid.f90:16) from phi = 0
  outer_<ng>_<prec>.npz  3 outer iterations of calcdefp / def += defp / relaxing (mmh/main.f90:95-101)
  fft_<ng>_f32.npz       scalar2gradient and defp2delta of a smooth field (real(4) routines)
Inputs: the lognormal field of the oracle's generator and a smoothed random def.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import refbind as R  # noqa: E402


def aos(rho):
    u = np.zeros((rho.size, 5), rho.dtype)
    u[:, 0] = rho
    return u.ravel()


def smooth_def(n, amp, seed, dt):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal(n ** 3).astype(dt)
    for _ in range(4):
        R.or_nsmooth(d, n)
    return (d * dt(amp)).astype(dt)


def main():
    for ng, var in [(16, "f64"), (16, "f32"), (32, "f32")]:
        dt = np.float64 if var == "f64" else np.float32
        lib = R.ref(ng, var)
        R.ref_set_compressmax(lib, 10.0, dt)
        n = ng
        rho = R.or_lognormal(n, 2.0, 7).astype(dt)
        u = aos(rho)
        d = smooth_def(n, 2.0, 4, dt)
        rhs = (1 - rho).astype(dt)
        R.or_nsmooth(rhs, n)
        a, b = np.zeros(n ** 3, dt), rhs.copy()
        for _ in range(3):
            R.ref_multigrid(lib, a, b, d, u, 5, 1.0, n)
        np.savez_compressed(os.path.join(HERE, "mg_%d_%s.npz" % (ng, var)), kind="multigrid", n=n, ng=ng, cycles=3,
                            u=u, rhs_in=rhs, phi_out=a, rhs_out=b, **{"def": d})
        if ng == 16:
            rho = R.or_lognormal(n, 1.0, 1234).astype(dt)
            u0 = aos(rho)
            uu, dd = u0.copy(), np.zeros(n ** 3, dt)
            for _ in range(3):
                p, t, q = np.zeros(n ** 3, dt), np.zeros(n ** 3, dt), np.zeros((n + 2) * n * n, dt)
                R.ref_calcdefp(lib, p, t, q, dd, uu)
                dd = (dd + p).astype(dt)
                R.ref_relaxing(lib, uu, dd, p)
            np.savez_compressed(os.path.join(HERE, "outer_%d_%s.npz" % (ng, var)), kind="outer", n=n, ng=ng, iters=3,
                                u=u0, def_out=dd, u_out=uu, **{"def": np.zeros(n ** 3, dt)})
        if var == "f32" and ng == 16:
            phi = smooth_def(n, 3.0, 12, np.float32)
            np.savez_compressed(os.path.join(HERE, "fft_%d_f32.npz" % ng), kind="fft", n=n, ng=ng, phi=phi,
                                grad=R.ref_gradient(lib, phi, n), delta=R.ref_delta(lib, phi, n),
                                u=phi, **{"def": phi})
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main()
