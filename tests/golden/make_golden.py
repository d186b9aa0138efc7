"""Writes the committed golden vectors under tests/golden/ from the reference Fortran built by
oracle/build_ref.sh (oracle/_ref/).  Run in a container that has the reference sources:

    python tests/golden/make_golden.py

Each .npz holds inputs and the reference's outputs (no code):
  mg_<ng>_<prec>.npz     3 multigrid V-cycles (mmh/multigrid.f90:16) from phi = 0
  outer_<ng>_<prec>.npz  3 outer iterations of calcdefp / def += defp / relaxing (mmh/main.f90:95-101)
  fft_<ng>_f32.npz       scalar2gradient and defp2delta of a smooth field (real(4) routines)
  mg_guard_32.npz        multigrid V-cycles whose divergence guard fires on the 16^3, 8^3 and 4^3
                         levels (mmh/mgtemplate.f90:46-55, mmh/multigrid.f90:113-119), FP64 (2
                         cycles) and FP32 (1 cycle), with the levels the reference reports diverged
Inputs: the lognormal field of the oracle's generator and a smoothed random def.
"""
import os
import re
import subprocess
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


def guard_state(n=32):
    """The second outer iteration of a sigma = 3 field (seed 1234) at 32^3: def and u after one
    reference iteration (oracle), and the limiter rhs of the next one.  Rounded to FP32 so that the
    same inputs serve the FP64 and FP32 builds."""
    rho = R.or_lognormal(n, 3.0, 1234)
    u = aos(rho)
    d = np.zeros(n ** 3)
    p, t, q = np.zeros(n ** 3), np.zeros(n ** 3), np.zeros((n + 2) * n * n)
    R.or_calcdefp(p, t, q, d, u, n)
    d += p
    R.or_relaxing(u, d, p, n)
    rhs = R.or_limiter_rhs(d, u, n)
    return u[0::5].astype(np.float32), d.astype(np.float32), rhs.astype(np.float32)


def guard_child(var, cycles, path):
    """Runs the reference's multigrid on the guard state (in a child process, so that its stdout,
    which reports each level whose guard fired, can be read) and saves the iterate."""
    n = 32
    dt = np.float64 if var == "f64" else np.float32
    rho, d, rhs = guard_state(n)
    lib = R.ref(n, var)
    u = aos(rho.astype(dt))
    a, b = np.zeros(n ** 3, dt), rhs.astype(dt)
    for c in range(cycles):
        R.ref_multigrid(lib, a, b, d.astype(dt), u, 5, 1.0, n)
    np.savez(path, phi=a, rhs=b)


def guard_golden():
    n = 32
    rho, d, rhs = guard_state(n)
    out = dict(kind="multigrid_guard", n=n, ng=n, rho=rho, rhs_in=rhs, **{"def": d})
    for var, cycles in (("f64", 2), ("f32", 1)):
        # The reference's stdout is flushed when its process ends, so each cycle count runs in a
        # child of its own; cycle c's report is the tail of the child that ran c + 1 cycles.
        levels, seen = [], 0
        for c in range(1, cycles + 1):
            tmp = os.path.join("/tmp", "guard_%s_%d.npz" % (var, c))
            log = subprocess.run([sys.executable, __file__, "--guard-child", var, str(c), tmp], check=True,
                                 capture_output=True, text=True).stdout
            # "mg: relaxation diverged, level=  N ..." (mmh/mgtemplate.f90:47-49, mmh/multigrid.f90:114-116)
            found = [int(m) for m in re.findall(r"diverged, level=\s*(\d+)", log)]
            levels.append(sorted(found[seen:]))
            seen = len(found)
        z = np.load(tmp)
        out["cycles_" + var] = cycles
        out["phi_out_" + var] = z["phi"]
        out["rhs_out_" + var] = z["rhs"]
        out["diverged_levels_" + var] = np.array(levels, np.int32)
    np.savez_compressed(os.path.join(HERE, "mg_guard_32.npz"), **out)


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
    if len(sys.argv) > 1 and sys.argv[1] == "--guard-child":
        guard_child(sys.argv[2], int(sys.argv[3]), sys.argv[4])
    elif len(sys.argv) > 1 and sys.argv[1] == "--guard":
        guard_golden()
    else:
        main()
        guard_golden()
