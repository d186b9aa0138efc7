"""Shared pieces of the drop-in driver tests (tests/dropin/dropin_driver.F90): inputs, output
layout, and the same call sequence restated with the oracle."""
import os
import subprocess

import numpy as np

import refbind as R

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NG = 32


def driver_path(kind, ng=NG):
    if kind == "ref":
        return os.path.join(REPO, "oracle", "_ref", "dropin_ref_ng%d" % ng)
    return os.path.join(REPO, "tests", "dropin", "_bin", "dropin_gpu_ng%d" % ng)


def inputs(ng=NG):
    rho = R.or_lognormal(ng, 1.0, 1234).astype(np.float32)
    rng = np.random.default_rng(21)
    d = rng.standard_normal(ng ** 3).astype(np.float32)
    for _ in range(4):
        R.or_nsmooth(d, ng)
    return rho, (d * np.float32(1.5)).astype(np.float32)


def run_driver(path, rho, d, tmpdir, ng=NG):
    fr, fd, fo = (os.path.join(str(tmpdir), f) for f in ("rho.bin", "def.bin", "out.bin"))
    rho.astype(np.float32).tofile(fr)
    d.astype(np.float32).tofile(fd)
    p = subprocess.run([path, fr, fd, fo], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    return split(np.fromfile(fo, dtype=np.float32), ng)


def split(a, ng):
    n3, h3 = ng ** 3, (ng // 2) ** 3
    names = [("relax_phi", n3), ("relax_resid", h3), ("relax_rinf", 1), ("mg_phi", n3), ("mg_rhs", n3),
             ("def", n3), ("defp", n3), ("u1", n3), ("grad", 3 * n3), ("delta", n3)]
    assert a.size == sum(s for _, s in names), a.size
    out, o = {}, 0
    for k, s in names:
        out[k] = a[o:o + s]
        o += s
    return out


def oracle_sequence(rho, d, sumorder, ng=NG):
    """The driver's calls restated with the oracle (FP32)."""
    f = np.float32
    n = ng
    u = np.zeros((n ** 3, 5), f)
    u[:, 0] = rho
    u = u.ravel()
    out = {}
    rhs = ((f(1) - rho) / f(30)).astype(f)
    R.or_nsmooth(rhs, n)
    phi, resid = np.zeros(n ** 3, f), np.zeros((n // 2) ** 3, f)
    rinf = R.or_relax(phi, resid, rhs, d, u, 5, 1.0, n, 8, 19)
    out["relax_phi"], out["relax_resid"], out["relax_rinf"] = phi, resid, np.array([rinf], f)
    phi = np.zeros(n ** 3, f)
    R.or_multigrid(phi, rhs, d, u, 5, 1.0, n, 2, n, sumorder)
    out["mg_phi"], out["mg_rhs"] = phi, rhs
    dd = np.zeros(n ** 3, f)
    ncomp = []
    for _ in range(2):
        p, t, t2 = np.zeros(n ** 3, f), np.zeros(n ** 3, f), np.zeros((n + 2) * n * n, f)
        st = R.or_calcdefp(p, t, t2, dd, u, n, sumorder=sumorder)
        ncomp.append(st.ncompress)
        dd = (dd + p).astype(f)
        R.or_relaxing(u, dd, p, n)
    out["def"], out["defp"], out["u1"] = dd, p, u[0::5].copy()
    out["grad"], out["delta"] = R.or_gradient(dd, n), R.or_delta(dd, n)
    return out, ncomp
