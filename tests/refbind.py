"""ctypes bindings to the oracle (oracle/liboracle.so) and to the flang build of the reference
(oracle/_ref/libmmhref_ng<NG>_<variant>.so).  Test infrastructure only.

Arrays are numpy, flat or (n, n, n) in Fortran order (x fastest), matching the reference layout.
The reference routines are Fortran externals: every argument by reference, symbols lowercase
with a trailing underscore (mmh/gauss1.f90:64 relax, mmh/multigrid.f90:16 multigrid,
mmh/limiter.f90:17 calcdefp, mmh/relaxing.f90:21 relaxing, mmh/gradient.f90 scalar2gradient,
mmh/reconstruct.f90 defp2delta, mmh/limiter.f90:677 nsmooth, mmh/mgutil.f90 restrict/inject).
"""
import ctypes as C
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "liboracle.so")
REF_DIR = os.path.join(REPO, "oracle", "_ref")

_P = C.c_void_p


def ptr(a):
    return a.ctypes.data_as(_P)


def build_oracle():
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle")], check=True)


_oracle = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            build_oracle()
        _oracle = C.CDLL(ORACLE_SO)
        for s in ("d", "f"):
            T = C.c_double if s == "d" else C.c_float
            getattr(_oracle, "or_residual_" + s).restype = T
            getattr(_oracle, "or_sum_" + s).restype = C.c_double
    return _oracle


def _sfx(dt):
    return "d" if np.dtype(dt) == np.float64 else "f"


def _T(dt):
    return C.c_double if np.dtype(dt) == np.float64 else C.c_float


class MgStats(C.Structure):
    _fields_ = [("r1", C.c_double * 16), ("relaxerr", C.c_double * 16), ("guard", C.c_int * 16),
                ("nlev", C.c_int), ("zerosum_entry", C.c_int)]


class LimStats(C.Structure):
    _fields_ = [("cmax", C.c_double), ("detmin", C.c_double), ("detmax", C.c_double),
                ("fluxmin", C.c_double), ("fluxmax", C.c_double), ("rho2", C.c_double),
                ("nexpand", C.c_int), ("ncompress", C.c_int), ("rhomin", C.c_double),
                ("rhomax", C.c_double), ("rms", C.c_double), ("mg", MgStats)]


# ---------------------------------------------------------------------------------------------
# oracle wrappers (in place on numpy arrays of the working dtype)

def or_metric(defp, u, nu, n, dx, lsquare=0):
    dt = defp.dtype
    G = np.zeros(6 * n ** 3, dt)
    getattr(oracle(), "or_metric_" + _sfx(dt))(ptr(G), ptr(defp), ptr(u), nu, n, _T(dt)(dx), lsquare)
    return G


def or_sweep(phi, rhs, G, n, dx, nsweep):
    getattr(oracle(), "or_sweep_" + _sfx(phi.dtype))(ptr(phi), ptr(rhs), ptr(G), n, _T(phi.dtype)(dx), nsweep)


def or_residual(phi, rhs, G, n, dx):
    r = np.zeros(n ** 3, phi.dtype)
    rinf = getattr(oracle(), "or_residual_" + _sfx(phi.dtype))(ptr(r), ptr(phi), ptr(rhs), ptr(G), n,
                                                               _T(phi.dtype)(dx))
    return r, rinf


def or_relax(arr, resid, rhs, defp, u, nu, dx, n, nrelax, iopt):
    dt = arr.dtype
    rinf = _T(dt)(0)
    getattr(oracle(), "or_relax_" + _sfx(dt))(ptr(arr), ptr(resid), ptr(rhs), ptr(defp), ptr(u), nu,
                                              C.byref(rinf), _T(dt)(dx), n, nrelax, iopt)
    return rinf.value


def or_multigrid(u, rhs, defpot, ubig, nu, dx, n, iopt=2, ng=None, sumorder=0):
    dt = u.dtype
    st = MgStats()
    rc = getattr(oracle(), "or_multigrid_" + _sfx(dt))(ptr(u), ptr(rhs), ptr(defpot), ptr(ubig), nu,
                                                       _T(dt)(dx), n, iopt, ng or n, sumorder, C.byref(st))
    assert rc == 0, rc
    return st


def or_calcdefp(defp, tmp, tmp2, defn, u, n, dtold=1.0, dtaumesh=1.0, nfluid=5, compressmax=10.0, ng=None,
                sumorder=0):
    dt = defp.dtype
    T = _T(dt)
    st = LimStats()
    rc = getattr(oracle(), "or_calcdefp_" + _sfx(dt))(ptr(defp), ptr(tmp), ptr(tmp2), ptr(defn), ptr(u), T(dtold),
                                                      T(dtaumesh), nfluid, n, T(compressmax), ng or n, sumorder,
                                                      C.byref(st))
    assert rc == 0, rc
    return st


def or_set_mesh_solve(mode, tol=1e-8, nu_pre=2, nu_post=2, nu_coarse=32, max_cycles=100):
    """calcdefp's mesh solve in the oracle: 0 the reference's one V-cycle, 1 V-cycles of a fixed
    schedule to the relative residual tol."""
    oracle().or_set_mesh_solve(mode, C.c_double(tol), nu_pre, nu_post, nu_coarse, max_cycles)


def or_mesh_cycles():
    return oracle().or_mesh_cycles()


def or_relaxing(u, defn, defp, n, dt=1.0, nfluid=5):
    rc = getattr(oracle(), "or_relaxing_" + _sfx(u.dtype))(ptr(u), ptr(defn), ptr(defp), _T(u.dtype)(dt), nfluid, n)
    assert rc == 0, rc


def or_nsmooth(a, n):
    getattr(oracle(), "or_nsmooth_" + _sfx(a.dtype))(ptr(a), n)


def or_restrict(full, n, stride=1):
    half = np.zeros((n // 2) ** 3, full.dtype)
    getattr(oracle(), "or_restrict_" + _sfx(full.dtype))(ptr(half), ptr(full), n, stride)
    return half


def or_inject(u, uh, n):
    getattr(oracle(), "or_inject_" + _sfx(u.dtype))(ptr(u), ptr(uh), n)


def or_prolong(uh, n):
    u = np.zeros(n ** 3, uh.dtype)
    getattr(oracle(), "or_prolong_" + _sfx(uh.dtype))(ptr(u), ptr(uh), n)
    return u


def or_rlft3(data, speq, n, isign):
    getattr(oracle(), "or_rlft3_" + _sfx(data.dtype))(ptr(data), ptr(speq), n, n, n, isign)


def or_fft3(z, d0, d1, d2, s):
    """z: complex128/complex64 array of d0*d1*d2, in place."""
    sfx = "d" if z.dtype == np.complex128 else "f"
    getattr(oracle(), "or_fft3_" + sfx)(ptr(z), d0, d1, d2, s)


def or_gradient(phi, n):
    out = np.zeros(3 * n ** 3, phi.dtype)
    getattr(oracle(), "or_gradient_" + _sfx(phi.dtype))(ptr(phi), ptr(out), n)
    return out


def or_delta(phi, n):
    out = np.zeros(n ** 3, phi.dtype)
    getattr(oracle(), "or_delta_" + _sfx(phi.dtype))(ptr(phi), ptr(out), n)
    return out


def or_rgzerosum(a, defn, n, dx=1.0, sumorder=2):
    """rgzerosum (mmh/gauss1.f90:1219-1417) in place: a -= (sum a / sum det) det; returns (dsum, asum)."""
    sums = np.zeros(2)
    getattr(oracle(), "or_rgzerosum_" + _sfx(a.dtype))(ptr(a), ptr(defn), _T(a.dtype)(dx), n, sumorder, ptr(sums))
    return sums


def or_sum(x, sumorder):
    return getattr(oracle(), "or_sum_" + _sfx(x.dtype))(ptr(x), C.c_int64(x.size), sumorder)


def or_lognormal(n, sigma, seed, slope=-1.5, ks_frac=0.5):
    rho = np.zeros(n ** 3, np.float64)
    oracle().or_lognormal(ptr(rho), n, C.c_double(sigma), C.c_double(slope), C.c_double(ks_frac),
                          C.c_uint64(seed))
    return rho


# ---------------------------------------------------------------------------------------------
# reference (flang build) wrappers

def ref_path(ng, variant):
    return os.path.join(REF_DIR, "libmmhref_ng%d_%s.so" % (ng, variant))


def ref_available(ng, variant="f64"):
    return os.path.exists(ref_path(ng, variant))


_refs = {}


def ref(ng, variant="f64"):
    key = (ng, variant)
    if key not in _refs:
        lib = C.CDLL(ref_path(ng, variant))
        _refs[key] = lib
    return _refs[key]


def _I(v):
    return C.byref(C.c_int(v))


def ref_set_compressmax(lib, value, dt):
    # COMMON /globalpa/ omega0, omegal, omegab, compressmax, ... (mmh/globalpa.f90:6-9)
    T = _T(dt)
    arr = (T * 4).in_dll(lib, "globalpa_")
    arr[3] = value


def ref_relax(lib, arr, resid, rhs, defp, u, nu, dx, n, nrelax, iopt):
    T = _T(arr.dtype)
    rinf = T(0)
    lib.relax_(ptr(arr), ptr(resid), ptr(rhs), ptr(defp), ptr(u), C.byref(rinf), C.byref(T(dx)), _I(n), _I(n), _I(n),
               _I(nu), _I(nrelax), _I(iopt))
    return rinf.value


def ref_multigrid(lib, u, rhs, defpot, ubig, nu, dx, n, iopt=2):
    T = _T(u.dtype)
    lib.multigrid_(ptr(u), ptr(rhs), ptr(defpot), ptr(ubig), C.byref(T(dx)), _I(n), _I(n), _I(n), _I(nu), _I(1),
                   _I(iopt))


def ref_calcdefp(lib, defp, tmp, tmp2, defn, u, dtold=1.0, dtaumesh=1.0, nfluid=5):
    T = _T(defp.dtype)
    lib.calcdefp_(ptr(defp), ptr(tmp), ptr(tmp2), ptr(defn), ptr(u), C.byref(T(dtold)), C.byref(T(dtaumesh)),
                  _I(nfluid))


def ref_relaxing(lib, u, defn, defp, dt=1.0, nfluid=5):
    T = _T(u.dtype)
    lib.relaxing_(ptr(u), ptr(defn), ptr(defp), C.byref(T(0)), C.byref(T(dt)), _I(nfluid))


def ref_nsmooth(lib, a):
    lib.nsmooth_(ptr(a))


def ref_restrict(lib, full, n):
    half = np.zeros((n // 2) ** 3, full.dtype)
    lib.restrict_(ptr(half), ptr(full), _I(n), _I(n), _I(n))
    return half


def ref_inject(lib, u, uh, n):
    lib.inject_(ptr(u), ptr(uh), _I(n), _I(n), _I(n))


def ref_gradient(lib, phi, n):
    out = np.zeros(3 * n ** 3, np.float32)
    nn = np.array([n, n, n], np.int32)
    lib.scalar2gradient_(ptr(phi), ptr(out), ptr(nn))
    return out


def ref_delta(lib, phi, n):
    out = np.zeros(n ** 3, np.float32)
    nn = np.array([n, n, n], np.int32)
    lib.defp2delta_(ptr(phi), ptr(out), ptr(nn))
    return out


def or_solve(u, rhs, defpot, ubig, nu, n, tol=1e-8, max_cycles=100, schedule=0, nu_pre=2, nu_post=2, nu_coarse=32,
             ng=0, sumorder=1):
    relres = np.zeros(256)
    cyc = C.c_int(0)
    rc = oracle().or_solve_d(ptr(u), ptr(rhs), ptr(defpot), ptr(ubig), nu, n, C.c_double(tol), max_cycles, schedule,
                             nu_pre, nu_post, nu_coarse, ng, sumorder, ptr(relres), C.byref(cyc))
    return rc, relres[:cyc.value]


def or_limiter_rhs(defn, u, n, dtold=1.0, dtaumesh=1.0, nfluid=5, compressmax=10.0):
    tmp = np.zeros(n ** 3)
    rc = oracle().or_limiter_rhs_d(ptr(tmp), ptr(defn), ptr(u), C.c_double(dtold), C.c_double(dtaumesh), nfluid, n,
                                   C.c_double(compressmax))
    assert rc == 0, rc
    return tmp


def or_particle_density(xp, ng, nf):
    """Oracle of haloden.f90's density assignment: xp (np, 3) float32 -> (den ng^3, wp np)."""
    xp = np.ascontiguousarray(xp, dtype=np.float32)
    npart = xp.shape[0]
    den = np.zeros(ng ** 3, np.float32)
    wp = np.zeros(npart, np.int32)
    oracle().or_particle_density(ptr(xp), C.c_long(npart), ng, C.c_float(nf), ptr(den), ptr(wp))
    return den, wp


def or_coarsen(den, n, levels):
    den = np.ascontiguousarray(den, dtype=np.float32)
    out = np.zeros((n >> levels) ** 3, np.float32)
    oracle().or_coarsen(ptr(den), n, levels, ptr(out))
    return out


def or_column_density(den, defn, ng, nplot):
    den = np.ascontiguousarray(den, dtype=np.float32)
    defn = np.ascontiguousarray(defn, dtype=np.float32)
    proj = np.zeros(ng * ng, np.float32)
    oracle().or_column_density(ptr(den), ptr(defn), ng, nplot, ptr(proj))
    return proj


def or_gridfigure(defn, nf, den, ng, nplot=1):
    """Oracle of gridfigure.f90:41-131: (pos (2 nf^2), proj (ng^2), dsp_slab (2 (nf+1)^2 4))."""
    defn = np.ascontiguousarray(defn, dtype=np.float32)
    den = np.ascontiguousarray(den, dtype=np.float32)
    pos = np.zeros(2 * nf * nf, np.float32)
    proj = np.zeros(ng * ng, np.float32)
    slab = np.zeros(2 * (nf + 1) ** 2 * 4, np.float32)
    rc = oracle().or_gridfigure(ptr(defn), nf, ptr(den), ng, nplot, ptr(pos), ptr(proj), ptr(slab))
    assert rc == 0, rc
    return pos, proj, slab
