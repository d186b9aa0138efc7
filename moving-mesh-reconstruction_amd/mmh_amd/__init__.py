"""Python host layer of the MI355X moving-mesh reconstruction library.

Loads the in-tree libmmh_amd.so (HIP, gfx950) and exposes the reference's operator interface on
device arrays (torch tensors on a ROCm device, or raw device pointers):

    relax(arr, resid, rhs, deformpot, u, dx, nrelax, iopt)      mmh/gauss1.f90:64
    multigrid(u, rhs, defpot, ubig, dx, iopt, ng)                mmh/multigrid.f90:16
    calcdefp(defp, tmp, tmp2, def_, u, dtold1, dtaumesh, nfluid) mmh/limiter.f90:17
    relaxing(u, def_, defp, cmax, dt, nfluidcmp)                 mmh/relaxing.f90:21
    scalar2gradient(indata, outdata)                             mmh/gradient.f90:1
    defp2delta(indata, outdata)                                  mmh/reconstruct.f90:1
    nsmooth(a)                                                   mmh/limiter.f90:677

Arguments keep the reference's meaning and in/out behaviour (multigrid projects rhs in place,
calcdefp overwrites tmp/tmp2).  Sizes come from the tensors instead of the compile-time NG.  Errors
raise MmhError carrying the library's message; there is no CPU fallback.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_DIR, "libmmh_amd.so")
FLIB_PATH = os.path.join(PKG_DIR, "libmmh_fortran.so")

MMH_OK, MMH_EINVAL, MMH_EDEVICE, MMH_ENOMEM, MMH_ENOCONV = 0, 1, 2, 3, 4
MAXLEV = 16


class MmhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("mmh error %d: %s" % (code, msg))
        self.code = code


class MgStats(C.Structure):
    _fields_ = [("nlev", C.c_int), ("n", C.c_int * MAXLEV), ("r1", C.c_double * MAXLEV),
                ("relaxerr", C.c_double * MAXLEV), ("guard", C.c_int * MAXLEV), ("entry_zerosum", C.c_int)]


class LimStats(C.Structure):
    _fields_ = [("cmax", C.c_double), ("detmin", C.c_double), ("detmax", C.c_double), ("fluxmin", C.c_double),
                ("fluxmax", C.c_double), ("nexpand", C.c_int), ("ncompress", C.c_int), ("mg", MgStats)]


class IterStats(C.Structure):
    """One outer iteration's diagnostics (the values mmh/main.f90:97-106 prints) and limiter record."""
    _fields_ = [("iter", C.c_int), ("defp_sigma", C.c_double), ("def_sigma", C.c_double), ("u_sigma", C.c_double),
                ("u_min", C.c_double), ("u_max", C.c_double), ("u_negative", C.c_longlong), ("lim", LimStats)]


class AdjointStats(C.Structure):
    _fields_ = [("loss", C.c_double), ("cycles_forward", C.c_int), ("cycles_adjoint", C.c_int),
                ("relres_forward", C.c_double), ("relres_adjoint", C.c_double), ("ncompress", C.c_int)]


class BatchStats(C.Structure):
    _fields_ = [("cycles", C.c_int), ("relres", C.c_double), ("converged", C.c_int)]


class SolveOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_cycles", C.c_int), ("schedule", C.c_int), ("nu_pre", C.c_int),
                ("nu_post", C.c_int), ("nu_coarse", C.c_int), ("ng", C.c_int), ("mixed", C.c_int),
                ("krylov", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("cycles", C.c_int), ("relres", C.c_double * 256), ("rhs_norm", C.c_double),
                ("nonneg_pLp", C.c_int), ("nonneg_zr", C.c_int), ("galerkin", C.c_int)]


_lib = None


def lib():
    """The loaded libmmh_amd.so.  Raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is None:
        # One HIP runtime per process: PyTorch wheels carry their own libamdhip64, so it is loaded
        # first and the library binds to the same runtime torch's device memory comes from.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise MmhError(MMH_EDEVICE, "libmmh_amd.so not built: run moving-mesh-reconstruction_amd/build.py")
        L = C.CDLL(LIB_PATH)
        L.mmh_error_string.restype = C.c_char_p
        L.mmh_default_ctx.restype = C.c_void_p
        _lib = L
    return _lib


def _ptr(x):
    """Device pointer of a torch tensor (or an int address)."""
    if isinstance(x, int):
        return C.c_void_p(x)
    if x is None:
        return C.c_void_p(0)
    if not x.is_contiguous():
        raise MmhError(MMH_EINVAL, "arrays must be contiguous")
    return C.c_void_p(x.data_ptr())


def _sfx(t):
    import torch
    if t.dtype == torch.float64:
        return "d", C.c_double
    if t.dtype == torch.float32:
        return "f", C.c_float
    raise MmhError(MMH_EINVAL, "dtype must be float64 or float32")


def _cube_side(t, comps=1):
    n3 = t.numel() // comps
    n = round(n3 ** (1.0 / 3.0))
    if n ** 3 * comps != t.numel():
        raise MmhError(MMH_EINVAL, "array is not an N^3 field")
    return n


class Context:
    """A library context bound to one device (one HIP stream, cached work buffers)."""

    def __init__(self, device=0):
        self._c = C.c_void_p()
        rc = lib().mmh_create(C.byref(self._c), device)
        if rc != MMH_OK:
            raise MmhError(rc, "mmh_create failed (no HIP device %d?)" % device)

    def close(self):
        if self._c:
            lib().mmh_destroy(self._c)
            self._c = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != MMH_OK:
            raise MmhError(rc, lib().mmh_error_string(self._c).decode())

    def call(self, name, *args):
        self._check(getattr(lib(), name)(self._c, *args))

    def set_stream(self, stream_ptr):
        self.call("mmh_set_stream", C.c_void_p(stream_ptr))

    def set_option(self, key, value):
        self.call("mmh_set_option", key.encode(), int(value))

    def get_option(self, key):
        v = C.c_int(0)
        self.call("mmh_get_option", key.encode(), C.byref(v))
        return v.value

    def mesh_cycles(self):
        """Cycles (or Krylov iterations) of the last mesh solve inside calcdefp (option mesh_solve)."""
        return self.get_option("mesh_cycles")

    def release_workspace(self):
        self.call("mmh_release_workspace")

    # ---- reference operator interface ------------------------------------------------------
    def relax(self, arr, resid, rhs, deformpot, u, dx, nrelax, iopt, nu=None):
        s, T = _sfx(arr)
        n = _cube_side(arr)
        nu = nu or u.numel() // arr.numel()
        rinf = T(0)
        self.call("mmh_relax_" + s, _ptr(arr), _ptr(resid), _ptr(rhs), _ptr(deformpot), _ptr(u), nu, C.byref(rinf),
                  T(dx), n, nrelax, iopt)
        return rinf.value

    def multigrid(self, u, rhs, defpot, ubig, dx=1.0, iopt=2, ng=0, nu=None):
        s, T = _sfx(u)
        n = _cube_side(u)
        nu = nu or ubig.numel() // u.numel()
        st = MgStats()
        self.call("mmh_multigrid_" + s, _ptr(u), _ptr(rhs), _ptr(defpot), _ptr(ubig), T(dx), n, nu, iopt, ng,
                  C.byref(st))
        return st

    def solve(self, u, rhs, defpot, ubig, tol=1e-8, max_cycles=100, schedule=0, nu_pre=2, nu_post=2, nu_coarse=32,
              ng=0, mixed=False, nu=None, krylov=False):
        n = _cube_side(u)
        nu = nu or ubig.numel() // u.numel()
        o = SolveOpts(tol, max_cycles, schedule, nu_pre, nu_post, nu_coarse, ng, int(mixed), 1 if krylov else 0)
        st = SolveStats()
        rc = lib().mmh_solve_d(self._c, _ptr(u), _ptr(rhs), _ptr(defpot), _ptr(ubig), n, nu, C.byref(o), C.byref(st))
        if rc not in (MMH_OK, MMH_ENOCONV):
            self._check(rc)
        return st, rc == MMH_OK

    def solve_batch(self, items, tol=1e-8, max_cycles=200, nu_pre=2, nu_coarse=8, nu=5):
        """items: list of (u, rhs, defpot, ubig) device tensors (sizes may differ).  One batched FP64
        PCG solve; returns (list of BatchStats, all_converged)."""
        k = len(items)
        ns = (C.c_int * k)(*[_cube_side(it[0]) for it in items])
        arr = [(C.c_void_p * k)(*[_ptr(it[j]).value for it in items]) for j in range(4)]
        o = SolveOpts(tol, max_cycles, 1, nu_pre, nu_pre, nu_coarse, 0, 0, 1)
        st = (BatchStats * k)()
        rc = lib().mmh_solve_batch_d(self._c, k, ns, arr[0], arr[1], arr[2], arr[3], nu, C.byref(o), st)
        if rc not in (MMH_OK, MMH_ENOCONV):
            self._check(rc)
        return list(st), rc == MMH_OK

    def calcdefp(self, defp, tmp, tmp2, def_, u, dtold1=1.0, dtaumesh=1.0, nfluid=5, compressmax=10.0, ng=0):
        s, T = _sfx(defp)
        n = _cube_side(defp)
        st = LimStats()
        self.call("mmh_calcdefp_" + s, _ptr(defp), _ptr(tmp), _ptr(tmp2), _ptr(def_), _ptr(u), T(dtold1), T(dtaumesh),
                  nfluid, n, T(compressmax), ng, C.byref(st))
        return st

    def limiter_rhs(self, tmp, def_, u, dtold1=1.0, dtaumesh=1.0, nfluid=5, compressmax=10.0):
        n = _cube_side(tmp)
        self.call("mmh_limiter_rhs_d", _ptr(tmp), _ptr(def_), _ptr(u), C.c_double(dtold1), C.c_double(dtaumesh), nfluid,
                  n, C.c_double(compressmax))

    def reconstruct(self, u, def_, defp, niter, dtold=1.0, dtaumesh=1.0, dt=1.0, compressmax=10.0, nfluid=5):
        """niter outer iterations (calcdefp, def += defp, relaxing; mmh/main.f90:89-110) on device
        tensors; returns one IterStats per iteration."""
        s, T = _sfx(u)
        n = _cube_side(def_)
        st = (IterStats * max(niter, 1))()
        self.call("mmh_reconstruct_" + s, _ptr(u), _ptr(def_), _ptr(defp), n, nfluid, niter, T(dtold), T(dtaumesh),
                  T(dt), T(compressmax), st)
        return list(st)[:niter]

    def residual(self, r, phi, rhs, def_, u, dx=1.0, nu=None):
        """r = rhs - L(def, u) phi at every point, FP64 (calcresid's formula)."""
        n = _cube_side(phi)
        nu = nu or u.numel() // phi.numel()
        self.call("mmh_residual_d", _ptr(r), _ptr(phi), _ptr(rhs), _ptr(def_), _ptr(u), nu, n, C.c_double(dx))

    def galerkin(self, A, def_, u, level, dx=1.0, nu=None):
        """Galerkin coarse operator of the Krylov preconditioner at `level` (27 fields of
        (n >> level)^3, Fortran order each) from the fine operator of def and u(1)."""
        n = _cube_side(def_)
        nu = nu or u.numel() // def_.numel()
        self.call("mmh_galerkin_d", _ptr(A), _ptr(def_), _ptr(u), nu, n, C.c_double(dx), level)

    def particle_density(self, xp, ng, nf, den, wp=None):
        """haloden.f90's density assignment: xp device float32 (np, 3) positions in [0, nf);
        den device float32 ng^3 (Fortran order) receives the mean-1 density."""
        npart = xp.numel() // 3
        self.call("mmh_particle_density_f", _ptr(xp), C.c_longlong(npart), ng, C.c_float(nf), _ptr(den), _ptr(wp))

    def coarsen(self, den, n, levels, out):
        """from1024to512.f90: levels rounds of 2x2x2 block sums, then mean 1 (float32 device)."""
        self.call("mmh_coarsen_f", _ptr(den), n, levels, _ptr(out))

    def column_density(self, den, def_, ng, nplot, proj):
        """column_density.f90: density integrated over the deformed column (float32 device)."""
        self.call("mmh_column_density_f", _ptr(den), _ptr(def_), ng, nplot, _ptr(proj))

    def gridfigure(self, def_, den, nplot, pos, proj, dsp_slab=None):
        """gridfigure.f90:41-131 (FP32 device tensors): def_ nf^3, den ng^3 -> pos (2 nf^2), proj (ng^2),
        optionally dsp_slab (2 (nf+1)^2 4)."""
        nf, ng = _cube_side(def_), _cube_side(den)
        self.call("mmh_gridfigure_f", _ptr(def_), nf, _ptr(den), ng, nplot, _ptr(pos), _ptr(proj),
                  _ptr(dsp_slab) if dsp_slab is not None else None)

    def power(self, field, nbins=16):
        """Binned power spectrum of an N^3 FP64 device field: (P_b, mode counts) as numpy arrays."""
        import numpy as np
        n = _cube_side(field)
        pk, nm = np.zeros(nbins), np.zeros(nbins)
        self.call("mmh_power_d", _ptr(field), n, nbins, pk.ctypes.data_as(C.c_void_p), nm.ctypes.data_as(C.c_void_p))
        return pk, nm

    def recon_power_grad(self, rho, weights, grad=None, tol=1e-12):
        """loss = sum_b w_b P_b(delta_rec(rho)) for the first outer iteration, its binned spectrum
        and (if grad is a device tensor) dloss/drho written into grad.  Returns (loss, pk, stats)."""
        import numpy as np
        n = _cube_side(rho)
        w = np.ascontiguousarray(weights, dtype=np.float64)
        pk = np.zeros(w.size)
        st = AdjointStats()
        self.call("mmh_recon_power_grad_d", _ptr(rho), n, w.size, w.ctypes.data_as(C.c_void_p), C.c_double(tol),
                  pk.ctypes.data_as(C.c_void_p), _ptr(grad), C.byref(st))
        return st.loss, pk, st

    def recon_power_grad_mesh(self, rho, def_, weights, grad_rho=None, grad_def=None, tol=1e-12):
        """The loss of recon_power_grad for the outer iteration that starts from the mesh potential
        def_ (held fixed): delta = defp2delta(def_ + defp).  dloss/drho into grad_rho and dloss/d(def)
        into grad_def where given.  Returns (loss, pk, stats)."""
        import numpy as np
        n = _cube_side(rho)
        w = np.ascontiguousarray(weights, dtype=np.float64)
        pk = np.zeros(w.size)
        st = AdjointStats()
        self.call("mmh_recon_power_grad_mesh_d", _ptr(rho), _ptr(def_), n, w.size, w.ctypes.data_as(C.c_void_p),
                  C.c_double(tol), pk.ctypes.data_as(C.c_void_p), _ptr(grad_rho), _ptr(grad_def), C.byref(st))
        return st.loss, pk, st

    def relaxing(self, u, def_, defp, cmax=0.0, dt=1.0, nfluidcmp=5):
        s, T = _sfx(u)
        n = _cube_side(def_)
        self.call("mmh_relaxing_" + s, _ptr(u), _ptr(def_), _ptr(defp), T(dt), nfluidcmp, n)

    def scalar2gradient(self, indata, outdata):
        s, _ = _sfx(indata)
        self.call("mmh_gradient_" + s, _ptr(indata), _ptr(outdata), _cube_side(indata))

    def defp2delta(self, indata, outdata):
        s, _ = _sfx(indata)
        self.call("mmh_delta_" + s, _ptr(indata), _ptr(outdata), _cube_side(indata))

    def nsmooth(self, a):
        s, _ = _sfx(a)
        self.call("mmh_nsmooth_" + s, _ptr(a), _cube_side(a))

    def restrict(self, half, full, stride=1):
        s, _ = _sfx(full)
        self.call("mmh_restrict_" + s, _ptr(half), _ptr(full), _cube_side(full, stride), stride)

    def inject(self, u, uh):
        s, _ = _sfx(u)
        self.call("mmh_inject_" + s, _ptr(u), _ptr(uh), _cube_side(u))

    def metric(self, G, def_, u, dx=1.0, lsquare=0, nu=None):
        s, T = _sfx(def_)
        n = _cube_side(def_)
        nu = nu or u.numel() // def_.numel()
        self.call("mmh_metric_" + s, _ptr(G), _ptr(def_), _ptr(u), nu, n, T(dx), lsquare)

    def time_sweeps(self, phi, rhs, G, dx=1.0, nsweeps=10):
        s, T = _sfx(phi)
        ms = C.c_float(0)
        self.call("mmh_time_sweeps_" + s, _ptr(phi), _ptr(rhs), _ptr(G), _cube_side(phi), T(dx), nsweeps, C.byref(ms))
        return ms.value

    def time_sweeps_bf16(self, phi, rhs, G, dx=1.0, nsweeps=10):
        """FP32 sweeps (phi, rhs float32) with G (FP64, 6 n^3) rounded to bfloat16; ms per sweep."""
        ms = C.c_float(0)
        self.call("mmh_time_sweeps_fb", _ptr(phi), _ptr(rhs), _ptr(G), _cube_side(phi), C.c_float(dx), nsweeps,
                  C.byref(ms))
        return ms.value

    def rlft3(self, data, speq, isign):
        self.call("mmh_rlft3_d", _ptr(data), _ptr(speq), _cube_side(data), isign)

    def lognormal(self, rho, sigma, seed, slope=-1.5, ks_frac=0.5):
        self.call("mmh_lognormal_d", _ptr(rho), _cube_side(rho), C.c_double(sigma), C.c_double(slope),
                  C.c_double(ks_frac), C.c_uint64(seed))


def fisher(pk, c):
    """calcfisherinfo-style statistics of nsim x nk binned spectra pk and derivative vector c:
    dict with cov, covnorm (nk x nk), error (nk) and the cumulative Fisher information (nk).
    Host computation (mmh_fisher_d); needs no device."""
    import numpy as np
    pk = np.ascontiguousarray(pk, dtype=np.float64)
    c = np.ascontiguousarray(c, dtype=np.float64)
    nsim, nk = pk.shape
    cov, covn = np.zeros(nk * nk), np.zeros(nk * nk)
    fis, err = np.zeros(nk), np.zeros(nk)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    rc = lib().mmh_fisher_d(vp(pk), nsim, nk, vp(c), vp(cov), vp(covn), vp(fis), vp(err))
    if rc != MMH_OK:
        raise MmhError(rc, "mmh_fisher_d: bad arguments")
    return {"cov": cov.reshape(nk, nk, order="F"), "covnorm": covn.reshape(nk, nk, order="F"), "fisher": fis,
            "error": err}


def header_symbols(header=None):
    """Function names declared in include/mmh.h."""
    import re
    header = header or os.path.join(os.path.dirname(PKG_DIR), "include", "mmh.h")
    txt = open(header).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mmh_[a-z0-9_]+)\s*\(", txt)))
