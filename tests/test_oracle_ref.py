"""The oracle pinned against the reference itself (CPU only).

Two sources, same checks:
  * oracle/_ref/libmmhref_ng<NG>_<f32|f64>.so -- the reference Fortran compiled by
    oracle/build_ref.sh (FP32 as shipped; FP64 via -fdefault-real-8), where it has been built;
  * tests/golden/*.npz -- outputs of those same builds written by tests/golden/make_golden.py,
    for machines without the reference sources or flang.

Parity status per oracle function (the oracle with sequential sums, sumorder=0, the order the
reference runs with one OpenMP thread):
  relax (metric, 8-colour GS, residual, calcresid restriction)  bit-exact FP32 and FP64
  multigrid (V-cycle, guards, rgzerosum, zerosum, transfers)      bit-exact FP32 and FP64
  calcdefp without the FFT limiter branch (ncompress = 0)         bit-exact FP32 and FP64
  calcdefp with the FFT limiter branch (ncompress > 0)            rel 1e-13 FP64, 1e-5 FP32: the
      reference FFT rounds its twiddles to FP32 (recipes12.f90:54-55,125); ours is exact
  relaxing (relaxing-TVD)                                         bit-exact FP32 and FP64
  nsmooth, restrict, inject                                       bit-exact FP32 and FP64
  scalar2gradient, defp2delta (reference is real(4) only)         rel 1e-5 against FP32
The reference's mod(iopt,8) = 1 relax mode leaves its residual buffer uninitialised
(gauss1.f90:645-724 computes rt only inside the norm branch), so that mode is not compared."""
import glob
import os

import numpy as np
import pytest

import refbind as R

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


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


VARIANTS = [(16, "f64"), (16, "f32"), (32, "f64"), (32, "f32")]


def need_ref(ng, var):
    if not R.ref_available(ng, var):
        pytest.skip("reference build %d/%s not present (oracle/build_ref.sh)" % (ng, var))
    return R.ref(ng, var)


@pytest.mark.parametrize("ng,var", VARIANTS)
@pytest.mark.parametrize("nrelax,iopt", [(1, 19), (2, 19), (3, 19), (4, 19), (6, 19), (8, 18), (4, 11), (4, 0)])
def test_relax_vs_reference(ng, var, nrelax, iopt):
    lib = need_ref(ng, var)
    dt = np.float64 if var == "f64" else np.float32
    n = ng
    rho = R.or_lognormal(n, 1.0, 1234).astype(dt)
    u = aos(rho)
    d = smooth_def(n, 1.5, 3, dt)
    rhs = np.random.default_rng(1).standard_normal(n ** 3).astype(dt)
    a1, r1 = np.zeros(n ** 3, dt), np.zeros((n // 2) ** 3, dt)
    a2, r2 = a1.copy(), r1.copy()
    ri1 = R.ref_relax(lib, a1, r1, rhs, d, u, 5, 1.0, n, nrelax, iopt)
    ri2 = R.or_relax(a2, r2, rhs, d, u, 5, 1.0, n, nrelax, iopt)
    assert np.array_equal(a1, a2)
    assert np.array_equal(r1, r2)
    assert ri1 == ri2


@pytest.mark.parametrize("ng,var", VARIANTS + [(64, "f64"), (64, "f32")])
@pytest.mark.parametrize("sigma,amp", [(1.0, 0.0), (2.0, 2.0)])
def test_multigrid_vs_reference(ng, var, sigma, amp):
    lib = need_ref(ng, var)
    dt = np.float64 if var == "f64" else np.float32
    n = ng
    rho = R.or_lognormal(n, sigma, 7).astype(dt)
    u = aos(rho)
    d = smooth_def(n, amp, 4, dt) if amp else np.zeros(n ** 3, dt)
    rhs = (1 - rho).astype(dt)
    R.or_nsmooth(rhs, n)
    u1, b1 = np.zeros(n ** 3, dt), rhs.copy()
    u2, b2 = u1.copy(), rhs.copy()
    for _ in range(2):  # two cycles: the second starts from the first's solution
        R.ref_multigrid(lib, u1, b1, d, u, 5, 1.0, n)
        R.or_multigrid(u2, b2, d, u, 5, 1.0, n, 2, ng, 0)
    assert np.array_equal(b1, b2)
    assert np.array_equal(u1, u2)


@pytest.mark.parametrize("ng,var", VARIANTS)
@pytest.mark.parametrize("amp", [0.0, 2.0, 8.0])
def test_calcdefp_relaxing_vs_reference(ng, var, amp):
    lib = need_ref(ng, var)
    dt = np.float64 if var == "f64" else np.float32
    R.ref_set_compressmax(lib, 10.0, dt)
    n = ng
    rho = R.or_lognormal(n, 2.0, 7).astype(dt)
    u = aos(rho)
    d = smooth_def(n, amp, 3, dt) if amp else np.zeros(n ** 3, dt)
    p1, t1, q1 = np.zeros(n ** 3, dt), np.zeros(n ** 3, dt), np.zeros((n + 2) * n * n, dt)
    p2, t2, q2 = p1.copy(), t1.copy(), q1.copy()
    R.ref_calcdefp(lib, p1, t1, q1, d, u)
    st = R.or_calcdefp(p2, t2, q2, d, u, n)
    if st.ncompress == 0:
        assert np.array_equal(p1, p2)
    else:
        tol = 1e-13 if dt == np.float64 else 1e-5
        assert np.abs(p1 - p2).max() <= tol * np.abs(p1).max()
    assert np.array_equal(t1[3:], t2[3:])
    d2 = (d + p1).astype(dt)
    ua, ub = u.copy(), u.copy()
    R.ref_relaxing(lib, ua, d2, p1)
    R.or_relaxing(ub, d2, p1, n)
    assert np.array_equal(ua, ub)


@pytest.mark.parametrize("ng", [16, 32])
def test_small_ops_vs_reference(ng):
    for var in ("f64", "f32"):
        lib = need_ref(ng, var)
        dt = np.float64 if var == "f64" else np.float32
        n = ng
        a = smooth_def(n, 1.0, 9, dt)
        b1, b2 = a.copy(), a.copy()
        R.ref_nsmooth(lib, b1)
        R.or_nsmooth(b2, n)
        assert np.array_equal(b1, b2)
        h1, h2 = R.ref_restrict(lib, a, n), R.or_restrict(a, n)
        assert np.array_equal(h1, h2)
        u1, u2 = a.copy(), a.copy()
        R.ref_inject(lib, u1, h1, n)
        R.or_inject(u2, h1, n)
        assert np.array_equal(u1, u2)


@pytest.mark.parametrize("ng", [16, 32])
def test_gradient_delta_vs_reference_fp32(ng):
    lib = need_ref(ng, "f32")
    n = ng
    phi = smooth_def(n, 3.0, 12, np.float32)
    g1, d1 = R.ref_gradient(lib, phi, n), R.ref_delta(lib, phi, n)
    g2, d2 = R.or_gradient(phi.astype(np.float64), n), R.or_delta(phi.astype(np.float64), n)
    assert np.abs(g2 - g1).max() <= 1e-5 * np.abs(g1).max()
    assert np.abs(d2 - d1).max() <= 1e-5 * np.abs(d1).max()


# ---------------------------------------------------------------------------------------------
# committed golden vectors (written by tests/golden/make_golden.py from the reference builds)

GOLDEN_FILES = sorted(glob.glob(os.path.join(GOLDEN, "*.npz")))


@pytest.mark.parametrize("path", GOLDEN_FILES, ids=[os.path.basename(p) for p in GOLDEN_FILES])
def test_golden(path):
    z = np.load(path, allow_pickle=False)
    kind = str(z["kind"])
    n = int(z["n"])
    ng = int(z["ng"])
    dt = z["def"].dtype
    u = z["u"] if "u" in z.files else None
    d = z["def"]
    if kind == "multigrid":
        a, b = np.zeros(n ** 3, dt), z["rhs_in"].copy()
        for _ in range(int(z["cycles"])):
            R.or_multigrid(a, b, d, u, 5, 1.0, n, 2, ng, 0)
        assert np.array_equal(a, z["phi_out"])
        assert np.array_equal(b, z["rhs_out"])
    elif kind == "outer":
        # K reference outer iterations (main.f90:90-110): calcdefp, def += defp, relaxing
        uu, dd = u.copy(), d.copy()
        for _ in range(int(z["iters"])):
            p, t, q = np.zeros(n ** 3, dt), np.zeros(n ** 3, dt), np.zeros((n + 2) * n * n, dt)
            st = R.or_calcdefp(p, t, q, dd, uu, n, ng=ng)
            assert st.ncompress == 0
            dd = (dd + p).astype(dt)
            R.or_relaxing(uu, dd, p, n)
        assert np.array_equal(dd, z["def_out"])
        assert np.array_equal(uu, z["u_out"])
    elif kind == "multigrid_guard":
        # V-cycles whose divergence guard fires (mmh/mgtemplate.f90:46-55) on the levels the
        # reference reported diverged, 4^3 included (mmh/multigrid.f90:113-119)
        for var, dt in (("f64", np.float64), ("f32", np.float32)):
            u = np.zeros((n ** 3, 5), dt)
            u[:, 0] = z["rho"]
            u = u.ravel()
            a, b = np.zeros(n ** 3, dt), z["rhs_in"].astype(dt)
            for c in range(int(z["cycles_" + var])):
                st = R.or_multigrid(a, b, z["def"].astype(dt), u, 5, 1.0, n, 2, ng, 0)
                fired = sorted(n >> l for l in range(st.nlev) if st.guard[l])
                assert fired == list(z["diverged_levels_" + var][c])
            assert np.array_equal(a, z["phi_out_" + var])
            assert np.array_equal(b, z["rhs_out_" + var])
    elif kind == "fft":
        phi = z["phi"]
        g = R.or_gradient(phi.astype(np.float64), n)
        dl = R.or_delta(phi.astype(np.float64), n)
        assert np.abs(g - z["grad"]).max() <= 1e-5 * np.abs(z["grad"]).max()
        assert np.abs(dl - z["delta"]).max() <= 1e-5 * np.abs(z["delta"]).max()
    else:
        raise AssertionError(kind)
