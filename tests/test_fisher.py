"""calcfisherinfo-style statistics (mmh_fisher_d, host code in the library) against the
reference's own calccov / calerror / inverse subroutines compiled from mmh/calcfisherinfo.f90
by oracle/build_ref.sh, and against numpy (np.cov, np.linalg.inv)."""
import ctypes as C
import os

import numpy as np
import pytest

import mmh_amd
import refbind as R

FISHER_REF = os.path.join(R.REF_DIR, "libmmhref_fisher.so")


def data(nsim=24, nk=10, seed=0):
    rng = np.random.default_rng(seed)
    base = np.exp(-np.arange(nk) / 4.0) + 0.5
    pk = base * (1 + 0.1 * rng.standard_normal((nsim, nk)) + 0.05 * rng.standard_normal((nsim, 1)))
    c = 1 + 0.2 * rng.standard_normal(nk)
    return pk, c


def test_fisher_against_numpy():
    pk, c = data()
    out = mmh_amd.fisher(pk, c)
    cov = np.cov(pk.T, ddof=1)
    mean = pk.mean(0)
    assert np.allclose(out["cov"], cov, rtol=1e-12, atol=0)
    assert np.allclose(out["covnorm"], cov / np.outer(mean, mean), rtol=1e-12, atol=0)
    # calerror divides by a single-precision sqrt(n_sim) (calcfisherinfo.f90:276)
    assert np.allclose(out["error"], pk.std(0, ddof=1) / np.sqrt(pk.shape[0]), rtol=1e-7, atol=0)
    for m in range(1, pk.shape[1] + 1):
        cn = out["covnorm"][:m, :m]
        assert np.isclose(out["fisher"][m - 1], c[:m] @ np.linalg.solve(cn, c[:m]), rtol=1e-9)


def test_fisher_bad_arguments():
    with pytest.raises(mmh_amd.MmhError):
        mmh_amd.fisher(np.ones((1, 3)), np.ones(3))


@pytest.mark.skipif(not os.path.exists(FISHER_REF), reason="reference calcfisherinfo build absent")
def test_fisher_against_reference_subroutines():
    lib = C.CDLL(FISHER_REF)
    pk, c = data(nsim=17, nk=7, seed=3)
    nsim, nk = pk.shape
    out = mmh_amd.fisher(pk, c)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    ns = C.c_int(nsim)
    for x in range(nk):
        for y in range(nk):
            p1, p2 = np.ascontiguousarray(pk[:, x]), np.ascontiguousarray(pk[:, y])
            cv, cn = C.c_double(0), C.c_double(0)
            lib.calccov_(p(p1), p(p2), C.byref(ns), C.byref(cv), C.byref(cn))
            assert abs(cv.value - out["cov"][x, y]) <= 1e-14 * abs(cv.value)
            assert abs(cn.value - out["covnorm"][x, y]) <= 1e-14 * abs(cn.value)
        err = C.c_double(0)
        lib.calerror_(p(np.ascontiguousarray(pk[:, x])), C.byref(err), C.byref(ns))
        assert abs(err.value - out["error"][x]) <= 1e-14 * err.value
    for m in range(1, nk + 1):
        a = np.asfortranarray(out["covnorm"][:m, :m])
        ia = np.zeros((m, m), order="F")
        lib.inverse_(p(a), p(ia), C.byref(C.c_int(m)))
        # the reference's cumulative sum (calcfisherinfo.f90:122-131)
        t = [sum(ia[i, j] * c[i] for i in range(m)) for j in range(m)]
        f = sum(t[j] * c[j] for j in range(m))
        assert abs(f - out["fisher"][m - 1]) <= 1e-12 * abs(f)
