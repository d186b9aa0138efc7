"""The drop-in boundary proven by a call: the reference-style driver of tests/dropin/ (implicit-
interface calls of multigrid, relax, calcdefp, relaxing, nsmooth, scalar2gradient and defp2delta with
the reference's argument lists, mmh/multigrid.f90:16, mmh/gauss1.f90:64, mmh/limiter.f90:17,677,
mmh/relaxing.f90:21, mmh/gradient.f90:1, mmh/reconstruct.f90:1) linked against libmmh_fortran.so,
against the same source linked against the reference's objects and against the oracle.

Bars: bit-exact against the oracle run with the device's summation order (sumorder 2), every output
including the FFT paths; against the reference's executable, bit-exact for relax (no global sums) and
within 1e-6 relative elsewhere (the reference sums rgzerosum/zerosum sequentially, FP32 outputs) and
1e-5 for the two spectral outputs (the reference's FFT rounds twiddles to FP32)."""
import os

import numpy as np
import pytest

import dropin_common as D

pytestmark = pytest.mark.gpu


def test_dropin_driver_on_gpu(tmp_path):
    path = D.driver_path("gpu")
    assert os.path.exists(path), "tests/dropin/build.sh has not been run"
    print("case drop-in driver ng=%d" % D.NG, flush=True)
    rho, d = D.inputs()
    (tmp_path / "gpu").mkdir()
    got = D.run_driver(path, rho, d, tmp_path / "gpu")
    want, ncomp = D.oracle_sequence(rho, d, sumorder=2)
    assert ncomp == [0, 0]
    for k in want:
        assert np.array_equal(got[k], want[k]), k
    ref = D.driver_path("ref")
    if not os.path.exists(ref):
        return
    (tmp_path / "ref").mkdir()
    r = D.run_driver(ref, rho, d, tmp_path / "ref")
    for k in ("relax_phi", "relax_resid", "relax_rinf"):
        assert np.array_equal(got[k], r[k]), k
    for k, tol in (("mg_phi", 1e-6), ("mg_rhs", 1e-6), ("def", 1e-6), ("defp", 1e-6), ("u1", 1e-6),
                   ("grad", 1e-5), ("delta", 1e-5)):
        err = np.abs(got[k] - r[k]).max() / np.abs(r[k]).max()
        print("  %s vs reference: rel %.2e" % (k, err))
        assert err <= tol, k
