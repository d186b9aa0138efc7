"""The reference-style driver of tests/dropin/ linked against the reference's own objects
(oracle/_ref/dropin_ref_ng32, FP32 as shipped) reproduces the oracle's restatement of the same calls:
the harness that the GPU drop-in test compares against is itself pinned here, on the CPU.
Bit-exact except scalar2gradient and defp2delta, whose FFT the reference runs with twiddles rounded
to FP32 (mmh/recipes12.f90:54-55): rel 1e-5."""
import os

import numpy as np
import pytest

import dropin_common as D


def test_reference_driver_matches_oracle(tmp_path):
    path = D.driver_path("ref")
    if not os.path.exists(path):
        pytest.skip("reference drop-in driver not built (oracle/build_ref.sh)")
    rho, d = D.inputs()
    got = D.run_driver(path, rho, d, tmp_path)
    want, ncomp = D.oracle_sequence(rho, d, sumorder=0)
    assert ncomp == [0, 0]
    for k in ("relax_phi", "relax_resid", "relax_rinf", "mg_phi", "mg_rhs", "def", "defp", "u1"):
        assert np.array_equal(got[k], want[k]), k
    for k in ("grad", "delta"):
        assert np.abs(got[k] - want[k]).max() <= 1e-5 * np.abs(want[k]).max(), k
