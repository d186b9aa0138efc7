"""The allocation-free multigrid entry point (mmh_multigrid_noalloc_d) against mmh_multigrid_d,
through the C ABI: two calls in a row, the second on a smaller grid, each bit-identical to the
allocating entry point (iterate, projected rhs, per-level norms and guard decisions); sizes beyond
the scratch array's maximum are refused."""
import ctypes as C

import numpy as np
import pytest

import refbind as R

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def case(n, seed):
    rho = R.or_lognormal(n, 2.0, seed)
    u = np.zeros((n ** 3, 5))
    u[:, 0] = rho
    rng = np.random.default_rng(seed)
    d = rng.standard_normal(n ** 3)
    for _ in range(4):
        R.or_nsmooth(d, n)
    rhs = 1 - rho
    R.or_nsmooth(rhs, n)
    return u.ravel(), d * 2.0, rhs


def run(ctx, fn, u, d, rhs, n, ng):
    import mmh_amd
    a, b = dev(np.zeros(n ** 3)), dev(rhs)
    st = mmh_amd.MgStats()
    rc = getattr(mmh_amd.lib(), fn)(ctx._c, C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()),
                                    C.c_void_p(dev(d).data_ptr()), C.c_void_p(dev(u).data_ptr()), C.c_double(1.0),
                                    n, 5, 2, ng, C.byref(st))
    torch.cuda.synchronize()
    return rc, a.cpu().numpy(), b.cpu().numpy(), st


def test_multigrid_noalloc_equals_multigrid(ctx):
    import mmh_amd
    for n, ng, seed in ((64, 64, 3), (32, 64, 5), (16, 16, 7)):
        print("case no-alloc multigrid n=%d ng=%d" % (n, ng), flush=True)
        u, d, rhs = case(n, seed)
        rc0, a0, b0, s0 = run(ctx, "mmh_multigrid_d", u, d, rhs, n, ng)
        rc1, a1, b1, s1 = run(ctx, "mmh_multigrid_noalloc_d", u, d, rhs, n, ng)
        assert rc0 == mmh_amd.MMH_OK and rc1 == mmh_amd.MMH_OK
        assert np.array_equal(a1, a0)
        assert np.array_equal(b1, b0)
        assert s1.nlev == s0.nlev
        for l in range(s0.nlev):
            assert (s1.r1[l], s1.relaxerr[l], s1.guard[l]) == (s0.r1[l], s0.relaxerr[l], s0.guard[l])
    u, d, rhs = case(128, 1)
    rc, _, _, _ = run(ctx, "mmh_multigrid_noalloc_d", u, d, rhs, 128, 128)
    assert rc == mmh_amd.MMH_EINVAL
    assert "n must be a power of two in [4, 64]" in mmh_amd.lib().mmh_error_string(ctx._c).decode()
