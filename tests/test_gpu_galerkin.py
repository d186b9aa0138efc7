"""GPU: the Galerkin coarse operators of the Krylov preconditioner against the CPU statement
(galerkin_ref: R A P from the oracle's metric), and the preconditioned solve with them."""
import numpy as np
import pytest

import galerkin_ref as GR
import refbind as R

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def smooth_def(n, amp, seed):
    d = np.random.default_rng(seed).standard_normal(n ** 3)
    for _ in range(4):
        R.or_nsmooth(d, n)
    return d * amp


@pytest.mark.parametrize("n,amp", [(16, 0.0), (16, 2.0), (32, 2.0)])
def test_galerkin_operators_match_rap(ctx, n, amp):
    """Levels 1.. of the device build (k_rap19, then k_rap27) against R A P: FP64 sums in another
    order, so to 1e-12 of the largest coefficient."""
    print("case galerkin operators n=%d def amp %.1f" % (n, amp), flush=True)
    rho = R.or_lognormal(n, 2.0, 7)
    d = smooth_def(n, amp, 4)
    nl = int(np.log2(n)) - 2
    ref = GR.galerkin_levels(d, rho, 1, n, 1.0, nl)
    for level in range(1, nl + 1):
        m = n >> level
        A = torch.zeros(27 * m ** 3, dtype=torch.float64, device="cuda")
        ctx.galerkin(A, dev(d), dev(rho), level, nu=1)
        got = A.cpu().numpy().reshape(27, m ** 3)
        want = GR.stencil_fields(ref[level - 1], m)
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max(), level


@pytest.mark.parametrize("mixed", [0, 1])
def test_galerkin_pcg(ctx, mixed):
    """At 64^3 (sigma 2, a smooth deformation) the PCG with Galerkin coarse levels converges to the
    same potential as with the rebuilt ones, in fewer iterations."""
    n = 64
    print("case galerkin PCG n=%d mixed=%d" % (n, mixed), flush=True)
    rho = R.or_lognormal(n, 2.0, 7)
    u = np.zeros((n ** 3, 5))
    u[:, 0] = rho
    d = smooth_def(n, 2.0, 4)
    rhs = (1 - rho) / 30.0
    out = []
    for gal in (1, 0):
        ctx.set_option("galerkin", gal)
        try:
            phi, b = dev(np.zeros(n ** 3)), dev(rhs)
            st, ok = ctx.solve(phi, b, dev(d), dev(u.ravel()), tol=1e-10, max_cycles=200, schedule=1, nu_pre=2,
                               nu_post=2, nu_coarse=8, mixed=mixed, nu=5, krylov=True)
        finally:
            ctx.set_option("galerkin", 1)
        assert ok
        out.append((phi.cpu().numpy(), st.cycles))
    (a, ca), (b, cb) = out
    print("iterations galerkin %d rebuilt %d" % (ca, cb), flush=True)
    assert ca < cb
    assert np.abs(a - b).max() <= 1e-7 * np.abs(b).max()
