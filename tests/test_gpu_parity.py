"""GPU parity: the HIP path against the oracle on the same seeded inputs.

Integer-exact bar for floating point: every kernel keeps the reference's association order and
the library is built without contraction, so FP64 and FP32 results must be bit-identical to the
oracle run with the sum order the device reductions use (sumorder=2: blocked tree sums, with the
per-level mean removal summing the level arrays in their sub-lattice storage order).  Exceptions,
each with its tolerance in the test: the limiter's eigenvalue gate uses device acos/cos
(calcdefp with ncompress > 0: rel 1e-12 FP64 / 1e-5 FP32), and the lognormal generator uses device
log/cos/exp (rel 1e-12)."""
import os

import numpy as np
import pytest

import refbind as R

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


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


def field(n, sigma, seed, dt):
    return R.or_lognormal(n, sigma, seed).astype(dt)


DT = [np.float64, np.float32]


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("n", [8, 16, 32])
def test_metric_and_sweeps_bitexact(ctx, dt, n):
    print("case metric/sweeps n=%d %s" % (n, dt.__name__), flush=True)
    rho = field(n, 1.5, 11, dt)
    u = aos(rho)
    d = smooth_def(n, 2.0, 3, dt)
    G = R.or_metric(d, u, 5, n, 1.0)
    Gd = dev(np.zeros(6 * n ** 3, dt))
    ctx.metric(Gd, dev(d), dev(u), 1.0, 0, nu=5)
    assert np.array_equal(host(Gd), G)
    rhs = (1 - rho).astype(dt)
    R.or_nsmooth(rhs, n)
    phi = np.zeros(n ** 3, dt)
    phid = dev(phi)
    R.or_sweep(phi, rhs, G, n, 1.0, 3)
    ctx.time_sweeps(phid, dev(rhs), dev(G), 1.0, 3)
    assert np.array_equal(host(phid), phi)


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("n,nrelax,iopt", [(16, 4, 19), (16, 3, 19), (32, 8, 19), (8, 32, 19), (4, 32, 18),
                                           (16, 4, 11), (16, 4, 18)])
def test_relax_bitexact(ctx, dt, n, nrelax, iopt):
    print("case relax n=%d nrelax=%d iopt=%d %s" % (n, nrelax, iopt, dt.__name__), flush=True)
    rho = field(n, 1.0, 5, dt)
    u = aos(rho)
    d = smooth_def(n, 1.5, 8, dt)
    rhs = np.random.default_rng(2).standard_normal(n ** 3).astype(dt)
    a = np.zeros(n ** 3, dt)
    r = np.zeros((n // 2) ** 3, dt)
    rinf = R.or_relax(a, r, rhs, d, u, 5, 1.0, n, nrelax, iopt)
    ad, rd = dev(np.zeros(n ** 3, dt)), dev(np.zeros((n // 2) ** 3, dt))
    rinf_d = ctx.relax(ad, rd, dev(rhs), dev(d), dev(u), 1.0, nrelax, iopt, nu=5)
    assert np.array_equal(host(ad), a)
    if iopt % 8 in (1, 3):
        assert np.array_equal(host(rd), r)
    assert rinf_d == rinf


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("n,ng,sigma", [(16, 16, 1.0), (32, 32, 2.0), (32, 64, 1.0), (64, 64, 2.0), (128, 128, 2.0)])
def test_multigrid_bitexact(ctx, dt, n, ng, sigma):
    print("case multigrid n=%d ng=%d %s" % (n, ng, dt.__name__), flush=True)
    rho = field(n, sigma, 7, dt)
    u = aos(rho)
    d = smooth_def(n, 2.0, 4, dt)
    rhs = (1 - rho).astype(dt)
    R.or_nsmooth(rhs, n)
    a, b = np.zeros(n ** 3, dt), rhs.copy()
    st = R.or_multigrid(a, b, d, u, 5, 1.0, n, 2, ng, 2)
    ad, bd = dev(np.zeros(n ** 3, dt)), dev(rhs)
    sd = ctx.multigrid(ad, bd, dev(d), dev(u), 1.0, 2, ng, nu=5)
    assert np.array_equal(host(bd), b)
    assert np.array_equal(host(ad), a)
    assert list(sd.guard[:sd.nlev]) == list(st.guard[:sd.nlev])
    assert [sd.relaxerr[l] for l in range(sd.nlev)] == [st.relaxerr[l] for l in range(sd.nlev)]


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("n,sigma,amp", [(16, 1.0, 0.0), (32, 2.0, 2.0), (32, 2.0, 10.0)])
def test_calcdefp(ctx, dt, n, sigma, amp):
    print("case calcdefp n=%d amp=%g %s" % (n, amp, dt.__name__), flush=True)
    rho = field(n, sigma, 7, dt)
    u = aos(rho)
    d = smooth_def(n, amp, 4, dt) if amp else np.zeros(n ** 3, dt)
    p, t, t2 = np.zeros(n ** 3, dt), np.zeros(n ** 3, dt), np.zeros((n + 2) * n * n, dt)
    ls = R.or_calcdefp(p, t, t2, d, u, n, sumorder=2)
    pd, td, t2d = dev(np.zeros(n ** 3, dt)), dev(np.zeros(n ** 3, dt)), dev(np.zeros((n + 2) * n * n, dt))
    lg = ctx.calcdefp(pd, td, t2d, dev(d), dev(u))
    assert lg.ncompress == ls.ncompress and lg.nexpand == ls.nexpand
    got = host(pd)
    if ls.ncompress == 0:
        assert np.array_equal(got, p)
    else:
        tol = 1e-12 if dt == np.float64 else 1e-5
        assert np.abs(got - p).max() <= tol * np.abs(p).max()
    assert np.array_equal(host(td)[3:], t[3:])


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("n", [16, 32])
def test_relaxing_bitexact(ctx, dt, n):
    print("case relaxing n=%d %s" % (n, dt.__name__), flush=True)
    rho = field(n, 1.5, 9, dt)
    u = aos(rho)
    d = smooth_def(n, 2.0, 5, dt)
    dp = smooth_def(n, 0.5, 6, dt)
    ua = u.copy()
    R.or_relaxing(ua, d, dp, n)
    ud = dev(u)
    ctx.relaxing(ud, dev(d), dev(dp))
    assert np.array_equal(host(ud), ua)


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("n", [16, 32])
def test_gradient_delta_bitexact(ctx, dt, n):
    print("case gradient/delta n=%d %s" % (n, dt.__name__), flush=True)
    phi = smooth_def(n, 3.0, 12, dt)
    g = R.or_gradient(phi, n)
    dl = R.or_delta(phi, n)
    gd, dd = dev(np.zeros(3 * n ** 3, dt)), dev(np.zeros(n ** 3, dt))
    ctx.scalar2gradient(dev(phi), gd)
    ctx.defp2delta(dev(phi), dd)
    assert np.array_equal(host(gd), g)
    assert np.array_equal(host(dd), dl)


@pytest.mark.parametrize("n", [16, 32])
def test_small_ops_bitexact(ctx, n):
    dt = np.float64
    a = smooth_def(n, 1.0, 1, dt)
    b = a.copy()
    R.or_nsmooth(b, n)
    ad = dev(a)
    ctx.nsmooth(ad)
    assert np.array_equal(host(ad), b)
    h = R.or_restrict(a, n)
    hd = dev(np.zeros((n // 2) ** 3, dt))
    ctx.restrict(hd, dev(a))
    assert np.array_equal(host(hd), h)
    u = a.copy()
    R.or_inject(u, h, n)
    ud = dev(a)
    ctx.inject(ud, dev(h))
    assert np.array_equal(host(ud), u)
    data = a.copy()
    speq = np.zeros(2 * n * n, dt)
    R.or_rlft3(data, speq, n, 1)
    dd, sd = dev(a), dev(np.zeros(2 * n * n, dt))
    ctx.rlft3(dd, sd, 1)
    assert np.array_equal(host(dd), data) and np.array_equal(host(sd), speq)


@pytest.mark.parametrize("n,sigma,seed", [(16, 1.0, 1234), (32, 2.0, 7)])
def test_lognormal_generator(ctx, n, sigma, seed):
    ref = R.or_lognormal(n, sigma, seed)
    rd = dev(np.zeros(n ** 3))
    ctx.lognormal(rd, sigma, seed)
    got = host(rd)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    assert abs(got.mean() - 1.0) < 1e-12


def test_error_paths(ctx):
    import mmh_amd
    x = dev(np.zeros(6 ** 3))
    with pytest.raises(mmh_amd.MmhError) as e:
        ctx.multigrid(x, x.clone(), x.clone(), x.clone())
    assert e.value.code == mmh_amd.MMH_EINVAL


@pytest.mark.parametrize("n,sched", [(32, 0), (32, 1), (64, 0)])
def test_solve_fp64_bitexact(ctx, n, sched):
    print("case solve n=%d sched=%d" % (n, sched), flush=True)
    rho = field(n, 2.0, 42, np.float64)
    u = aos(rho)
    d = smooth_def(n, 1.0, 4, np.float64)
    rhs = np.zeros(n ** 3)
    rd = dev(rhs)
    ctx.limiter_rhs(rd, dev(d), dev(u))
    rhs = host(rd)
    a, b = np.zeros(n ** 3), rhs.copy()
    rc, rr = R.or_solve(a, b, d, u, 5, n, 1e-8, 100, sched, 2, 2, 16, 0, 2)
    assert rc == 0
    ad, bd = dev(np.zeros(n ** 3)), dev(rhs)
    st, ok = ctx.solve(ad, bd, dev(d), dev(u), 1e-8, 100, sched, 2, 2, 16, nu=5)
    assert ok and st.cycles == len(rr)
    assert list(st.relres[:st.cycles]) == list(rr)
    assert np.array_equal(host(ad), a)


@pytest.mark.parametrize("n,sched", [(32, 0), (64, 0)])
def test_solve_mixed_matches_fp64(ctx, n, sched):
    print("case mixed n=%d sched=%d" % (n, sched), flush=True)
    rho = field(n, 2.0, 42, np.float64)
    u = aos(rho)
    d = smooth_def(n, 1.0, 4, np.float64)
    rd = dev(np.zeros(n ** 3))
    ctx.limiter_rhs(rd, dev(d), dev(u))
    rhs = host(rd)
    a64, b64 = dev(np.zeros(n ** 3)), dev(rhs)
    st64, ok64 = ctx.solve(a64, b64, dev(d), dev(u), 1e-10, 200, sched, 2, 2, 16, nu=5)
    am, bm = dev(np.zeros(n ** 3)), dev(rhs)
    stm, okm = ctx.solve(am, bm, dev(d), dev(u), 1e-8, 200, sched, 2, 2, 16, mixed=True, nu=5)
    assert ok64 and okm
    assert stm.relres[stm.cycles - 1] <= 1e-8
    x64, xm = host(a64), host(am)
    err = np.abs(xm - x64).max() / np.abs(x64).max()
    print("mixed cycles %d fp64 cycles %d rel err %.3e" % (stm.cycles, st64.cycles, err))
    assert err <= 1e-7


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("n", [16, 64, 128])
def test_sublattice_sweep_equals_column_major(ctx, dt, n):
    print("case sub-lattice sweep n=%d %s" % (n, dt.__name__), flush=True)
    rho = field(n, 2.0, 3, dt)
    u = aos(rho)
    d = smooth_def(n, 2.0, 4, dt)
    Gd = dev(np.zeros(6 * n ** 3, dt))
    ctx.metric(Gd, dev(d), dev(u), 1.0, 0, nu=5)
    rhs = dev((1 - rho).astype(dt))
    out = []
    for impl in (0, 1):
        ctx.set_option("sweep_impl", impl)
        phi = dev(np.zeros(n ** 3, dt))
        ctx.time_sweeps(phi, rhs, Gd, 1.0, 3)
        out.append(host(phi))
    ctx.set_option("sweep_impl", 0)
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("n,ns", [(128, 2), (128, 3), (256, 2)])
def test_wavefront_sweep_equals_colour_kernels(ctx, dt, n, ns):
    """The whole-sweep wavefront kernel (sweep_impl 0 at n >= 128) against the colour-per-launch
    kernels (sweep_impl 2) on the same sub-lattice arrays: bit-identical after ns sweeps (an odd
    count ends with one colour-by-colour sweep)."""
    print("case wavefront sweep n=%d ns=%d %s" % (n, ns, dt.__name__), flush=True)
    rho = field(n, 2.0, 3, dt)
    u = aos(rho)
    d = smooth_def(n, 2.0, 4, dt)
    Gd = dev(np.zeros(6 * n ** 3, dt))
    ctx.metric(Gd, dev(d), dev(u), 1.0, 0, nu=5)
    rhs = dev((1 - rho).astype(dt))
    phi0 = smooth_def(n, 0.5, 9, dt)
    out = []
    ctx.set_option("wave", 1)
    try:
        for impl in (0, 2):
            ctx.set_option("sweep_impl", impl)
            phi = dev(phi0)
            ctx.time_sweeps(phi, rhs, Gd, 1.0, ns)
            out.append(host(phi))
    finally:
        ctx.set_option("sweep_impl", 0)
        ctx.set_option("wave", 0)
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("S", [2, 4, 8, 16])
def test_column_colour_kernel_equals_colour_kernel(ctx, dt, S):
    """The colour kernel with S points per thread along k (option colour_col; here on every level
    with n/2 divisible by S, option colour_col_min_n = 8) against one point per thread:
    bit-identical after 3 sweeps from a non-zero start at 128^3 and 256^3, through a reference
    V-cycle, and (FP64 data) through PCG solves with FP64 and FP32 preconditioners, the latter in
    plain and difference form; the solves' first sweeps from zero take the masked form."""
    print("case colour_col S=%d %s" % (S, dt.__name__), flush=True)
    ctx.set_option("colour_col_min_n", 8)
    try:
        _column_colour_cases(ctx, dt, S)
    finally:
        ctx.set_option("colour_col_min_n", 256)
        ctx.set_option("colour_col", 8)


def _column_colour_cases(ctx, dt, S):
    res = {1: [], S: []}
    for n in (128, 256):
        rho = field(n, 2.0, 3, dt)
        u = aos(rho)
        d = smooth_def(n, 2.0, 4, dt)
        Gd = dev(np.zeros(6 * n ** 3, dt))
        ctx.metric(Gd, dev(d), dev(u), 1.0, 0, nu=5)
        rhs = dev((1 - rho).astype(dt))
        phi0 = smooth_def(n, 0.5, 9, dt)
        for cc in (1, S):
            ctx.set_option("colour_col", cc)
            try:
                phi = dev(phi0)
                ctx.time_sweeps(phi, rhs, Gd, 1.0, 3)
                res[cc].append(host(phi))
            finally:
                ctx.set_option("colour_col", 1)
    n = 128
    rho = field(n, 2.0, 7, dt)
    u = aos(rho)
    d = smooth_def(n, 2.0, 4, dt)
    rhs = (1 - rho).astype(dt)
    R.or_nsmooth(rhs, n)
    for cc in (1, S):
        ctx.set_option("colour_col", cc)
        try:
            ad, bd = dev(np.zeros(n ** 3, dt)), dev(rhs)
            ctx.multigrid(ad, bd, dev(d), dev(u), 1.0, 2, n, nu=5)
            res[cc].append(host(ad))
            if dt == np.float64:
                for mixed, delta in ((0, 0), (1, 0), (1, 1)):
                    ctx.set_option("mixed_delta", delta)
                    phi, b = dev(np.zeros(n ** 3)), dev(rhs)
                    st, ok = ctx.solve(phi, b, dev(d), dev(u), tol=1e-8, max_cycles=100, schedule=1, nu_pre=2,
                                       nu_post=2, nu_coarse=8, mixed=mixed, nu=5, krylov=True)
                    assert ok
                    res[cc].append(host(phi))
                    res[cc].append(np.array([st.relres[i] for i in range(st.cycles)]))
        finally:
            ctx.set_option("colour_col", 1)
            ctx.set_option("mixed_delta", 0)
    assert len(res[1]) == len(res[S])
    for a, b in zip(res[1], res[S]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("dt", DT)
def test_wavefront_vcycle_bitexact(ctx, dt):
    """With the wavefront kernel switched on, a 128^3 V-cycle (forward sweeps at 128^3) and a
    symmetric-preconditioner PCG solve (reverse sweeps) give the colour kernels' results."""
    n = 128
    print("case wavefront V-cycle/PCG n=%d %s" % (n, dt.__name__), flush=True)
    rho = field(n, 2.0, 7, dt)
    u = aos(rho)
    d = smooth_def(n, 2.0, 4, dt)
    rhs = (1 - rho).astype(dt)
    R.or_nsmooth(rhs, n)
    res = []
    for wave in (0, 1):
        ctx.set_option("wave", wave)
        try:
            ad, bd = dev(np.zeros(n ** 3, dt)), dev(rhs)
            ctx.multigrid(ad, bd, dev(d), dev(u), 1.0, 2, n, nu=5)
            res.append(host(ad))
            if dt == np.float64:
                phi, b = dev(np.zeros(n ** 3)), dev(rhs)
                ctx.solve(phi, b, dev(d), dev(u), tol=1e-8, max_cycles=100, schedule=1, nu_pre=2, nu_post=2,
                          nu_coarse=8, nu=5, krylov=True)
                res.append(host(phi))
        finally:
            ctx.set_option("wave", 0)
    half = len(res) // 2
    for a, b in zip(res[:half], res[half:]):
        assert np.array_equal(a, b)


def test_mixed_preconditioner_1024(ctx):
    """At 1024^3 the FP32 copy of the metric is 6 * 2^30 values, more than one dispatch can cover;
    three PCG iterations with the FP32 preconditioner must track the FP64 ones: iterates to 1e-6
    relative, residual histories to 1e-3 (FP32 rounding of the V-cycle at |phi| ~ 255 moves the
    first residuals by about 1e-4; a truncated metric copy diverges instead).  Both runs use the
    rebuilt coarse operators (option galerkin = 0): at 1024^3 the Galerkin ones fit beside the
    FP64 run's workspace for one precision only, and the comparison needs the same preconditioner.
    The mixed run is repeated with the 16-bit coefficient copies (option precond_bf16), whose
    conversion is split the same way, with the looser bounds of their 8-bit significands.  These
    three runs sweep the finest level colour by colour (option sweep_slabs = 0), so only the
    precision differs; a fourth, 16-bit with the default slab order (another symmetric
    Gauss-Seidel order), must stay within 1e-2 of the third."""
    n = 1024
    print("case mixed preconditioner n=%d" % n, flush=True)
    f64 = dict(dtype=torch.float64, device="cuda:0")
    rho = torch.zeros(n ** 3, **f64)
    ctx.lognormal(rho, 2.0, 42)
    d = torch.zeros(n ** 3, **f64)
    rhs0 = (1.0 - rho) / 30.0
    out = []
    ctx.set_option("galerkin", 0)
    try:
        for mixed, bf, slabs in ((0, 0, 0), (1, 0, 0), (1, 1, 0), (1, 1, 4)):
            ctx.set_option("precond_bf16", bf)
            ctx.set_option("sweep_slabs", slabs)
            phi, rhs = torch.zeros(n ** 3, **f64), rhs0.clone()
            st, ok = ctx.solve(phi, rhs, d, rho, tol=1e-30, max_cycles=3, schedule=1, nu_pre=2, nu_post=2,
                               nu_coarse=8, mixed=mixed, nu=1, krylov=True)
            out.append((phi, [st.relres[i] for i in range(st.cycles)]))
            del rhs
            ctx.release_workspace()
    finally:
        ctx.set_option("galerkin", 1)
        ctx.set_option("precond_bf16", 1)
        ctx.set_option("sweep_slabs", 4)
    (a, ra), (b, rb), (c, rc), (e, re) = out
    print("relres fp64 %s mixed %s mixed-bf16 %s slab order %s" % (ra, rb, rc, re), flush=True)
    assert float((a - b).abs().max() / a.abs().max()) <= 1e-6
    assert all(abs(x - y) <= 1e-3 * abs(x) for x, y in zip(ra, rb))
    # 16-bit coefficients (8-bit significands) perturb the preconditioner by ~1e-3 relative
    assert float((a - c).abs().max() / a.abs().max()) <= 1e-4
    assert all(abs(x - y) <= 1e-2 * abs(x) for x, y in zip(ra, rc))
    assert float((c - e).abs().max() / c.abs().max()) <= 1e-2
    assert all(abs(x - y) <= 1e-2 * abs(x) for x, y in zip(rc, re))


@pytest.mark.parametrize("mixed", [0, 1])
def test_marching_operator_pcg(ctx, mixed):
    """The k-marching residual/operator kernels (levels >= 64^3) against the cell-lockstep ones in a
    128^3 PCG solve: r and q are the same values, only the order of the p.q block sums differs, so
    the iteration counts agree and the solutions agree to 1e-12 relative."""
    n = 128
    print("case marching operator PCG n=%d mixed=%d" % (n, mixed), flush=True)
    rho = field(n, 2.0, 7, np.float64)
    u = aos(rho)
    d = smooth_def(n, 2.0, 4, np.float64)
    rhs = (1 - rho) / 30.0
    out = []
    for march in (1, 0):
        ctx.set_option("march_op", march)
        try:
            phi, b = dev(np.zeros(n ** 3)), dev(rhs)
            st, ok = ctx.solve(phi, b, dev(d), dev(u), tol=1e-10, max_cycles=100, schedule=1, nu_pre=2, nu_post=2,
                               nu_coarse=8, mixed=mixed, nu=5, krylov=True)
        finally:
            ctx.set_option("march_op", 1)
        assert ok
        out.append((host(phi), st.cycles))
    (a, ca), (b, cb) = out
    assert ca == cb
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


@pytest.mark.parametrize("mixed,galerkin", [(0, 1), (1, 1), (0, 0)])
@pytest.mark.parametrize("n", [32, 64])
def test_krylov_solution_solves_the_reference_equation(ctx, n, mixed, galerkin):
    """The PCG solve has no reference counterpart: what pins it is the equation it solves.  Its
    result must satisfy the reference operator (the oracle's metric and calcresid residual) to the
    requested tolerance (the recursive residual reached 1e-10; the true one may differ by rounding),
    after the same rgzerosum projection of the right-hand side, and it must agree with the oracle's
    own solution of the same system (reference V-cycles to 1e-8) as far as that residual fixes it."""
    print("case Krylov solution vs oracle operator n=%d mixed=%d galerkin=%d" % (n, mixed, galerkin), flush=True)
    rho = field(n, 2.0, 42, np.float64)
    u = aos(rho)
    d = smooth_def(n, 1.0, 4, np.float64)
    rd = dev(np.zeros(n ** 3))
    ctx.limiter_rhs(rd, dev(d), dev(u))
    rhs = host(rd)
    ctx.set_option("galerkin", galerkin)
    try:
        ad, bd = dev(np.zeros(n ** 3)), dev(rhs)
        st, ok = ctx.solve(ad, bd, dev(d), dev(u), tol=1e-10, max_cycles=200, schedule=1, nu_pre=2, nu_post=2,
                           nu_coarse=8, mixed=mixed, nu=5, krylov=True)
    finally:
        ctx.set_option("galerkin", 1)
    assert ok
    phi, bproj = host(ad), host(bd)
    G = R.or_metric(d, u, 5, n, 1.0)
    _, rinf = R.or_residual(phi, bproj, G, n, 1.0)
    rel = rinf / np.abs(bproj).max()
    a, b = np.zeros(n ** 3), rhs.copy()
    rc, rr = R.or_solve(a, b, d, u, 5, n, 1e-8, 100, 0, 2, 2, 16, 0, 2)
    err = np.abs((phi - phi.mean()) - (a - a.mean())).max() / np.abs(a - a.mean()).max()
    print("true relres %.3e (recursive %.3e, %d iterations); oracle solve %d cycles, difference %.3e"
          % (rel, st.relres[st.cycles - 1], st.cycles, len(rr), err), flush=True)
    assert rel <= 1e-9
    assert rc == 0
    assert np.array_equal(b, bproj)
    assert err <= 1e-5


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("n,ng", [(32, 32), (32, 64), (64, 64)])
def test_reference_vcycle_fused_coarse_levels_equal_launches(ctx, dt, n, ng):
    """The reference V-cycle's coarse levels in one workgroup (option coarse_fused_ref: levels up
    to 32, 16, 8 or 4; with the smallest of them held in LDS, option coarse_lds, and without)
    against the launch-by-launch cycle: the same iterate, per-level residual norms r1 and relaxerr,
    and divergence-guard decisions (the finest level is guarded too when NG > N)."""
    print("case fused reference V-cycle n=%d ng=%d %s" % (n, ng, dt.__name__), flush=True)
    rho = field(n, 2.0, 7, dt)
    u = aos(rho)
    d = smooth_def(n, 2.0, 4, dt)
    rhs = (1 - rho).astype(dt)
    R.or_nsmooth(rhs, n)
    out = []
    for cf, lds in ((32, 1), (32, 0), (16, 1), (16, 0), (8, 1), (4, 1), (0, 1)):
        ctx.set_option("coarse_fused_ref", cf)
        ctx.set_option("coarse_lds", lds)
        try:
            ad, bd = dev(np.zeros(n ** 3, dt)), dev(rhs)
            st = ctx.multigrid(ad, bd, dev(d), dev(u), 1.0, 2, ng, nu=5)
        finally:
            ctx.set_option("coarse_fused_ref", 16)
            ctx.set_option("coarse_lds", 1)
        out.append((host(ad), [(st.r1[l], st.relaxerr[l], st.guard[l]) for l in range(st.nlev)]))
    for a, s in out[:-1]:
        assert s == out[-1][1]
        assert np.array_equal(a, out[-1][0])


GUARD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mg_guard_32.npz")


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("ng", [32, 64])
@pytest.mark.parametrize("cf", [16, 0])
def test_multigrid_divergence_guard(ctx, dt, ng, cf):
    """V-cycles whose divergence guard fires: the second outer iteration of a sigma = 3 field at
    32^3, where the reference reports 'relaxation diverged' on the 16^3, 8^3 and 4^3 levels
    (mmh/mgtemplate.f90:46-55; 4^3 through mg4's 1.001 factor and r1 > 1e-5 test,
    mmh/multigrid.f90:113-119).  With NG = 64 the 32^3 grid is a coarse level and is guarded too.
    Both forms of the coarse levels (one workgroup up to 16^3, and launch by launch) must give the
    oracle's iterate, norms and guard decisions bit for bit; with NG = 32 the decisions must also be
    the levels of the committed reference run."""
    print("case guard ng=%d coarse_fused_ref=%d %s" % (ng, cf, dt.__name__), flush=True)
    z = np.load(GUARD, allow_pickle=False)
    n = int(z["n"])
    var = "f64" if dt == np.float64 else "f32"
    if dt == np.float32 and ng == 64:
        pytest.skip("the FP32 cycle with a guarded 32^3 level overflows to inf in the oracle and the reference")
    u = aos(z["rho"].astype(dt))
    d = z["def"].astype(dt)
    a, b = np.zeros(n ** 3, dt), z["rhs_in"].astype(dt)
    ad, bd, dd, ud = dev(a), dev(b), dev(d), dev(u)
    ctx.set_option("coarse_fused_ref", cf)
    try:
        for c in range(int(z["cycles_" + var])):
            st = R.or_multigrid(a, b, d, u, 5, 1.0, n, 2, ng, 2)
            sd = ctx.multigrid(ad, bd, dd, ud, 1.0, 2, ng, nu=5)
            assert sd.nlev == st.nlev
            assert [sd.guard[l] for l in range(sd.nlev)] == [st.guard[l] for l in range(st.nlev)]
            assert [sd.r1[l] for l in range(sd.nlev)] == [st.r1[l] for l in range(st.nlev)]
            assert [sd.relaxerr[l] for l in range(sd.nlev)] == [st.relaxerr[l] for l in range(st.nlev)]
            assert sd.entry_zerosum == st.zerosum_entry
            fired = sorted(n >> l for l in range(st.nlev) if st.guard[l])
            if ng == n:
                assert fired == list(z["diverged_levels_" + var][c])
            else:
                assert n in fired
            assert np.array_equal(host(ad), a)
            assert np.array_equal(host(bd), b)
    finally:
        ctx.set_option("coarse_fused_ref", 16)
    if ng == n:
        # the reference sums sequentially, the device in blocked trees: FP32 agrees bit for bit
        # here, FP64 to the rounding of the mean removals
        ref = z["phi_out_" + var]
        if dt == np.float32:
            assert np.array_equal(a, ref)
        else:
            assert np.abs(a - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("n", [64, 128])
def test_mixed_bf16_coefficients(ctx, n):
    """The mixed Krylov solve with the preconditioner's coefficients in 16 bits (option precond_bf16)
    reaches the same FP64 tolerance as with FP32 copies, agrees with the FP64 solve within 1e-7 and
    needs at most two more iterations; its true residual under the oracle's operator is within the
    tolerance."""
    print("case mixed bf16 coefficients n=%d" % n, flush=True)
    rho = field(n, 2.0, 42, np.float64)
    u = aos(rho)
    d = smooth_def(n, 1.0, 4, np.float64)
    rd = dev(np.zeros(n ** 3))
    ctx.limiter_rhs(rd, dev(d), dev(u))
    rhs = host(rd)
    res = {}
    try:
        for mixed, bf in ((0, 1), (1, 0), (1, 1)):
            ctx.set_option("precond_bf16", bf)
            ad, bd = dev(np.zeros(n ** 3)), dev(rhs)
            st, ok = ctx.solve(ad, bd, dev(d), dev(u), tol=1e-9, max_cycles=200, schedule=1, nu_pre=2, nu_post=2,
                               nu_coarse=4, mixed=mixed, nu=5, krylov=True)
            assert ok
            res[(mixed, bf)] = (host(ad), host(bd), st.cycles)
    finally:
        ctx.set_option("precond_bf16", 1)
    x64, b64, c64 = res[(0, 1)]
    x32, _, c32 = res[(1, 0)]
    x16, b16, c16 = res[(1, 1)]
    print("iterations fp64 %d mixed-fp32 %d mixed-bf16 %d" % (c64, c32, c16), flush=True)
    assert c16 <= c32 + 2
    assert np.abs(x16 - x64).max() <= 1e-7 * np.abs(x64).max()
    G = R.or_metric(d, u, 5, n, 1.0)
    _, rinf = R.or_residual(x16, b16, G, n, 1.0)
    assert rinf <= 1.05e-9 * np.abs(b16).max()


@pytest.mark.parametrize("n", [64, 256])
def test_bf16_coefficient_sweeps_equal_fp32_sweeps(ctx, n):
    """The FP32 Gauss-Seidel sweeps that read 16-bit coefficients (mmh_time_sweeps_fb, the finest-level
    smoother of the mixed solve; the column kernel at 256^3) give bit for bit the FP32 sweeps
    (mmh_time_sweeps_f) run on the same metric rounded to bfloat16 by torch and widened back: the
    16-bit path changes where the coefficients are read from, not the arithmetic."""
    print("case bf16 sweeps n=%d" % n, flush=True)
    rho = field(n, 2.0, 3, np.float64)
    u = aos(rho)
    d = smooth_def(n, 2.0, 4, np.float64)
    Gd = dev(np.zeros(6 * n ** 3))
    ctx.metric(Gd, dev(d), dev(u), 1.0, 0, nu=5)
    rhs = dev((1 - rho).astype(np.float32))
    pa = dev(np.zeros(n ** 3, np.float32))
    ctx.time_sweeps_bf16(pa, rhs, Gd, 1.0, 3)
    Gr = Gd.float().to(torch.bfloat16).float().contiguous()
    pb = dev(np.zeros(n ** 3, np.float32))
    ctx.time_sweeps(pb, rhs, Gr, 1.0, 3)
    a, b = host(pa), host(pb)
    assert np.abs(a).max() > 0
    assert np.array_equal(a, b)
