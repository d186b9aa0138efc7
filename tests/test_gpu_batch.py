"""GPU: the batched solve (many realisations of uneven sizes in one call) against single solves.
Every realisation must be bit-identical to its own mmh_solve_d (FP64 PCG), with graph replay and
with eager launches alike (the graph is replayed tens of times per solve and checked against the
eager result), and with or without launch-shape tuning."""
import numpy as np
import pytest

import refbind as R

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

F64 = dict(dtype=torch.float64, device="cuda:0")
SOLVE = dict(tol=1e-8, max_cycles=200, nu_pre=2, nu_coarse=8)


def problem(ctx, n, sigma, seed):
    rho = torch.zeros(n ** 3, **F64)
    ctx.lognormal(rho, sigma, seed)
    u = torch.zeros(n ** 3, 5, **F64)
    u[:, 0] = rho
    u = u.contiguous().view(-1)
    d = torch.zeros(n ** 3, **F64)
    rhs = torch.zeros(n ** 3, **F64)
    ctx.limiter_rhs(rhs, d, u)
    return u, rhs, d


CASES = [(32, 0.5, 0), (32, 2.0, 1), (64, 1.2, 2), (32, 1.0, 3), (64, 1.8, 4), (128, 1.0, 5)]


def run_batch(ctx, probs):
    items = [(torch.zeros_like(r), r.clone(), d, u) for (u, r, d) in probs]
    st, ok = ctx.solve_batch(items, **SOLVE)
    torch.cuda.synchronize()
    return [it[0].cpu().numpy() for it in items], st, ok


@pytest.fixture(scope="module")
def probs(ctx):
    return [problem(ctx, *c) for c in CASES]


@pytest.mark.parametrize("galerkin,nu_top", [(1, 0), (1, 1), (0, 0)])
def test_batch_bit_identical_to_single_solves(ctx, probs, galerkin, nu_top):
    """With Galerkin or rebuilt coarse operators, and with one sweep on the finest level or the
    common count, every realisation of the batch equals its single solve."""
    print("case batch vs single: galerkin %d nu_top %d %s" % (galerkin, nu_top, CASES), flush=True)
    ctx.set_option("galerkin", galerkin)
    ctx.set_option("nu_top", nu_top)
    try:
        xs, st, ok = run_batch(ctx, probs)
        singles = []
        for (u, r, d) in probs:
            phi = torch.zeros_like(r)
            st1, ok1 = ctx.solve(phi, r.clone(), d, u, tol=1e-8, max_cycles=200, schedule=1, nu_pre=2, nu_post=2,
                                 nu_coarse=8, nu=5, krylov=True)
            singles.append((phi, st1, ok1))
    finally:
        ctx.set_option("galerkin", 1)
        ctx.set_option("nu_top", 0)
    assert ok
    for (phi, st1, ok1), x, s, case in zip(singles, xs, st, CASES):
        assert ok1 and s.converged
        assert s.cycles == st1.cycles, (case, s.cycles, st1.cycles)
        assert s.relres == st1.relres[st1.cycles - 1]
        assert np.array_equal(x, phi.cpu().numpy()), case


def test_batch_graph_replay_equals_eager(ctx, probs):
    print("case batch graph vs eager", flush=True)
    xg, stg, okg = run_batch(ctx, probs)
    ctx.set_option("batch_graph", 0)
    try:
        xe, ste, oke = run_batch(ctx, probs)
    finally:
        ctx.set_option("batch_graph", 1)
    assert okg and oke
    for a, b, s1, s2 in zip(xg, xe, stg, ste):
        assert np.array_equal(a, b)
        assert s1.cycles == s2.cycles


def test_batch_tuning_does_not_change_results(ctx, probs):
    print("case batch autotune off", flush=True)
    xa, _, _ = run_batch(ctx, probs)
    ctx.set_option("batch_autotune", 0)
    try:
        xb, _, _ = run_batch(ctx, probs)
    finally:
        ctx.set_option("batch_autotune", 1)
    for a, b in zip(xa, xb):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("pers", [0, 64])
def test_batch_zero_rhs_item(ctx, probs, pers):
    """A grid with a zero right-hand side beside one that iterates: zero result, no iteration, with
    the batched launches and with one workgroup per grid (batch_persistent)."""
    print("case batch zero rhs, batch_persistent %d" % pers, flush=True)
    u, r, d = probs[0]
    items = [(torch.ones_like(r), torch.zeros_like(r), d, u), (torch.zeros_like(r), r.clone(), d, u)]
    ctx.set_option("batch_persistent", pers)
    try:
        st, ok = ctx.solve_batch(items, **SOLVE)
    finally:
        ctx.set_option("batch_persistent", 0)
    assert ok and st[0].cycles == 0 and st[0].converged
    assert float(items[0][0].abs().max()) == 0.0
    assert st[1].converged and st[1].cycles > 0 and st[1].relres <= SOLVE["tol"]


@pytest.mark.parametrize("pers", [0, 64])
def test_batch_group_converged_from_the_start(ctx, probs, pers):
    """A size group whose grids all have a zero right-hand side (converged before any iteration)
    beside a group that iterates: the first gives zeros with no iteration, the second its single
    solve's result (bit for bit with the batched launches; to rounding, within 2 iterations, with
    one workgroup per grid)."""
    print("case batch all-converged group, batch_persistent %d" % pers, flush=True)
    u0, r0, d0 = probs[2]  # 64^3
    u1, r1, d1 = probs[0]  # 32^3
    items = [(torch.ones_like(r0), torch.zeros_like(r0), d0, u0),
             (torch.ones_like(r0), torch.zeros_like(r0), d0, u0),
             (torch.zeros_like(r1), r1.clone(), d1, u1)]
    ctx.set_option("batch_persistent", pers)
    try:
        st, ok = ctx.solve_batch(items, **SOLVE)
    finally:
        ctx.set_option("batch_persistent", 0)
    torch.cuda.synchronize()
    assert ok
    for k in (0, 1):
        assert st[k].cycles == 0 and st[k].converged
        assert float(items[k][0].abs().max()) == 0.0
    phi = torch.zeros_like(r1)
    st1, ok1 = ctx.solve(phi, r1.clone(), d1, u1, tol=1e-8, max_cycles=200, schedule=1, nu_pre=2, nu_post=2,
                         nu_coarse=8, nu=5, krylov=True)
    x, ref = items[2][0].cpu().numpy(), phi.cpu().numpy()
    if pers == 0:
        assert ok1 and st[2].cycles == st1.cycles
        assert np.array_equal(x, ref)
    else:
        assert ok1 and abs(st[2].cycles - st1.cycles) <= 2
        assert np.abs(x - ref).max() <= 1e-6 * np.abs(ref).max()


@pytest.mark.parametrize("mixed", [0, 1])
def test_single_solve_graph_replay_equals_eager(ctx, probs, mixed):
    """The single solve replays its PCG preconditioner (one V-cycle, FP64 or FP32) as a captured
    graph from the second call on, tens of times per solve with eager CG work in between; every
    iterate must equal the all-eager solve's, so the residual histories and solutions match."""
    print("case single-solve graph vs eager, mixed=%d" % mixed, flush=True)
    for (u, r, d), case in zip(probs, CASES):
        out = []
        for graph in (1, 0):
            ctx.set_option("pcg_graph", graph)
            try:
                phi = torch.zeros_like(r)
                st, ok = ctx.solve(phi, r.clone(), d, u, tol=1e-8, max_cycles=200, schedule=1, nu_pre=2, nu_post=2,
                                   nu_coarse=8, mixed=mixed, nu=5, krylov=True)
            finally:
                ctx.set_option("pcg_graph", 1)
            assert ok, case
            out.append((phi.cpu().numpy(), [st.relres[i] for i in range(st.cycles)]))
        (xg, rg), (xe, re) = out
        assert len(rg) > 3, case
        assert rg == re, case
        assert np.array_equal(xg, xe), case


@pytest.mark.parametrize("mixed", [0, 1])
def test_single_solve_coarse_cycle_equals_launches(ctx, probs, mixed):
    """The preconditioner's coarse levels run in one workgroup (k_coarse_cycle, levels up to 32^3 by
    option, with and without the smallest of them held in LDS); every iterate must equal the
    launch-by-launch cycle's (the same per-point functions, the same sum order)."""
    print("case single-solve fused coarse cycle vs launches, mixed=%d" % mixed, flush=True)
    for (u, r, d), case in zip(probs, CASES):
        out = []
        for fused, lds in ((32, 1), (16, 1), (16, 0), (8, 1), (4, 1), (0, 1)):
            ctx.set_option("coarse_fused", fused)
            ctx.set_option("coarse_lds", lds)
            try:
                phi = torch.zeros_like(r)
                st, ok = ctx.solve(phi, r.clone(), d, u, tol=1e-8, max_cycles=200, schedule=1, nu_pre=2, nu_post=2,
                                   nu_coarse=8, mixed=mixed, nu=5, krylov=True)
            finally:
                ctx.set_option("coarse_fused", 16)
                ctx.set_option("coarse_lds", 1)
            assert ok, case
            out.append((phi.cpu().numpy(), [st.relres[i] for i in range(st.cycles)]))
        xl, rl = out[-1]
        for xf, rf in out[:-1]:
            assert rf == rl, case
            assert np.array_equal(xf, xl), case


@pytest.mark.parametrize("mixed,galerkin", [(0, 1), (1, 1), (0, 0), (1, 0)])
def test_zero_start_first_sweep_equals_full_sweep(ctx, probs, mixed, galerkin):
    """A level's first sweep from u = 0 loads only the colour classes updated before it in that
    sweep (option zero_start): the iterates and solutions equal those of the full first sweep, for
    the 19-point and Galerkin levels, FP64 and FP32 preconditioners, and the batched solve."""
    print("case zero_start on/off, mixed=%d galerkin=%d" % (mixed, galerkin), flush=True)
    ctx.set_option("galerkin", galerkin)
    try:
        for (u, r, d), case in zip(probs, CASES):
            out = []
            for zs in (1, 0):
                ctx.set_option("zero_start", zs)
                try:
                    phi = torch.zeros_like(r)
                    st, ok = ctx.solve(phi, r.clone(), d, u, tol=1e-8, max_cycles=200, schedule=1, nu_pre=2,
                                       nu_post=2, nu_coarse=8, mixed=mixed, nu=5, krylov=True)
                finally:
                    ctx.set_option("zero_start", 1)
                assert ok, case
                out.append((phi.cpu().numpy(), [st.relres[i] for i in range(st.cycles)]))
            (x1, r1), (x0, r0) = out
            assert len(r1) > 3, case
            assert r1 == r0, case
            assert np.array_equal(x1, x0), case
        if mixed == 0:
            xs1, st1, ok1 = run_batch(ctx, probs)
            ctx.set_option("zero_start", 0)
            try:
                xs0, st0, ok0 = run_batch(ctx, probs)
            finally:
                ctx.set_option("zero_start", 1)
            assert ok1 and ok0
            for a, b, s1, s0 in zip(xs1, xs0, st1, st0):
                assert s1.cycles == s0.cycles
                assert np.array_equal(a, b)
    finally:
        ctx.set_option("galerkin", 1)


def test_batch_group_streams_equal_one_stream(ctx, probs):
    """The size groups iterate side by side on streams of their own (option batch_streams); with
    the option off they take turns on one stream.  Every result and statistic is the same."""
    print("case batch_streams on/off %s" % CASES, flush=True)
    out = []
    for bs in (1, 0):
        ctx.set_option("batch_streams", bs)
        try:
            xs, st, ok = run_batch(ctx, probs)
        finally:
            ctx.set_option("batch_streams", 1)
        assert ok
        out.append((xs, [(s.cycles, s.relres, s.converged) for s in st]))
    (x1, s1), (x0, s0) = out
    assert s1 == s0
    for a, b in zip(x1, x0):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("pers,nu_top", [(32, 0), (64, 1)])
def test_batch_one_workgroup_per_grid(ctx, probs, pers, nu_top):
    """Groups of n <= batch_persistent solved whole in one workgroup per grid (k_pcg_stack) reach
    the tolerance and agree with the multi-launch batched solve to 1e-6 of the solution (the two
    differ in summation order and in the first sweep from zero); the other groups are unchanged;
    the true residual of every realisation is within the tolerance."""
    print("case batch persistent <= %d nu_top %d" % (pers, nu_top), flush=True)
    ctx.set_option("nu_top", nu_top)
    try:
        xs0, st0, ok0 = run_batch(ctx, probs)
        ctx.set_option("batch_persistent", pers)
        xs1, st1, ok1 = run_batch(ctx, probs)
    finally:
        ctx.set_option("batch_persistent", 0)
        ctx.set_option("nu_top", 0)
    assert ok0 and ok1
    for x0, x1, s0, s1, (u, r, d), case in zip(xs0, xs1, st0, st1, probs, CASES):
        n = case[0]
        assert s1.converged and s1.relres <= SOLVE["tol"] and np.isfinite(x1).all()
        if n > pers:
            assert np.array_equal(x0, x1) and s0.cycles == s1.cycles
        else:
            assert abs(s1.cycles - s0.cycles) <= 2, (case, s0.cycles, s1.cycles)
            assert np.abs(x1 - x0).max() <= 1e-6 * np.abs(x0).max(), case
        # true residual of the returned solution against the projected right-hand side
        b = r.clone()
        ctx.solve(torch.zeros_like(r), b, d, u, tol=1e-30, max_cycles=1, schedule=1, nu_pre=1, nu_post=1,
                  nu_coarse=1, nu=5)
        res = torch.zeros_like(r)
        ctx.residual(res, torch.from_numpy(x1).cuda(), b, d, u)
        assert float(res.abs().max()) <= 1.05 * SOLVE["tol"] * float(b.abs().max()), case


@pytest.mark.parametrize("fused,galerkin", [(4, 1), (8, 1), (16, 0), (32, 1)])
def test_batch_coarse_levels_one_workgroup_per_grid(ctx, probs, fused, galerkin):
    """The preconditioner's levels up to `fused` run one workgroup per grid (k_coarse_stack): every
    realisation bit-identical to the batched launches of those levels (batch_coarse_fused 0)."""
    print("case batch_coarse_fused %d galerkin %d" % (fused, galerkin), flush=True)
    ctx.set_option("galerkin", galerkin)
    try:
        ctx.set_option("batch_coarse_fused", 0)
        x0, s0, ok0 = run_batch(ctx, probs)
        ctx.set_option("batch_coarse_fused", fused)
        x1, s1, ok1 = run_batch(ctx, probs)
    finally:
        ctx.set_option("batch_coarse_fused", 16)
        ctx.set_option("galerkin", 1)
    assert ok0 and ok1
    for a, b, t0, t1, case in zip(x0, x1, s0, s1, CASES):
        assert t0.cycles == t1.cycles and t0.relres == t1.relres, case
        assert np.array_equal(a, b), case


@pytest.mark.parametrize("pers", [0, 64])
def test_batch_breakdown_item_reported_not_converged(ctx, probs, pers):
    """A realisation whose operator vanishes (zero density: G = 0, so L p = 0 and the V-cycle divides
    by a zero diagonal) breaks the CG recurrence at its first step.  The batch must report it as not
    converged (MMH_ENOCONV for the call) with its iterate left finite, and still solve the realisation
    beside it to the tolerance, with the batched launches and with one workgroup per grid."""
    print("case batch breakdown item, batch_persistent %d" % pers, flush=True)
    u, r, d = probs[0]  # 32^3
    n3 = r.numel()
    u0 = torch.zeros_like(u)
    b0 = torch.from_numpy(np.random.default_rng(5).standard_normal(n3)).cuda()
    items = [(torch.zeros_like(r), r.clone(), d, u), (torch.zeros_like(r), b0, torch.zeros_like(d), u0)]
    ctx.set_option("batch_persistent", pers)
    try:
        st, ok = ctx.solve_batch(items, **SOLVE)
    finally:
        ctx.set_option("batch_persistent", 0)
    torch.cuda.synchronize()
    assert not ok
    assert st[0].converged and st[0].relres <= SOLVE["tol"]
    assert not st[1].converged
    assert bool(torch.isfinite(items[1][0]).all())
    assert np.isfinite(st[1].relres) and st[1].relres > SOLVE["tol"]
