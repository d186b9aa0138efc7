"""GPU: binned power, the loss gradient by the hand-written adjoint, central finite differences
at the north-star configuration (128^3 and 256^3 FP64, lognormal sigma = 1.5, seed 99, 16
log-spaced bins, N(0,1) weights, eps = 1e-6 on 32 random cells), the torch.autograd harness and
the Fisher pipeline.  Tolerances are stated per test."""
import numpy as np
import pytest

import adjoint_ref as A
import refbind as R

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

NB = 16


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def weights(seed=2024):
    return np.random.default_rng(seed).standard_normal(NB)


@pytest.mark.parametrize("n", [32, 64])
def test_power_matches_numpy(ctx, n):
    print("case power n=%d" % n, flush=True)
    x = R.or_lognormal(n, 1.0, 5) - 1.0
    pk, nm = ctx.power(dev(x), NB)
    pr, nr = A.power(x, n, NB)
    assert np.array_equal(nm, nr)
    assert np.abs(pk - pr).max() <= 1e-12 * np.abs(pr).max()


def test_gradient_matches_cpu_restatement(ctx):
    """GPU adjoint (PCG solves to 1e-12) against the CPU restatement (V-cycle solves to 1e-13):
    loss within 1e-10, gradient within 1e-7 of its max-norm."""
    n = 32
    print("case adjoint vs CPU n=%d" % n, flush=True)
    rho = R.or_lognormal(n, 1.5, 99)
    w = weights()
    g = torch.zeros(n ** 3, dtype=torch.float64, device="cuda:0")
    loss, pk, st = ctx.recon_power_grad(dev(rho), w, g)
    lref, gref = A.gradient(rho, n, w)
    assert st.ncompress == 0
    assert abs(loss - lref) <= 1e-10 * abs(lref)
    gg = host(g)
    print("max |g - g_cpu| / max|g_cpu| = %.2e" % (np.abs(gg - gref).max() / np.abs(gref).max()), flush=True)
    assert np.abs(gg - gref).max() <= 1e-7 * np.abs(gref).max()


def fd_gpu(ctx, rho_d, n, w, cells, eps, phi0):
    """Central differences of the loss through the device forward stages, without cancellation:
    each perturbed solution is phi0 + e with L(rho+-) e = rhs(rho+-) - L(rho+-) phi0 solved for the
    correction e, and loss(+) - loss(-) is formed from the difference and the sum of the two
    reconstructed densities (the loss is a quadratic form in delta).  phi0 and the correction solves
    use the rebuilt coarse operators (option galerkin = 0), whose rounding floor on these right-hand
    sides (about 1e-9 of the operator scale) is the one characterised below; the gradient under
    test uses the library's defaults."""
    f64 = dict(dtype=torch.float64, device="cuda:0")
    b = A.mode_bins(n, NB)
    nm = np.array([(b == q).sum() for q in range(NB)], dtype=np.float64)
    W = np.zeros(b.shape)
    for q in range(NB):
        W[b == q] = w[q] / (nm[q] * n ** 3)
    Wd = dev(W)
    d0 = torch.zeros(n ** 3, **f64)
    out = []
    for x in cells:
        phis = []
        for sgn in (1.0, -1.0):
            r = rho_d.clone()
            r[x] += sgn * eps
            u = torch.zeros(n ** 3, 5, **f64)
            u[:, 0] = r
            u = u.contiguous().view(-1)
            rhs = torch.zeros(n ** 3, **f64)
            ctx.limiter_rhs(rhs, d0, u)
            res = torch.zeros(n ** 3, **f64)
            ctx.residual(res, phi0, rhs, d0, u)
            e = torch.zeros(n ** 3, **f64)
            # the correction's residual floors at 1e-8 (128^3) to 4e-8 (256^3) of its own right-hand
            # side (FP64 rounding of L e against the operator's conditioning); 1e-7 is reached
            ctx.set_option("galerkin", 0)
            try:
                st, ok = ctx.solve(e, res, d0, u, tol=1e-7, max_cycles=200, schedule=1, nu_pre=2, nu_post=2,
                                   nu_coarse=8, nu=5, krylov=True)
            finally:
                ctx.set_option("galerkin", 1)
            assert ok, st.relres[st.cycles - 1]
            phis.append(e)
        dd, sd = phis[0] - phis[1], 2 * phi0 + phis[0] + phis[1]
        for v in (dd, sd):
            ctx.nsmooth(v)
            ctx.nsmooth(v)
        ddl, sdl = torch.zeros_like(dd), torch.zeros_like(sd)
        ctx.defp2delta(dd, ddl)
        ctx.defp2delta(sd, sdl)
        Fd = torch.fft.fftn(ddl.view(n, n, n))
        Fs = torch.fft.fftn(sdl.view(n, n, n))
        out.append(float((Wd * (Fd * Fs.conj()).real).sum()) / (2 * eps))
    return np.array(out)


@pytest.mark.parametrize("n", [128, 256])
def test_gradient_central_differences(ctx, n):
    """dloss/drho from the adjoint against central differences (eps = 1e-6) at 32 random cells:
    agreement within 1e-5 of the largest of the 32 finite differences."""
    print("case adjoint FD n=%d" % n, flush=True)
    rho = R.or_lognormal(n, 1.5, 99)
    w = weights()
    rho_d = dev(rho)
    g = torch.zeros(n ** 3, dtype=torch.float64, device="cuda:0")
    loss, pk, st = ctx.recon_power_grad(rho_d, w, g, tol=1e-12)
    print("loss %.6e cycles fwd %d adj %d ncompress %d" % (loss, st.cycles_forward, st.cycles_adjoint,
                                                             st.ncompress), flush=True)
    # the unperturbed mesh solve, as the starting point of the perturbed ones
    f64 = dict(dtype=torch.float64, device="cuda:0")
    u = torch.zeros(n ** 3, 5, **f64)
    u[:, 0] = rho_d
    u = u.contiguous().view(-1)
    d0 = torch.zeros(n ** 3, **f64)
    rhs = torch.zeros(n ** 3, **f64)
    ctx.limiter_rhs(rhs, d0, u)
    phi0 = torch.zeros(n ** 3, **f64)
    ctx.set_option("galerkin", 0)  # the finite-difference reference runs on the rebuilt coarse operators
    try:
        st0, ok = ctx.solve(phi0, rhs, d0, u, tol=1e-12, max_cycles=200, schedule=1, nu_pre=2, nu_post=2,
                            nu_coarse=8, nu=5, krylov=True)
    finally:
        ctx.set_option("galerkin", 1)
    assert ok
    cells = np.random.default_rng(99).choice(n ** 3, 32, replace=False)
    fd = fd_gpu(ctx, rho_d, n, w, cells, 1e-6, phi0)
    gg = host(g)[cells]
    err = np.abs(gg - fd).max() / np.abs(fd).max()
    print("max |adjoint - FD| / max |FD| = %.2e over 32 cells" % err, flush=True)
    assert err <= 1e-5


def test_torch_autograd_harness(ctx):
    """loss.backward() through the autograd.Function gives the adjoint gradient, and its
    directional derivative along a smooth direction matches central differences (eps = 1e-6)."""
    from torch_recon import recon_power_loss
    n = 64
    print("case torch harness n=%d" % n, flush=True)
    rho = dev(R.or_lognormal(n, 1.5, 99)).requires_grad_()
    w = torch.tensor(weights(), dtype=torch.float64)
    loss = recon_power_loss(rho, w, ctx)
    loss.backward()
    g = torch.zeros(n ** 3, dtype=torch.float64, device="cuda:0")
    l2, _, _ = ctx.recon_power_grad(rho.detach(), w.numpy(), g)
    assert float(loss) == l2
    assert torch.equal(rho.grad, g)
    v = np.random.default_rng(7).standard_normal(n ** 3)
    for _ in range(3):
        R.or_nsmooth(v, n)
    vd = dev(v)
    eps = 1e-6
    lp, _, _ = ctx.recon_power_grad(rho.detach() + eps * vd, w.numpy(), None, tol=1e-13)
    lm, _, _ = ctx.recon_power_grad(rho.detach() - eps * vd, w.numpy(), None, tol=1e-13)
    dd = (lp - lm) / (2 * eps)
    gv = float((rho.grad * vd).sum())
    print("directional: adjoint %.10e FD %.10e" % (gv, dd), flush=True)
    assert abs(gv - dd) <= 1e-5 * abs(dd)


def test_fisher_pipeline(ctx):
    """Binned reconstructed power of 8 realisations and its derivative with respect to the
    density-contrast amplitude A (rho_A = 1 + A (rho - 1), at A = 1) from one adjoint per bin,
    fed to the calcfisherinfo-style statistics.  The derivative vector is checked against central
    differences in A."""
    import mmh_amd
    n = 32
    print("case fisher n=%d" % n, flush=True)
    pks, cs = [], []
    for s in range(8):
        rho = R.or_lognormal(n, 1.5, 99 + s)
        rd = dev(rho)
        c = np.zeros(NB)
        g = torch.zeros(n ** 3, dtype=torch.float64, device="cuda:0")
        for b in range(NB):
            onehot = np.zeros(NB)
            onehot[b] = 1.0
            _, pk, _ = ctx.recon_power_grad(rd, onehot, g)
            c[b] = float((g * (rd - 1.0)).sum())
        pks.append(pk)
        cs.append(c)
        if s == 0:
            eps = 1e-4
            _, pp, _ = ctx.recon_power_grad(1.0 + (1 + eps) * (rd - 1.0), np.ones(NB), None, tol=1e-13)
            _, pm, _ = ctx.recon_power_grad(1.0 + (1 - eps) * (rd - 1.0), np.ones(NB), None, tol=1e-13)
            fd = (pp - pm) / (2 * eps)
            assert np.all(np.isfinite(c))
            print("dP/dA adjoint vs FD max rel %.2e" % (np.abs(c - fd).max() / np.abs(fd).max()), flush=True)
            assert np.abs(c - fd).max() <= 1e-4 * np.abs(fd).max()
    # bins without modes at this size are dropped, as calcfisherinfo drops bins (calcfisherinfo.f90:62)
    keep = np.array(pks).min(0) > 0
    pk_all = np.array(pks)[:, keep]
    cmean = np.array(cs).mean(0)[keep]
    out = mmh_amd.fisher(pk_all, cmean)
    assert np.all(np.isfinite(out["fisher"]))
    assert np.allclose(out["cov"], np.cov(pk_all.T, ddof=1), rtol=1e-10, atol=0)
    print("cumulative Fisher information:", np.array2string(out["fisher"], precision=4), flush=True)


def curved_state(ctx, n, sigma=1.5, seed=99):
    """The start of the second outer iteration: def and density after one reconstruction iteration
    whose mesh solve ran to 1e-10 (option mesh_solve = 2)."""
    f64 = dict(dtype=torch.float64, device="cuda:0")
    rho = dev(R.or_lognormal(n, sigma, seed))
    u = torch.zeros(n ** 3, 5, **f64)
    u[:, 0] = rho
    u = u.contiguous().view(-1)
    d, p = torch.zeros(n ** 3, **f64), torch.zeros(n ** 3, **f64)
    ctx.set_option("mesh_solve", 2)
    ctx.set_option("mesh_tol", 10)
    try:
        ctx.reconstruct(u, d, p, 1)
    finally:
        ctx.set_option("mesh_solve", 0)
        ctx.set_option("mesh_tol", 8)
    return u.view(n ** 3, 5)[:, 0].contiguous(), d


def fd_mesh(ctx, rho_d, d_d, n, w, cells, eps, phi0, which):
    """Central differences of the curved-mesh loss (delta = defp2delta(def + defp)) with respect to
    rho or def at the given cells, formed as in fd_gpu: perturbed solutions phi0 + e with the
    correction e solved from the perturbed residual, the loss difference from the sum and the
    difference of the two reconstructed densities."""
    f64 = dict(dtype=torch.float64, device="cuda:0")
    b = A.mode_bins(n, NB)
    nm = np.array([(b == q).sum() for q in range(NB)], dtype=np.float64)
    W = np.zeros(b.shape)
    for q in range(NB):
        W[b == q] = w[q] / (nm[q] * n ** 3)
    Wd = dev(W)
    out = []
    for x in cells:
        fields = []
        for sgn in (1.0, -1.0):
            r, dd = rho_d.clone(), d_d.clone()
            if which == "rho":
                r[x] += sgn * eps
            else:
                dd[x] += sgn * eps
            u = torch.zeros(n ** 3, 5, **f64)
            u[:, 0] = r
            u = u.contiguous().view(-1)
            rhs = torch.zeros(n ** 3, **f64)
            ctx.limiter_rhs(rhs, dd, u)
            # the correction equation L e = rhs - L phi0; the solve projects its right-hand side with
            # rgzerosum at this def, which leaves L phi0 (sum 0) unchanged, so e + phi0 solves the
            # projected equation of the perturbed input
            res = torch.zeros(n ** 3, **f64)
            ctx.residual(res, phi0, rhs, dd, u)
            e = torch.zeros(n ** 3, **f64)
            # the correction by two rounds of iterative refinement: 1e-7 of the right-hand side, then
            # 1e-4 of the first round's residual (recomputed in FP64).  One round leaves the finite
            # differences of d/d(def) a bias of about 1e-9 absolute (4e-5 of the largest at 128^3),
            # and single solves floor near 1e-8 in FP64
            for rnd, ctol in ((0, 1e-7), (1, 1e-4)):
                de = torch.zeros(n ** 3, **f64)
                r = res
                if rnd:
                    r = torch.zeros(n ** 3, **f64)
                    ctx.residual(r, e, res, dd, u)
                ctx.set_option("galerkin", 0)  # as fd_gpu: its floor on such right-hand sides is known
                try:
                    st, ok = ctx.solve(de, r, dd, u, tol=ctol, max_cycles=200, schedule=1, nu_pre=2, nu_post=2,
                                       nu_coarse=8, nu=5, krylov=True)
                finally:
                    ctx.set_option("galerkin", 1)
                assert ok, st.relres[st.cycles - 1]
                e = e + de
            fields.append((e, dd))
        (ep, dp), (em, dm) = fields
        diff, summ = ep - em, 2 * phi0 + ep + em
        for v in (diff, summ):
            ctx.nsmooth(v)
            ctx.nsmooth(v)
        diff = diff + (dp - dm)
        summ = summ + (dp + dm)
        ddl, sdl = torch.zeros_like(diff), torch.zeros_like(summ)
        ctx.defp2delta(diff, ddl)
        ctx.defp2delta(summ, sdl)
        Fd = torch.fft.fftn(ddl.view(n, n, n))
        Fs = torch.fft.fftn(sdl.view(n, n, n))
        out.append(float((Wd * (Fd * Fs.conj()).real).sum()) / (2 * eps))
    return np.array(out)


@pytest.mark.parametrize("n", [64, 128])
def test_curved_mesh_gradient_central_differences(ctx, n):
    """The adjoint at a curved mesh (mmh_recon_power_grad_mesh_d, the second outer iteration of the
    sigma = 1.5, seed 99 field): dloss/drho and dloss/d(def) against central differences
    (eps = 1e-6) at 16 random cells each, within 1e-5 of the largest finite difference.  The
    Cartesian-mesh entry and the mesh entry at def = 0 give the same loss and dloss/drho."""
    print("case curved-mesh adjoint FD n=%d" % n, flush=True)
    f64 = dict(dtype=torch.float64, device="cuda:0")
    rho_d, d_d = curved_state(ctx, n)
    w = weights()
    gr, gd = torch.zeros(n ** 3, **f64), torch.zeros(n ** 3, **f64)
    loss, pk, st = ctx.recon_power_grad_mesh(rho_d, d_d, w, gr, gd, tol=1e-12)
    print("loss %.6e cycles fwd %d adj %d ncompress %d; |def| max %.3f" % (
        loss, st.cycles_forward, st.cycles_adjoint, st.ncompress, float(d_d.abs().max())), flush=True)
    assert st.ncompress == 0
    # the unperturbed solution
    u = torch.zeros(n ** 3, 5, **f64)
    u[:, 0] = rho_d
    u = u.contiguous().view(-1)
    rhs = torch.zeros(n ** 3, **f64)
    ctx.limiter_rhs(rhs, d_d, u)
    phi0 = torch.zeros(n ** 3, **f64)
    st0, ok = ctx.solve(phi0, rhs, d_d, u, tol=1e-13, max_cycles=250, schedule=1, nu_pre=2, nu_post=2,
                        nu_coarse=8, nu=5, krylov=True)
    assert ok
    rng = np.random.default_rng(n)
    for which, g in (("rho", gr), ("def", gd)):
        cells = rng.choice(n ** 3, 16, replace=False)
        fd = fd_mesh(ctx, rho_d, d_d, n, w, cells, 1e-6, phi0, which)
        ga = host(g)[cells]
        err = np.abs(ga - fd).max() / np.abs(fd).max()
        print("  d/d%s: max |adjoint - FD| / max |FD| = %.2e over 16 cells" % (which, err), flush=True)
        assert err <= 1e-5
    # def = 0 reduces to the Cartesian entry
    z = torch.zeros(n ** 3, **f64)
    g1, g2 = torch.zeros(n ** 3, **f64), torch.zeros(n ** 3, **f64)
    l1, _, _ = ctx.recon_power_grad(rho_d, w, g1, tol=1e-12)
    l2, _, _ = ctx.recon_power_grad_mesh(rho_d, z, w, g2, None, tol=1e-12)
    assert abs(l1 - l2) <= 1e-12 * abs(l1)
    assert float((g1 - g2).abs().max()) <= 1e-9 * float(g1.abs().max())


def test_torch_autograd_harness_curved_mesh(ctx):
    """backward() through the curved-mesh autograd.Function gives both adjoint gradients, and the
    directional derivative along smooth directions in (rho, def) jointly matches central differences
    (eps = 1e-6) to 1e-5."""
    from torch_recon import recon_power_loss_mesh
    n = 64
    print("case torch harness curved mesh n=%d" % n, flush=True)
    rho0, d0 = curved_state(ctx, n)
    rho, d = rho0.clone().requires_grad_(), d0.clone().requires_grad_()
    w = torch.tensor(weights(), dtype=torch.float64)
    loss = recon_power_loss_mesh(rho, d, w, ctx)
    loss.backward()
    rng = np.random.default_rng(11)
    vr, vd = rng.standard_normal(n ** 3), rng.standard_normal(n ** 3)
    for _ in range(3):
        R.or_nsmooth(vr, n)
        R.or_nsmooth(vd, n)
    vr, vd = dev(vr), dev(vd * 0.1)
    eps = 1e-6
    lp, _, _ = ctx.recon_power_grad_mesh(rho0 + eps * vr, d0 + eps * vd, w.numpy(), None, None, tol=1e-13)
    lm, _, _ = ctx.recon_power_grad_mesh(rho0 - eps * vr, d0 - eps * vd, w.numpy(), None, None, tol=1e-13)
    fd = (lp - lm) / (2 * eps)
    gv = float((rho.grad * vr).sum() + (d.grad * vd).sum())
    print("directional: adjoint %.10e FD %.10e" % (gv, fd), flush=True)
    assert abs(gv - fd) <= 1e-5 * abs(fd)
