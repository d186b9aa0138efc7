"""GPU: particle to grid density assignment (haloden.f90) against the oracle's linked-list
restatement.  The per-particle cell counts (nearest-particle fill) are integers and must match
exactly; the density matches to 1 ulp of FP32 (the final FP64 normalising sum is a blocked tree on
the device and sequential in the reference)."""
import numpy as np
import pytest

import refbind as R
from test_particle_density import particles

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.mark.parametrize("ng,npart,seed", [(16, 500, 1), (32, 20000, 2), (64, 150000, 3)])
def test_particle_density_matches_oracle(ctx, ng, npart, seed):
    print("case particle density ng=%d np=%d" % (ng, npart), flush=True)
    xp = particles(npart, ng, seed) * np.float32(2.0)
    den_o, wp_o = R.or_particle_density(xp, ng, 2.0 * ng)
    xd = torch.from_numpy(xp).cuda()
    den = torch.zeros(ng ** 3, dtype=torch.float32, device="cuda")
    wp = torch.zeros(npart, dtype=torch.int32, device="cuda")
    ctx.particle_density(xd, ng, 2.0 * ng, den, wp)
    torch.cuda.synchronize()
    assert np.array_equal(wp.cpu().numpy(), wp_o)
    d = den.cpu().numpy()
    assert np.abs(d - den_o).max() <= np.spacing(den_o).max()
    assert np.allclose(d, den_o, rtol=1.2e-7, atol=0)


@pytest.mark.parametrize("n,levels", [(32, 1), (64, 3)])
def test_coarsen_matches_oracle(ctx, n, levels):
    print("case coarsen n=%d levels=%d" % (n, levels), flush=True)
    x = (np.random.default_rng(n).random(n ** 3) + 0.5).astype(np.float32)
    ref = R.or_coarsen(x, n, levels)
    out = torch.zeros((n >> levels) ** 3, dtype=torch.float32, device="cuda")
    ctx.coarsen(torch.from_numpy(x).cuda(), n, levels, out)
    got = out.cpu().numpy()
    # block sums are bit-identical; the mean-1 scaling uses an FP64 total summed in another order
    assert np.abs(got - ref).max() <= np.spacing(ref).max()


@pytest.mark.parametrize("ng,nplot", [(32, 4), (64, 1)])
def test_column_density_matches_oracle(ctx, ng, nplot):
    print("case column density ng=%d nplot=%d" % (ng, nplot), flush=True)
    rng = np.random.default_rng(ng)
    den = (rng.random(ng ** 3) + 0.5).astype(np.float32)
    d = rng.standard_normal(ng ** 3)
    for _ in range(3):
        R.or_nsmooth(d, ng)
    defn = (d * 3.0).astype(np.float32)
    ref = R.or_column_density(den, defn, ng, nplot)
    proj = torch.zeros(ng * ng, dtype=torch.float32, device="cuda")
    ctx.column_density(torch.from_numpy(den).cuda(), torch.from_numpy(defn).cuda(), ng, nplot, proj)
    assert np.allclose(proj.cpu().numpy(), ref, rtol=1e-6, atol=0)


@pytest.mark.parametrize("nf,ng,nplot", [(8, 32, 1), (16, 128, 3), (32, 256, 1)])
def test_gridfigure_matches_oracle(ctx, nf, ng, nplot):
    """gridfigure.f90:41-131 on the device against the oracle: corner displacements (dsp_proj.dat)
    and node positions bit for bit; the column density to 1e-6 (den's mean-1 scaling uses an FP64
    total summed in another order)."""
    print("case gridfigure nf=%d ng=%d nplot=%d" % (nf, ng, nplot), flush=True)
    rng = np.random.default_rng(nf + ng)
    d = rng.standard_normal(nf ** 3).astype(np.float32)
    for _ in range(3):
        R.or_nsmooth(d, nf)
    defn = (d * np.float32(3.0)).astype(np.float32)
    den = (rng.random(ng ** 3) + 0.5).astype(np.float32)
    pos_o, proj_o, slab_o = R.or_gridfigure(defn, nf, den, ng, nplot)
    f32 = dict(dtype=torch.float32, device="cuda")
    pos, proj, slab = torch.zeros(2 * nf * nf, **f32), torch.zeros(ng * ng, **f32), torch.zeros(slab_o.size, **f32)
    ctx.gridfigure(torch.from_numpy(defn).cuda(), torch.from_numpy(den).cuda(), nplot, pos, proj, slab)
    torch.cuda.synchronize()
    assert np.array_equal(slab.cpu().numpy(), slab_o)
    assert np.array_equal(pos.cpu().numpy(), pos_o)
    assert np.allclose(proj.cpu().numpy(), proj_o, rtol=1e-6, atol=0)
