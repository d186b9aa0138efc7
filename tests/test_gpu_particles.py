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
