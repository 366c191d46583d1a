"""Host-array staging.  Pageable numpy arrays of 4 MB or more cross PCIe
through the context's ring of pinned buffers (16 MB chunks, 4 buffers, host
copies split over threads); smaller ones through HIP.  The results must be
bit-identical to the same calls on device arrays, for arrays that end just
past a chunk boundary and for more data per call than the ring holds (the
ring wraps several times within one array)."""
import numpy as np
import pytest

import openspacecharge_amd as osc
from oracle.oracle import gaussian_bunch

pytestmark = pytest.mark.gpu

CHUNK = 16 << 20


@pytest.mark.parametrize("n", [CHUNK // 8 + 3, 9 * CHUNK // 16 + 1, 1000])
def test_host_arrays_match_device_arrays(ctx, n):
    import torch

    x, y, z = gaussian_bunch(n, (1e-3, 1e-3, 3e-3), seed=41)
    nhi = (32, 32, 48)
    mh = osc.Mesh3D(nhi=nhi)
    nd_h = osc.deposit_particles(x, y, z, mh, total_charge=1e-9, ctx=ctx)
    osc.space_charge_3d(mh, calc_bfield=True, ctx=ctx)
    Eh, Bh = osc.interpolate_field(x, y, z, mh, E=True, B=True, ctx=ctx)

    dev = torch.device("cuda:0")
    xd, yd, zd = (torch.from_numpy(a).to(dev) for a in (x, y, z))
    md = osc.Mesh3D(nhi=nhi, device="cuda")
    nd_d = osc.deposit_particles(xd, yd, zd, md, total_charge=1e-9, ctx=ctx)
    osc.space_charge_3d(md, calc_bfield=True, ctx=ctx)
    Ed, Bd = osc.interpolate_field(xd, yd, zd, md, E=True, B=True, ctx=ctx)

    assert nd_h == nd_d
    assert np.array_equal(mh.min, md.min) and np.array_equal(mh.delta, md.delta)
    for a, b in ((mh.rho, md.rho), (mh.efield, md.efield), (mh.bfield, md.bfield),
                 (Eh, Ed), (Bh, Bd)):
        assert np.array_equal(np.asarray(a), b.cpu().numpy())
    assert np.any(np.asarray(Eh))
