"""The gather's return to particle order writes E(3,n) and B(3,n) from
64-particle chunks of result records.  The outputs must be bit-identical
whatever the alignment of the caller's E and B (8-byte aligned pointers
included) and whichever of them is asked for, including a last chunk of
fewer than 64 particles."""
import ctypes as C

import numpy as np
import pytest

import openspacecharge_amd as osc
from oracle.oracle import gaussian_bunch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [64 * 157 + 37, 64 * 3])
def test_unpack_alignment(ctx, n):
    import torch

    dev = torch.device("cuda:0")
    x, y, z = (torch.from_numpy(a).to(dev) for a in gaussian_bunch(n, (1e-3, 1e-3, 3e-3), seed=5))
    m = osc.Mesh3D(nhi=(24, 24, 32), device="cuda")
    osc.deposit_particles(x, y, z, m, total_charge=1e-9, ctx=ctx)
    osc.space_charge_3d(m, calc_bfield=True, ctx=ctx)
    E0, B0 = osc.interpolate_field(x, y, z, m, E=True, B=True, ctx=ctx)
    E1 = osc.interpolate_field(x, y, z, m, E=True, B=False, ctx=ctx)
    B1 = osc.interpolate_field(x, y, z, m, E=False, B=True, ctx=ctx)
    assert torch.equal(E0, E1) and torch.equal(B0, B1)

    # E and B one double past a 16-byte boundary, through the C-ABI
    eb = torch.full((3 * n + 1,), np.nan, dtype=torch.float64, device=dev)
    bb = torch.full((3 * n + 1,), np.nan, dtype=torch.float64, device=dev)
    nt = m.n
    mc = m.to_c()
    ctx.check(ctx.lib.osc_interpolate_field(
        ctx.h, C.byref(mc), C.c_void_p(m.efield.data_ptr()), C.c_void_p(m.bfield.data_ptr()), n,
        C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), C.c_void_p(z.data_ptr()),
        C.c_void_p(eb.data_ptr() + 8), C.c_void_p(bb.data_ptr() + 8), None, None))
    torch.cuda.synchronize()
    assert m.efield.numel() == 3 * nt
    assert torch.equal(eb[1:].view(n, 3), E0) and torch.equal(bb[1:].view(n, 3), B0)
    assert torch.isnan(eb[0]) and torch.isnan(bb[0])
    assert torch.count_nonzero(E0) > 0
