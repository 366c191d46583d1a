"""osc_interpolate_field_noalloc (the gather that allocates no device memory:
its scratch is a fixed device array per GPU) called through the C-ABI
against osc_interpolate_field on the same device arrays: E, B, the clamp
flags and the clamped count bit-identical, on two calls in a row, the second
with a smaller mesh and fewer points (the scratch keeps the first call's
records, and its count must start again from zero); a mesh larger than the
scratch is rejected with an error."""
import ctypes as C

import numpy as np
import pytest

import openspacecharge_amd as osc
from openspacecharge_amd import _capi as K
from oracle.oracle import gaussian_bunch

pytestmark = pytest.mark.gpu


def _case(nhi, npts, seed):
    import torch
    rng = np.random.default_rng(seed)
    x, y, z = gaussian_bunch(npts, (1e-3, 1e-3, 2e-3), seed)
    m = osc.Mesh3D(nhi=nhi, gamma=4.0)
    m.min = np.array([-3e-3, -3e-3, -6e-3])
    m.max = -m.min
    m.delta = (m.max - m.min) / (np.array(nhi) - 1)
    # some points past every face, so that clamping is exercised
    k = npts // 50
    x[:k] *= 3.0
    z[k:2 * k] *= 4.0
    ef = rng.standard_normal(tuple(nhi) + (3,))
    bf = rng.standard_normal(tuple(nhi) + (3,))
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    # fields in Fortran order (x fastest, component slowest)
    t = dict(ef=dev(ef.transpose(3, 2, 1, 0)), bf=dev(bf.transpose(3, 2, 1, 0)),
             x=dev(x), y=dev(y), z=dev(z))
    return m, t


def _call(lib, ctx, name, m, t):
    import torch
    n = t["x"].shape[0]
    E = torch.full((n, 3), np.nan, dtype=torch.float64, device="cuda")
    B = torch.full((n, 3), np.nan, dtype=torch.float64, device="cuda")
    cl = torch.full((n,), -1, dtype=torch.int8, device="cuda")
    ncl = C.c_int64(-1)
    mc = m.to_c()
    rc = getattr(lib, name)(ctx.h, C.byref(mc), t["ef"].data_ptr(), t["bf"].data_ptr(), n,
                            t["x"].data_ptr(), t["y"].data_ptr(), t["z"].data_ptr(),
                            E.data_ptr(), B.data_ptr(), cl.data_ptr(), C.byref(ncl))
    assert rc == K.OSC_OK, (name, rc, ctx.lib.osc_ctx_last_error(ctx.h))
    torch.cuda.synchronize()
    return E.cpu().numpy(), B.cpu().numpy(), cl.cpu().numpy(), ncl.value


def test_noalloc_gather_equals_interpolate_field():
    lib = K.load()
    ctx = osc.Context(0)
    for nhi, npts, seed in (((96, 80, 128), 300_000, 1), ((20, 24, 16), 7_000, 2)):
        print(f"noalloc gather {nhi} {npts} points", flush=True)
        m, t = _case(nhi, npts, seed)
        E0, B0, c0, n0 = _call(lib, ctx, "osc_interpolate_field_noalloc", m, t)
        E1, B1, c1, n1 = _call(lib, ctx, "osc_interpolate_field", m, t)
        assert np.array_equal(E0, E1) and np.array_equal(B0, B1)
        assert np.array_equal(c0, c1) and n0 == n1 == int(c1.sum()) > 0
    ctx.close()


def test_noalloc_gather_rejects_a_mesh_beyond_its_scratch():
    lib = K.load()
    ctx = osc.Context(0)
    m, t = _case((8, 8, 8), 100, 3)
    m.nhi = (128, 128, 65)  # 1064960 nodes > OSC_NOALLOC_MAX_NODES
    ncl = C.c_int64(0)
    mc = m.to_c()
    rc = lib.osc_interpolate_field_noalloc(ctx.h, C.byref(mc), t["ef"].data_ptr(), None, 100,
                                           t["x"].data_ptr(), t["y"].data_ptr(), t["z"].data_ptr(),
                                           t["x"].data_ptr(), None, None, C.byref(ncl))
    assert rc == K.OSC_ERR_ARG
    assert b"OSC_NOALLOC_MAX_NODES" in ctx.lib.osc_ctx_last_error(ctx.h)
    ctx.close()
