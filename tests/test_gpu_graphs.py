"""Graph replay of the fused step (osc_ctx_set_graphs): after the first step
of a (shape, precision, solver, B, gamma) with a resized mesh, the Green
build and the spectral solve are recorded as a HIP graph and replayed, the
Green primitive reading the spacing of each new mesh on the device.  Every replay must equal
the eager launches bit for bit (mesh, rho, fields, E and B at the particles):
a sequence of steps with different bunches (so different resized meshes),
fixed meshes and per-particle charges, with eager work of other shapes and
synchronisations between the replays, in FP64 and FP32, free space and at
the cathode, with and without B."""
import os

import numpy as np
import pytest

import openspacecharge_amd as osc
from oracle.oracle import gaussian_bunch

pytestmark = pytest.mark.gpu

GAMMA = 4.0


def _step(ctx, x, y, z, nhi, cath, calcb, q=None, fixed=None):
    m = osc.Mesh3D(nhi=nhi, gamma=GAMMA)
    kw = dict(qa=q, total_charge=None if q is not None else 1e-9)
    if fixed is not None:
        m.min, m.max = fixed[0].copy(), fixed[1].copy()
        m.delta = (m.max - m.min) / (np.array(nhi) - 1)
        kw["resize_mesh"] = False
    out = osc.space_charge_step(x, y, z, m, at_cathode=cath, calc_bfield=calcb, ctx=ctx, **kw)
    E, B = out if calcb else (out, None)
    return m, E, B


def _same(a, b):
    (ma, Ea, Ba), (mb, Eb, Bb) = a, b
    assert np.array_equal(ma.min, mb.min) and np.array_equal(ma.delta, mb.delta)
    assert np.array_equal(ma.rho, mb.rho)
    assert np.array_equal(ma.efield, mb.efield)
    assert np.array_equal(Ea, Eb)
    if Ba is not None:
        assert np.array_equal(ma.bfield, mb.bfield) and np.array_equal(Ba, Bb)


@pytest.mark.parametrize("prec", [osc.OSC_PREC_FP64, osc.OSC_PREC_FP32])
@pytest.mark.parametrize("cath,calcb", [(False, True), (True, True), (False, False)])
def test_graph_replays_equal_eager(prec, cath, calcb):
    cg = osc.Context(0, precision=prec)
    cg.set_graphs(True)
    # graphs off (the default); a context that captures solves builds every
    # FP32 Green cell from the primitive, so the eager one does too
    old = os.environ.get("OSC_FAR_GREEN")
    os.environ["OSC_FAR_GREEN"] = "0"
    try:
        ce = osc.Context(0, precision=prec)
    finally:
        if old is None:
            del os.environ["OSC_FAR_GREEN"]
        else:
            os.environ["OSC_FAR_GREEN"] = old
    nhi = (24, 20, 32)
    ctr = (0, 0, 3e-3) if cath else (0, 0, 0)
    for k in range(5):
        print(f"prec={prec} cathode={cath} B={calcb} step {k}", flush=True)
        sig = (1e-3 * (1 + 0.1 * k), 0.8e-3, 5e-4 * (1 + 0.3 * k))
        x, y, z = gaussian_bunch(30000 + 1000 * k, sig, seed=50 + k, center=ctr)
        q = np.random.default_rng(k).uniform(0.5, 1.5, x.size) * 3e-14 if k == 3 else None
        fixed = None
        if k == 4:
            fixed = (np.array([-5e-3, -4e-3, 5e-4 if cath else -4e-3]),
                     np.array([5e-3, 4e-3, 6e-3 if cath else 4e-3]))
        _same(_step(cg, x, y, z, nhi, cath, calcb, q, fixed),
              _step(ce, x, y, z, nhi, cath, calcb, q, fixed))
        # eager work of another shape on the graph context between replays
        m2 = osc.Mesh3D(nhi=(16, 12, 10), gamma=GAMMA)
        osc.deposit_particles(x, y, z, m2, total_charge=1e-9, ctx=cg)
        osc.space_charge_3d(m2, calc_bfield=True, ctx=cg)
        cg.synchronize()
    st = cg.graph_stats()
    # steps 1-3 replay; the fixed mesh of step 4 runs eagerly (cached Green)
    assert st["captured"] == 1 and st["replays"] == 3, st
    assert ce.graph_stats()["replays"] == 0
    cg.close()
    ce.close()


def test_graph_replay_on_device_arrays():
    """Device-resident particles and mesh arrays: replays equal eager."""
    import torch
    cg = osc.Context(0)
    cg.set_graphs(True)
    ce = osc.Context(0)
    nhi = (32, 32, 32)
    for k in range(4):
        x, y, z = (torch.from_numpy(a).cuda() for a in
                   gaussian_bunch(50000, (1e-3, 1e-3, 2e-3 / (k + 1)), seed=80 + k))
        res = []
        for c in (cg, ce):
            m = osc.Mesh3D(nhi=nhi, gamma=GAMMA)
            E, B = osc.space_charge_step(x, y, z, m, total_charge=1e-9, calc_bfield=True, ctx=c)
            res.append((m.delta.copy(), m.efield.cpu().numpy(), E.cpu().numpy(), B.cpu().numpy()))
        (dg, eg, Eg, Bg), (de, ee, Ee, Be) = res
        assert np.array_equal(dg, de) and np.array_equal(eg, ee)
        assert np.array_equal(Eg, Ee) and np.array_equal(Bg, Be)
    assert cg.graph_stats()["replays"] == 3
    cg.close()
    ce.close()
