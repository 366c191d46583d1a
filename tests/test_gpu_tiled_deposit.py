"""The tiled charge deposit (particles binned by tiles of cells, each tile's
nodes summed in LDS as 64+64-bit fixed-point integers, the default where the
tile node block fits) against the cell-sorted deposit (OSC_TILED_DEPOSIT=0 at
context creation) and the oracle: rho within 1e-14 of max rho of both (FP32
deposition: 1e-6 of FP64), the same dropped count, bit-identical reruns, and
the fused step's E and B within 1e-12 of the sorted path (relative to the
mesh maximum of each component) (FP32 context: both
within 1e-5 of FP64).  Cases: resized and fixed meshes (particles off the
mesh), per-particle charges of both signs, every particle in one cell, meshes
with one tile and with partial edge tiles, 10^6 particles on 128^3."""
import os

import numpy as np
import pytest

import openspacecharge_amd as osc
from oracle.oracle import gaussian_bunch

pytestmark = pytest.mark.gpu

GAMMA = 4.0


def _ctx(prec, tiled):
    old = os.environ.get("OSC_TILED_DEPOSIT")
    os.environ["OSC_TILED_DEPOSIT"] = "1" if tiled else "0"
    try:
        return osc.Context(0, precision=prec)
    finally:
        if old is None:
            del os.environ["OSC_TILED_DEPOSIT"]
        else:
            os.environ["OSC_TILED_DEPOSIT"] = old


@pytest.fixture(scope="module")
def ctxs():
    c = {(p, t): _ctx(p, t) for p in (osc.OSC_PREC_FP64, osc.OSC_PREC_FP32) for t in (True, False)}
    yield c
    for v in c.values():
        v.close()


def relmax(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def relcomp(a, b):
    return max(float(np.max(np.abs(a[..., c] - b[..., c])) / np.max(np.abs(b[..., c])))
               for c in range(a.shape[-1]))


CASES = {
    "small": dict(nhi=(20, 18, 24), n=30000, sig=(1e-3, 1e-3, 2e-3)),
    "charges": dict(nhi=(33, 17, 40), n=50000, sig=(1e-3, 0.5e-3, 2e-3), q=True),
    "fixed_offmesh": dict(nhi=(24, 24, 24), n=40000, sig=(1e-3, 1e-3, 1e-3), fixed=2e-3),
    "one_tile": dict(nhi=(6, 5, 9), n=20000, sig=(1e-3, 1e-3, 1e-3)),
    "one_cell": dict(nhi=(16, 16, 16), n=20000, sig=(1e-9, 1e-9, 1e-9), fixed=4e-3),
    "full128": dict(nhi=(128, 128, 128), n=1_000_000, sig=(1e-3, 1e-3, 2e-3)),
}


def _bunch(case):
    s = CASES[case]
    x, y, z = gaussian_bunch(s["n"], s["sig"], seed=sorted(CASES).index(case) + 40)
    q = None
    if s.get("q"):
        q = np.random.default_rng(5).uniform(-0.5, 1.5, x.size) * 3e-14
    return x, y, z, q


def _step(ctx, case, x, y, z, q):
    s = CASES[case]
    m = osc.Mesh3D(nhi=s["nhi"], gamma=GAMMA)
    kw = dict(qa=q, total_charge=None if q is not None else 1e-9)
    if "fixed" in s:
        h = s["fixed"]
        m.min, m.max = np.array([-h, -h, -h]), np.array([h, h, h])
        m.delta = (m.max - m.min) / (np.array(s["nhi"]) - 1)
        kw["resize_mesh"] = False
    E, B = osc.space_charge_step(x, y, z, m, calc_bfield=True, ctx=ctx, **kw)
    return m, E, B


@pytest.mark.parametrize("case", sorted(CASES))
def test_tiled_deposit_fp64(ctxs, oracle, case):
    x, y, z, q = _bunch(case)
    print(f"case {case}", flush=True)
    mt, Et, Bt = _step(ctxs[(osc.OSC_PREC_FP64, True)], case, x, y, z, q)
    ms, Es, Bs = _step(ctxs[(osc.OSC_PREC_FP64, False)], case, x, y, z, q)
    assert np.array_equal(mt.min, ms.min) and np.array_equal(mt.delta, ms.delta)
    assert mt.charge == ms.charge
    assert relmax(mt.rho, ms.rho) < 1e-14
    # fields: relative to the mesh maximum of each component (particles at a
    # point charge see a sum of large cancelling node terms)
    assert relcomp(mt.efield, ms.efield) < 1e-12
    for c in range(3):
        emax = float(np.max(np.abs(ms.efield[..., c])))
        assert float(np.max(np.abs(Et[:, c] - Es[:, c]))) < 1e-12 * emax
    for c in range(2):
        bmax = float(np.max(np.abs(ms.bfield[..., c])))
        assert float(np.max(np.abs(Bt[:, c] - Bs[:, c]))) < 1e-12 * bmax
    if CASES[case]["n"] <= 100000:
        s = CASES[case]
        kw = dict(q=q) if q is not None else dict(total_charge=1e-9)
        if "fixed" in s:
            kw.update(resize=False, mn=mt.min, mx=mt.max, delta=mt.delta)
        d = oracle.deposit_particles(x, y, z, (1, 1, 1), s["nhi"], **kw)
        assert relmax(mt.rho, d["rho"]) < 1e-14
    # bit-identical rerun
    mt2, Et2, Bt2 = _step(ctxs[(osc.OSC_PREC_FP64, True)], case, x, y, z, q)
    assert np.array_equal(mt.rho, mt2.rho) and np.array_equal(Et, Et2) and np.array_equal(Bt, Bt2)


@pytest.mark.parametrize("case", ["small", "charges", "fixed_offmesh", "full128"])
def test_tiled_deposit_fp32(ctxs, case):
    x, y, z, q = _bunch(case)
    m64, E64, B64 = _step(ctxs[(osc.OSC_PREC_FP64, False)], case, x, y, z, q)
    mt, Et, Bt = _step(ctxs[(osc.OSC_PREC_FP32, True)], case, x, y, z, q)
    ms, Es, Bs = _step(ctxs[(osc.OSC_PREC_FP32, False)], case, x, y, z, q)
    for m, E, B in ((mt, Et, Bt), (ms, Es, Bs)):
        assert np.array_equal(m.delta, m64.delta)
        assert relmax(m.rho, m64.rho) < 1e-6
        for c in range(3):
            emax = float(np.max(np.abs(m64.efield[..., c])))
            assert float(np.max(np.abs(E[:, c] - E64[:, c]))) < 1e-5 * emax
    mt2, Et2, _ = _step(ctxs[(osc.OSC_PREC_FP32, True)], case, x, y, z, q)
    assert np.array_equal(mt.rho, mt2.rho) and np.array_equal(Et, Et2)


def test_tiled_deposit_dropped_count(ctxs, oracle):
    x, y, z, _ = _bunch("fixed_offmesh")
    res = []
    for tiled in (True, False):
        m = osc.Mesh3D(nhi=(24, 24, 24), gamma=GAMMA)
        m.min, m.max = np.array([-2e-3] * 3), np.array([2e-3] * 3)
        m.delta = (m.max - m.min) / 23
        nd = osc.deposit_particles(x, y, z, m, total_charge=1e-9, resize_mesh=False,
                                   ctx=ctxs[(osc.OSC_PREC_FP64, tiled)])
        res.append((m.rho.copy(), nd))
    d = oracle.deposit_particles(x, y, z, (1, 1, 1), (24, 24, 24), total_charge=1e-9, resize=False,
                                 mn=m.min, mx=m.max, delta=m.delta)
    assert res[0][1] == res[1][1] == d["dropped"] > 0
    assert relmax(res[0][0], res[1][0]) < 1e-14 and relmax(res[0][0], d["rho"]) < 1e-14


def test_tiled_order_reused_on_another_mesh(ctxs, oracle):
    """A deposit keeps its tile-binned order; interpolate_field at the same
    points on a mesh of another origin and spacing takes that order (content
    hash) and must still be bit-exact with the oracle: particles whose cell
    lies outside their bin's tile, and clamped ones, read the planar fields."""
    ctx = ctxs[(osc.OSC_PREC_FP64, True)]
    nhi = (24, 24, 24)
    x, y, z = gaussian_bunch(100000, (1e-3, 1e-3, 1e-3), seed=61)
    m = osc.Mesh3D(nhi=nhi, gamma=GAMMA)
    osc.deposit_particles(x, y, z, m, total_charge=1e-9, ctx=ctx)
    rng = np.random.default_rng(8)
    m2 = osc.Mesh3D(nhi=nhi, min=m.min + 0.3 * m.delta, delta=m.delta * 1.1)
    m2.allocate()
    m2.efield[...] = rng.standard_normal(m2.efield.shape)
    m2.bfield[...] = rng.standard_normal(m2.bfield.shape)
    (E, B), cl = osc.interpolate_field(x, y, z, m2, E=True, B=True, ctx=ctx, return_clamped=True)
    Er, Br, clr = oracle.interpolate_field(m2.nlo, m2.nhi, m2.min, m2.delta, m2.efield, m2.bfield,
                                           x, y, z)
    assert np.array_equal(E, Er.T) and np.array_equal(B, Br.T)
    assert np.array_equal(cl, clr) and clr.sum() > 0
    # and on the deposit's own mesh
    m.allocate()
    m.efield[...] = rng.standard_normal(m.efield.shape)
    m.bfield[...] = rng.standard_normal(m.bfield.shape)
    E, B = osc.interpolate_field(x, y, z, m, E=True, B=True, ctx=ctx)
    Er, Br, _ = oracle.interpolate_field(m.nlo, m.nhi, m.min, m.delta, m.efield, m.bfield, x, y, z)
    assert np.array_equal(E, Er.T) and np.array_equal(B, Br.T)


@pytest.mark.parametrize("prec", [osc.OSC_PREC_FP64, osc.OSC_PREC_FP32])
def test_tiled_deposit_lower_bound_not_one(ctxs, oracle, prec):
    """A mesh whose lower index bound is not 1 (nlo = (3, 2, 5)): the tiled
    deposit's cell rule and weights follow upstream's node offsets; rho vs
    the oracle (FP32 deposition: within 1e-6 of max rho) and the tiled path
    against the cell-sorted one."""
    nlo, nhi = (3, 2, 5), (26, 21, 36)
    x, y, z = gaussian_bunch(120000, (1e-3, 1e-3, 2e-3), seed=71)
    res = {}
    for tiled in (True, False):
        m = osc.Mesh3D(nlo=nlo, nhi=nhi, gamma=GAMMA)
        osc.deposit_particles(x, y, z, m, total_charge=1e-9, ctx=ctxs[(prec, tiled)])
        res[tiled] = m
    d = oracle.deposit_particles(x, y, z, nlo, nhi, total_charge=1e-9)
    mt = res[True]
    assert np.array_equal(mt.delta, d["delta"]) and np.array_equal(mt.min, d["min"])
    tol = 1e-14 if prec == osc.OSC_PREC_FP64 else 1e-6
    assert relmax(mt.rho, d["rho"]) < tol
    assert relmax(mt.rho, res[False].rho) < tol
