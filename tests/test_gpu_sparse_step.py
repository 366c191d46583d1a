"""The fused step with fewer particles than cells on the cell-sorted
deposit, taken by meshes whose tiles do not fit (cubic meshes above 161^3
nodes) and by contexts created with OSC_TILED_DEPOSIT=0: the node sums load
only the corner sums of occupied cells, and below two particles per node the
gather reads the planar fields instead of packing node records of the whole
mesh.

FP64: rho, the fields and the particles' E and B bit-identical to the three
separate calls (whose interpolate_field reads the planar fields in particle
order below nodes/8 points, and packed node records in the deposit's cell
order above: dense192), rho within 1e-14 of max rho of the oracle's
deposit and of the tiled deposit; a fixed mesh drops and clamps off-mesh
particles as the separate calls do.  FP32: E within 1e-5 of FP64 (per
component, relative to the mesh maximum), two runs bit-identical (the clamp
rule extrapolates particles off a fixed mesh, which multiplies FP32 rounding
by weights far above 1: the fixed mesh leaves them within 1.3 sigma of its
faces, as FP32 users of fixed meshes must)."""
import os

import numpy as np
import pytest

import openspacecharge_amd as osc
from oracle.oracle import gaussian_bunch

pytestmark = pytest.mark.gpu

GAMMA = 6.0


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


# n below the cells in every case (below nodes / 8 but for dense192)
CASES = {
    "resized": dict(nhi=(64, 48, 80), n=20000, sig=(1e-3, 1e-3, 2e-3)),
    "charges": dict(nhi=(48, 40, 56), n=8000, sig=(1e-3, 0.6e-3, 2e-3), q=True),
    "fixed_offmesh": dict(nhi=(40, 40, 40), n=5000, sig=(1e-3, 1e-3, 1e-3), fixed=2.5e-3),
    "cathode": dict(nhi=(40, 36, 64), n=10000, sig=(1e-3, 1e-3, 1e-3), ctr=(0, 0, 6e-3),
                    cath=True),
    "mesh192": dict(nhi=(192, 192, 192), n=100000, sig=(1e-3, 1e-3, 2e-3)),
    "dense192": dict(nhi=(192, 192, 192), n=2_000_000, sig=(1e-3, 1e-3, 2e-3)),
}


def _bunch(case):
    s = CASES[case]
    assert s["n"] < np.prod(s["nhi"]) // (1 if case == "dense192" else 8)
    x, y, z = gaussian_bunch(s["n"], s["sig"], seed=sorted(CASES).index(case) + 70,
                             center=s.get("ctr", (0, 0, 0)))
    q = None
    if s.get("q"):
        q = np.random.default_rng(9).uniform(-0.5, 1.5, x.size) * 3e-14
    return x, y, z, q


def _mesh(case):
    s = CASES[case]
    m = osc.Mesh3D(nhi=s["nhi"], gamma=GAMMA)
    kw = {}
    if "fixed" in s:
        h = s["fixed"]
        m.min, m.max = np.array([-h, -h, -h]), np.array([h, h, h])
        m.delta = (m.max - m.min) / (np.array(s["nhi"]) - 1)
        kw["resize_mesh"] = False
    return m, kw


def _step(ctx, case, x, y, z, q):
    m, kw = _mesh(case)
    E, B = osc.space_charge_step(x, y, z, m, qa=q, total_charge=None if q is not None else 1e-9,
                                 at_cathode=CASES[case].get("cath", False), calc_bfield=True,
                                 ctx=ctx, **kw)
    return m, E, B


@pytest.mark.parametrize("case", sorted(CASES))
def test_sparse_step_fp64_equals_separate_calls(ctxs, oracle, case):
    s = CASES[case]
    x, y, z, q = _bunch(case)
    print(f"case {case}", flush=True)
    # 192^3 takes the cell-sorted deposit in a default context
    ctx = ctxs[(osc.OSC_PREC_FP64, case in ("mesh192", "dense192"))]
    m2, E2, B2 = _step(ctx, case, x, y, z, q)
    m1, kw = _mesh(case)
    osc.deposit_particles(x, y, z, m1, qa=q, total_charge=None if q is not None else 1e-9,
                          ctx=ctx, **kw)
    osc.space_charge_3d(m1, at_cathode=s.get("cath", False), calc_bfield=True, ctx=ctx)
    E1, B1 = osc.interpolate_field(x, y, z, m1, E=True, B=True, ctx=ctx)
    assert np.array_equal(m1.rho, m2.rho)
    assert np.array_equal(m1.efield, m2.efield) and np.array_equal(m1.bfield, m2.bfield)
    assert np.array_equal(E1, E2) and np.array_equal(B1, B2)
    mt, _, _ = _step(ctxs[(osc.OSC_PREC_FP64, True)], case, x, y, z, q)
    assert relmax(m2.rho, mt.rho) < 1e-14
    if s["n"] <= 20000:
        kwo = dict(q=q) if q is not None else dict(total_charge=1e-9)
        if "fixed" in s:
            kwo.update(resize=False, mn=m2.min, mx=m2.max, delta=m2.delta)
        d = oracle.deposit_particles(x, y, z, (1, 1, 1), s["nhi"], **kwo)
        assert relmax(m2.rho, d["rho"]) < 1e-14


@pytest.mark.parametrize("case", ["resized", "fixed_offmesh", "cathode", "mesh192", "dense192"])
def test_sparse_step_fp32_within_1e5(ctxs, case):
    x, y, z, q = _bunch(case)
    print(f"case {case}", flush=True)
    m64, E64, B64 = _step(ctxs[(osc.OSC_PREC_FP64, False)], case, x, y, z, q)
    c32 = ctxs[(osc.OSC_PREC_FP32, case in ("mesh192", "dense192"))]
    m, E, B = _step(c32, case, x, y, z, q)
    assert np.array_equal(m.delta, m64.delta)
    assert relmax(m.rho, m64.rho) < 1e-6
    for c in range(3):
        emax = float(np.max(np.abs(m64.efield[..., c])))
        assert float(np.max(np.abs(E[:, c] - E64[:, c]))) < 1e-5 * emax
    for c in range(2):
        bmax = float(np.max(np.abs(m64.bfield[..., c])))
        assert float(np.max(np.abs(B[:, c] - B64[:, c]))) < 1e-5 * bmax
    m2, E2, B2 = _step(c32, case, x, y, z, q)
    assert np.array_equal(m.rho, m2.rho) and np.array_equal(E, E2) and np.array_equal(B, B2)
