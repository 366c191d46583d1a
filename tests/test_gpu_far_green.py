"""FP32 Green builds take the cells far from the origin (|w| >= 12 times the
longest cell edge) from the second-order expansion of the field kernel's
cell integral instead of the 8-corner difference of the FP64 primitive, and
evaluate the primitive only on the planes the near cells difference.  The
fused step's E and B of an FP32 context must stay within 1e-5 of FP64 (per
component, relative to the mesh maximum), and within 2e-6 of the same FP32
context with the far-field cells off (OSC_FAR_GREEN=0 at creation), and
differ from it (the far path is taken); two runs bit-identical.  The far
cells apply where the doubled x length has a compiled plan (the x pass that
differences the primitive as it loads); every mesh here has one.  Cases: cells
about 20 times longer in w than across (a relativistic bunch), near-cubic
cells, a cathode-image solve, a bunch off the mesh centre, a 96x96x192 mesh."""
import os

import numpy as np
import pytest

import openspacecharge_amd as osc
from oracle.oracle import gaussian_bunch

pytestmark = pytest.mark.gpu


def _ctx(prec, far):
    old = os.environ.get("OSC_FAR_GREEN")
    os.environ["OSC_FAR_GREEN"] = "1" if far else "0"
    try:
        return osc.Context(0, precision=prec)
    finally:
        if old is None:
            del os.environ["OSC_FAR_GREEN"]
        else:
            os.environ["OSC_FAR_GREEN"] = old


@pytest.fixture(scope="module")
def ctxs():
    c = {"fp64": _ctx(osc.OSC_PREC_FP64, True), "far": _ctx(osc.OSC_PREC_FP32, True),
         "exact": _ctx(osc.OSC_PREC_FP32, False)}
    yield c
    for v in c.values():
        v.close()


CASES = {
    "relativistic": dict(nhi=(48, 48, 96), sig=(0.5e-3, 0.5e-3, 2e-3), gamma=9.78),
    "cubic_cells": dict(nhi=(36, 36, 36), sig=(1e-3, 1e-3, 1e-3), gamma=1.2),
    "cathode": dict(nhi=(32, 32, 96), sig=(0.5e-3, 0.5e-3, 1e-3), gamma=5.0, ctr=(0, 0, 4e-3),
                    cath=True),
    "off_centre": dict(nhi=(48, 32, 72), sig=(0.6e-3, 0.4e-3, 1.5e-3), gamma=8.0,
                       ctr=(0.3e-3, -0.2e-3, 1e-3)),
    "mesh96x192": dict(nhi=(96, 96, 192), sig=(0.5e-3, 0.5e-3, 2e-3), gamma=9.78, n=400000),
}


def _step(ctx, case, x, y, z):
    s = CASES[case]
    m = osc.Mesh3D(nhi=s["nhi"], gamma=s["gamma"])
    E, B = osc.space_charge_step(x, y, z, m, total_charge=1e-9, at_cathode=s.get("cath", False),
                                 calc_bfield=True, ctx=ctx)
    return m, E, B


def _err(a, b, mesh_field, comps):
    return max(float(np.max(np.abs(a[:, c] - b[:, c]))) / float(np.max(np.abs(mesh_field[..., c])))
               for c in comps)


@pytest.mark.parametrize("case", sorted(CASES))
def test_far_green_fp32(ctxs, case):
    s = CASES[case]
    x, y, z = gaussian_bunch(s.get("n", 100000), s["sig"], seed=sorted(CASES).index(case) + 90,
                             center=s.get("ctr", (0, 0, 0)))
    print(f"case {case}", flush=True)
    m64, E64, B64 = _step(ctxs["fp64"], case, x, y, z)
    mf, Ef, Bf = _step(ctxs["far"], case, x, y, z)
    me, Ee, Be = _step(ctxs["exact"], case, x, y, z)
    e64, b64 = _err(Ef, E64, m64.efield, range(3)), _err(Bf, B64, m64.bfield, range(2))
    eoff = _err(Ef, Ee, me.efield, range(3))
    print(f"{case}: E vs FP64 {e64:.2e}, B vs FP64 {b64:.2e}, E vs exact-cell FP32 {eoff:.2e}")
    assert e64 < 1e-5 and b64 < 1e-5
    assert eoff < 2e-6
    assert not np.array_equal(Ef, Ee)  # the far-field cells were taken
    mf2, Ef2, Bf2 = _step(ctxs["far"], case, x, y, z)
    assert np.array_equal(Ef, Ef2) and np.array_equal(Bf, Bf2)
