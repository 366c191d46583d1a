"""The FP32 particle path of the fused step (an FP32 context): deposition with
FP32 weight records and FP32 cell sums, the FP32 convolution, and the gather
with FP32 node records, the deposit's cells and weights reused.  Cell
indices stay the FP64 upstream ones, so the resized mesh is bit-identical to
FP64, and the fields must stay within 1e-5 of the FP64 path (max over
components of max|E32 - E64| / max|E64|; rho within 1e-6 of max rho), also
against the oracle's deposit -> solve -> gather.  Off-mesh particles of a
fixed mesh are dropped from the deposit and clamped by the gather, as in FP64.
Two runs are bit-identical."""
import numpy as np
import pytest

import openspacecharge_amd as osc
from oracle.oracle import gaussian_bunch

pytestmark = pytest.mark.gpu

GAMMA = 9.78476


def relcomp(a, b):
    return max(float(np.max(np.abs(a[..., c] - b[..., c])) / np.max(np.abs(b[..., c])))
               for c in range(a.shape[-1]))


@pytest.fixture(scope="module")
def ctxs():
    c64 = osc.Context(0)
    c32 = osc.Context(0, precision=osc.OSC_PREC_FP32)
    yield c64, c32
    c64.close()
    c32.close()


def _step(ctx, x, y, z, nhi, cath=False, q=None, fixed=None):
    m = osc.Mesh3D(nhi=nhi, gamma=GAMMA)
    kw = dict(qa=q, total_charge=None if q is not None else 1e-9)
    if fixed is not None:
        m.min, m.max = fixed[0].copy(), fixed[1].copy()
        m.delta = (m.max - m.min) / (np.array(nhi) - 1)
        kw["resize_mesh"] = False
    E, B = osc.space_charge_step(x, y, z, m, at_cathode=cath, calc_bfield=True, ctx=ctx, **kw)
    return m, E, B


@pytest.mark.parametrize("case", ["free", "cathode", "charges", "fixed_offmesh", "heavy_cells"])
def test_fp32_step_within_1e5_of_fp64(ctxs, case):
    c64, c32 = ctxs
    nhi, n = (32, 28, 40), 200_000
    cath, q, fixed = False, None, None
    sig, ctr = (1e-3, 0.8e-3, 2e-3), (0, 0, 0)
    if case == "cathode":
        cath, ctr = True, (0, 0, 9e-3)
    if case == "heavy_cells":  # many particles per cell: long runs across waves
        nhi, n, sig = (12, 12, 12), 400_000, (1e-3, 1e-3, 1e-3)
    seed = {"free": 1, "cathode": 2, "charges": 3, "fixed_offmesh": 4, "heavy_cells": 5}[case]
    x, y, z = gaussian_bunch(n, sig, seed=seed, center=ctr)
    if case == "charges":
        q = np.random.default_rng(4).uniform(0.5, 1.5, n) * 5e-15
    if case == "fixed_offmesh":
        fixed = (np.array([-2.5e-3, -2e-3, -5e-3]), np.array([2.5e-3, 2e-3, 5e-3]))
    m64, E64, B64 = _step(c64, x, y, z, nhi, cath, q, fixed)
    m32, E32, B32 = _step(c32, x, y, z, nhi, cath, q, fixed)
    assert np.array_equal(m32.min, m64.min) and np.array_equal(m32.delta, m64.delta)
    assert float(np.max(np.abs(m32.rho - m64.rho))) <= 1e-6 * float(np.max(np.abs(m64.rho)))
    assert relcomp(m32.efield, m64.efield) < 1e-5
    e, b = relcomp(E32, E64), relcomp(B32[:, :2], B64[:, :2])
    print(f"{case}: E at particles {e:.2e}, B {b:.2e}")
    assert e < 1e-5 and b < 1e-5
    assert np.all(B32[:, 2] == 0)
    # deterministic
    _, E32b, B32b = _step(c32, x, y, z, nhi, cath, q, fixed)
    assert np.array_equal(E32, E32b) and np.array_equal(B32, B32b)


def test_fp32_step_vs_oracle(ctxs, oracle):
    """deposit -> space_charge_3d -> interpolate_field of the oracle (FP64,
    raw Green planes zeroed as on the GPU) against the FP32 step."""
    _, c32 = ctxs
    nhi = (20, 18, 24)
    x, y, z = gaussian_bunch(100_000, (1e-3, 1e-3, 3e-3), seed=12)
    m, E32, B32 = _step(c32, x, y, z, nhi)
    d = oracle.deposit_particles(x, y, z, (1, 1, 1), nhi, total_charge=1e-9)
    assert np.array_equal(d["delta"], m.delta)
    oracle.set_zero_unused_planes(True)
    try:
        e, b = oracle.space_charge_3d(d["rho"], (1, 1, 1), nhi, d["min"], d["max"], d["delta"], GAMMA,
                                      calc_bfield=True)
    finally:
        oracle.set_zero_unused_planes(False)
    Eo, Bo, _ = oracle.interpolate_field((1, 1, 1), nhi, d["min"], d["delta"], e, b, x, y, z)
    assert relcomp(E32, Eo.T) < 1e-5 and relcomp(B32[:, :2], Bo.T[:, :2]) < 1e-5
