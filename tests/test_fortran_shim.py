"""The Fortran drop-in module (openspacecharge_amd/fortran/open_spacecharge_mod.f90)
driven as a Fortran caller drives the upstream module (tests/fortran/drive_shim.f90),
compared bit for bit with the same calls through the Python binding of the
C-ABI: both reach the same GPU kernels, so any difference would be a
marshalling error (array order, bounds, optional arguments, mesh struct)."""
import os
import subprocess

import numpy as np
import pytest

import openspacecharge_amd as osc
from openspacecharge_amd.build import F_DRIVER
from oracle.oracle import gaussian_bunch

pytestmark = pytest.mark.gpu

GAMMA = 5.0e6 / 0.51099891e6


def _read(path, counts):
    raw = np.fromfile(path, dtype=np.float64)
    out, o = [], 0
    for c in counts:
        out.append(raw[o:o + c])
        o += c
    assert o == raw.size, (o, raw.size)
    return out


@pytest.mark.skipif(not os.path.exists(F_DRIVER), reason="Fortran drop-in not built (amdflang)")
@pytest.mark.parametrize("shape,n", [((16, 12, 20), 5000), ((8, 8, 8), 5000), ((64, 64, 64), 20000)])
def test_fortran_dropin_matches_c_abi(ctx, oracle, tmp_path, shape, n):
    nx, ny, nz = shape
    x, y, z = gaussian_bunch(n, (1e-3, 0.6e-3, 2e-3), 31)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        f.write(np.int64(n).tobytes())
        f.write(x.tobytes())
        f.write(y.tobytes())
        f.write(z.tobytes())
    print(f"drive_shim {shape}", flush=True)
    r = subprocess.run([F_DRIVER, str(fin), str(fout), str(nx), str(ny), str(nz)], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Error, must specify qa or total_charge" in r.stdout
    assert "drive_shim ok" in r.stdout
    ng = nx * ny * nz
    m1 = min(n, 10000)
    (mn, mx, dl, charge, rho, ef, bf, E, B, e4, Em, Bm, secs, eo, bo, ef_leg, pphi, pef) = _read(
        fout, [3, 3, 3, 1, ng, 3 * ng, 3 * ng, 3 * n, 3 * n, 12, 3 * m1, 3 * m1, 1, 3, 3, 3 * ng,
               ng, 3 * ng])
    F = lambda a, s: a.reshape(s, order="F")

    m = osc.Mesh3D(nhi=shape, gamma=GAMMA)
    osc.deposit_particles(x, y, z, m, total_charge=1e-9)
    assert np.array_equal(mn, m.min) and np.array_equal(mx, m.max) and np.array_equal(dl, m.delta)
    assert charge[0] == m.charge
    assert np.array_equal(F(rho, shape), m.rho)
    osc.space_charge_3d(m, calc_bfield=True)
    assert np.array_equal(F(ef, shape + (3,)), m.efield)
    assert np.array_equal(F(bf, shape + (3,)), m.bfield)
    Ep, Bp = osc.interpolate_field(x, y, z, m, E=True, B=True)
    assert np.array_equal(E.reshape(n, 3), Ep) and np.array_equal(B.reshape(n, 3), Bp)
    assert np.array_equal(e4.reshape(4, 3), Ep[:4])
    # per-point calls on the host (upstream order): bit-identical to the
    # device gather and to the oracle, in microseconds per call
    assert np.array_equal(Em.reshape(m1, 3), Ep[:m1]) and np.array_equal(Bm.reshape(m1, 3), Bp[:m1])
    Eo, Bo, _ = oracle.interpolate_field((1, 1, 1), shape, m.min, m.delta, m.efield, m.bfield,
                                         x[:m1], y[:m1], z[:m1])
    assert np.array_equal(Em.reshape(m1, 3), Eo.T) and np.array_equal(Bm.reshape(m1, 3), Bo.T)
    print(f"{m1} per-point interpolate_field calls: {secs[0] * 1e3:.2f} ms")
    assert secs[0] < 0.05 * m1 / 1e4
    # a point past the upper x edge is clamped, with upstream's message
    xo = np.array([m.max[0] + 2 * m.delta[0]])
    Ec, Bc, cl = oracle.interpolate_field((1, 1, 1), shape, m.min, m.delta, m.efield, m.bfield, xo,
                                          y[:1], z[:1])
    assert cl[0] == 1 and np.array_equal(eo, Ec[:, 0]) and np.array_equal(bo, Bc[:, 0])
    assert "ierror: ip=" in r.stdout
    osc.space_charge_freespace(m, direct_field_calc=True, integrated_green_function=True)
    assert np.array_equal(F(ef_leg, shape + (3,)), m.efield)

    a, b = 10e-3, 6e-3
    umin = np.array([-a / 2, -b / 2, -6 * 3e-3])
    pm = osc.Mesh3D(nhi=shape, gamma=GAMMA, min=umin, max=-umin,
                    delta=(-2 * umin) / (np.array(shape) - 1))
    osc.deposit_particles(x, y, z, pm, total_charge=1e-9, resize_mesh=False)
    osc.osc_alloc_rectpipe_arrays(pm.nlo, pm.nhi, pm.npad)
    osc.osc_getgrnpipe(GAMMA, a, b, pm.delta, umin, pm.npad)
    osc.space_charge_rectpipe(pm, a, b, direct_field_calc=False)
    assert np.array_equal(F(pphi, shape), pm.phi)
    assert np.array_equal(F(pef, shape + (3,)), pm.efield)
    # the Fortran run wrote and read back the upstream Green file
    assert os.path.getsize(tmp_path / "rectpipegrn.dat") > 0
