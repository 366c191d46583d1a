"""Writes the golden fixtures in this directory from the compiled upstream
library (oracle/_ref/libopensc_ref.so, built by ``make -C oracle ref`` from
the unchanged upstream Fortran sources).

    python tests/golden/make_golden.py            # all groups
    python tests/golden/make_golden.py free       # one group

Each group runs in its own process: upstream keeps its Green arrays in
module globals (code/open_spacecharge_core_mod.f90:22-23) that are allocated
once per grid size, so one process builds one legacy/pipe grid.

Every fixture stores its inputs next to the outputs, so nothing depends on a
random-number stream staying stable across numpy versions. Sizes are small
(8x6x10 physical grids, 2000 particles): the files total a few hundred KB.

Upstream semantics recorded here (file:line of the reference):
  * deposit_particles resize/charge rules   code/open_spacecharge_mod.f90:156-288
  * get_mesh_quantities / depose_rho_scalar code/test_mod.f90:288-358, 410-474
  * osc_get_cgrn_freespace (raw planes kept) code/open_spacecharge_mod.f90:633-707
  * osc_freespace_solver2 / space_charge_3d code/open_spacecharge_mod.f90:431-597
  * legacy/cathode/pipe solvers and Greens  code/open_spacecharge_core_mod.f90
  * rectpipegrn.dat layout                  code/open_spacecharge_core_mod.f90:825-870
  * interpolate_field clamping              code/open_spacecharge_mod.f90:311-390
  * ccfft3d per-axis signs                  code/fft_interface_mod.f90:96-167
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

GAMMA = 5.0e6 / 0.51099891e6          # 5 MeV electrons (code/test_opensc.f90:64,66,102)
NLO, NHI, NPAD = (1, 1, 1), (8, 6, 10), (1, 1, 1)
NHI_LEGACY_GRN = (4, 3, 5)
NP = 2000


def bunch_free():
    from oracle.oracle import gaussian_bunch
    return gaussian_bunch(NP, (1e-3, 1e-3, 1e-4), 11)


def bunch_cathode():
    """sigma=(1,1,0.5) mm centred 1.5 mm from the cathode, z<=0 rejected."""
    from oracle.oracle import gaussian_bunch
    x, y, z = gaussian_bunch(2 * NP, (1e-3, 1e-3, 0.5e-3), 12, center=(0.0, 0.0, 1.5e-3))
    keep = z > 0.0
    return x[keep][:NP].copy(), y[keep][:NP].copy(), z[keep][:NP].copy()


def bunch_pipe():
    from oracle.oracle import gaussian_bunch
    return gaussian_bunch(NP, (1e-3, 0.6e-3, 3e-3), 13)


def part_free(R, out):
    x, y, z = bunch_free()
    q = np.random.Generator(np.random.Philox(14)).uniform(0.5e-12, 1.5e-12, NP)
    d = dict(x=x, y=y, z=z, q=q, gamma=GAMMA)
    r = R.deposit_particles(x, y, z, NLO, NHI, total_charge=1e-9, resize=True)
    d.update(dep_rho=r["rho"], dep_charge=r["charge"], dep_min=r["min"], dep_max=r["max"],
             dep_delta=r["delta"])
    rq = R.deposit_particles(x, y, z, NLO, NHI, q=q, resize=True)
    d.update(depq_rho=rq["rho"], depq_charge=rq["charge"])
    # explicit +-4 sigma mesh, no resize (every particle is inside by the 4 sigma cut)
    mn = -4 * np.array([1e-3, 1e-3, 1e-4]) * (1 + 1e-6)
    mx = -mn
    dl = (mx - mn) / (np.array(NHI) - np.array(NLO))
    rf = R.deposit_particles(x, y, z, NLO, NHI, total_charge=1e-9, resize=False, mn=mn, mx=mx,
                             delta=dl)
    d.update(depf_min=mn, depf_max=mx, depf_delta=dl, depf_rho=rf["rho"])
    # harness mesh (1 guard node per side) and dropping deposit
    dlt, umn = R.get_mesh_quantities(x, y, z, NLO, NHI, 0)
    rh, nf = R.depose_rho_scalar(x, y, z, NLO, NHI, 1e-9 / NP, dlt, umn)
    d.update(mq_delta=dlt, mq_umin=umn, mq_rho=rh, mq_fail=nf)
    # depose onto a mesh one node too small in x: the outer particles are dropped
    rh2, nf2 = R.depose_rho_scalar(x, y, z, NLO, NHI, 1e-9 / NP, dlt, umn + np.array([2 * dlt[0], 0, 0]))
    d.update(mqd_rho=rh2, mqd_fail=nf2)
    # offset-grid Green function (upstream: raw last planes kept)
    rho, dl = r["rho"], r["delta"]
    sh2 = tuple(2 * s for s in rho.shape)
    d["cgrn"] = np.stack([R.get_cgrn_freespace(sh2, dl, GAMMA, ic).real for ic in range(4)])
    off = (1e-4, -2e-4, 3e-3)
    d["cgrn_off_offset"] = np.array(off)
    d["cgrn_off"] = np.stack([R.get_cgrn_freespace(sh2, dl, GAMMA, ic, off).real for ic in range(4)])
    e2, phi2 = R.freespace_solver2(rho, GAMMA, dl)
    d.update(s2_e=e2, s2_phi=phi2)
    e3, b3 = R.space_charge_3d(rho, NLO, NHI, r["min"], r["max"], dl, GAMMA, calc_bfield=True)
    d.update(sc3d_e=e3, sc3d_b=b3)
    # cathode (space_charge_3d at_cathode)
    xc, yc, zc = bunch_cathode()
    rc = R.deposit_particles(xc, yc, zc, NLO, NHI, total_charge=1e-9)
    ec, bc = R.space_charge_3d(rc["rho"], NLO, NHI, rc["min"], rc["max"], rc["delta"], GAMMA,
                               at_cathode=True, calc_bfield=True)
    d.update(xc=xc, yc=yc, zc=zc, cat_rho=rc["rho"], cat_min=rc["min"], cat_max=rc["max"],
             cat_delta=rc["delta"], cat_e=ec, cat_b=bc)
    np.savez_compressed(out, **d)


def part_legacy(R, out):
    x, y, z = bunch_free()
    d = dict(x=x, y=y, z=z, gamma=GAMMA)
    r = R.deposit_particles(x, y, z, NLO, NHI, total_charge=1e-9)
    rho, mn, mx, dl = r["rho"], r["min"], r["max"], r["delta"]
    d.update(rho=rho, min=mn, max=mx, delta=dl)
    for direct, igf in ((1, 1), (0, 1), (0, 0)):
        phi, e, b = R.space_charge_freespace(rho, NLO, NHI, NPAD, mn, mx, dl, GAMMA, direct, igf)
        tag = f"free_d{direct}i{igf}"
        d.update({tag + "_phi": phi, tag + "_e": e, tag + "_b": b})
    np.savez_compressed(out, **d)


def part_legacy_pad(R, out):
    """npad != 1 changes the period and the circular shift (core:151-182, 471-484)."""
    x, y, z = bunch_free()
    npad = (2, 1, 3)
    r = R.deposit_particles(x, y, z, NLO, NHI, total_charge=1e-9)
    phi, e, b = R.space_charge_freespace(r["rho"], NLO, NHI, npad, r["min"], r["max"], r["delta"],
                                         GAMMA, 1, 1)
    np.savez_compressed(out, x=x, y=y, z=z, gamma=GAMMA, npad=np.array(npad), rho=r["rho"],
                        min=r["min"], max=r["max"], delta=r["delta"], e=e, b=b)


def part_cathode(R, out, method):
    xc, yc, zc = bunch_cathode()
    d = dict(x=xc, y=yc, z=zc, gamma=GAMMA, method=method)
    r = R.deposit_particles(xc, yc, zc, NLO, NHI, total_charge=1e-9)
    rho, mn, mx, dl = r["rho"], r["min"], r["max"], r["delta"]
    d.update(rho=rho, min=mn, max=mx, delta=dl)
    for direct in (1, 0):
        phi, e, b, rho_after = R.space_charge_cathodeimages(rho, NLO, NHI, NPAD, mn, mx, dl, GAMMA,
                                                            direct, 1, method)
        d.update({f"d{direct}_phi": phi, f"d{direct}_e": e, f"d{direct}_b": b,
                  f"d{direct}_rho_after": rho_after})
    np.savez_compressed(out, **d)


def part_grn(R, out):
    nhi = NHI_LEGACY_GRN
    dl = np.array([2.1e-4, 3.3e-4, 2.7e-5])
    umin = np.array([-8.4e-4, -9.9e-4, 1.2e-4])
    d = dict(nhi=np.array(nhi), delta=dl, umin=umin, gamma=GAMMA, zshft=1.3e-3)
    d["free"] = np.stack([R.getgrnfree(NLO, nhi, NPAD, dl, GAMMA, ic, 1) for ic in range(4)])
    d["free_coulomb"] = R.getgrnfree(NLO, nhi, NPAD, dl, GAMMA, 0, 0)
    d["shift"] = np.stack([R.getgrnimageshift(NLO, nhi, NPAD, dl, GAMMA, 1.3e-3, ic, 1)
                           for ic in range(4)])
    d["convcorr"] = np.stack([R.getgrnimageconvcorr(NLO, nhi, NPAD, umin, dl, GAMMA, ic, 1)
                              for ic in range(4)])
    np.savez_compressed(out, **d)


def part_pipe(R, out):
    from oracle.oracle import Oracle
    nhi = (6, 6, 8)
    a, b, sz = 10e-3, 6e-3, 3e-3
    umin = np.array([-a / 2, -b / 2, -6 * sz])
    umax = -umin
    dl = np.array([a / (nhi[0] - 1), b / (nhi[1] - 1), 12 * sz / (nhi[2] - 1)])
    x, y, z = bunch_pipe()
    rho, nf = R.depose_rho_scalar(x, y, z, NLO, nhi, 1e-9 / NP, dl, umin)
    cwd = os.getcwd()
    with tempfile.TemporaryDirectory() as td:
        os.chdir(td)
        try:
            R.pipe_green(NLO, nhi, NPAD, GAMMA, a, b, dl, umin, umax, True)
            raw = open("rectpipegrn.dat", "rb").read()
        finally:
            os.chdir(cwd)
    per = Oracle.legacy_period(NLO, nhi, NPAD)
    ne = int(np.prod(per))
    rec1 = int(np.frombuffer(raw[:4], np.int32)[0])
    head = raw[:rec1 + 8]
    body = np.frombuffer(raw[rec1 + 12:rec1 + 12 + 4 * 16 * ne], np.complex128)
    phi, e, bf = R.space_charge_rectpipe(rho, NLO, nhi, NPAD, umin, umax, dl, GAMMA, a, b, False, True)
    np.savez_compressed(out, x=x, y=y, z=z, gamma=GAMMA, nhi=np.array(nhi), apipe=a, bpipe=b,
                        umin=umin, umax=umax, delta=dl, rho=rho, fail=nf,
                        file_head=np.frombuffer(head, np.uint8), file_len=len(raw),
                        greens=body.reshape((4,) + per[::-1]).transpose(0, 3, 2, 1),
                        phi=phi, e=e, b=bf)


def part_interp(R, out):
    x, y, z = bunch_free()
    r = R.deposit_particles(x, y, z, NLO, NHI, total_charge=1e-9)
    g = np.random.Generator(np.random.Philox(15))
    ef = np.asfortranarray(g.standard_normal(r["rho"].shape + (3,)))
    bf = np.asfortranarray(g.standard_normal(r["rho"].shape + (3,)))
    px, py, pz = x[:500].copy(), y[:500].copy(), z[:500].copy()
    # points outside the mesh on every side (clamped, with extrapolating weights)
    mn, mx = r["min"], r["max"]
    outside = [(mn[0] - 1e-3, 0, 0), (mx[0] + 1e-3, 0, 0), (0, mn[1] - 2e-4, 0), (0, mx[1] + 5e-4, 0),
               (0, 0, mn[2] - 1e-4), (0, 0, mx[2] + 1e-5), (mn[0], mn[1], mn[2]), (mx[0], mx[1], mx[2])]
    px = np.concatenate([px, [p[0] for p in outside]])
    py = np.concatenate([py, [p[1] for p in outside]])
    pz = np.concatenate([pz, [p[2] for p in outside]])
    E, B = R.interpolate_field(NLO, NHI, mn, mx, r["delta"], ef, bf, px, py, pz)
    np.savez_compressed(out, min=mn, max=mx, delta=r["delta"], efield=ef, bfield=bf, px=px, py=py,
                        pz=pz, E=E, B=B)


def part_fft(R, out):
    g = np.random.Generator(np.random.Philox(16))
    d = {}
    for t, (shape, idir) in enumerate((((6, 10, 7), (1, -1, 1)), ((8, 9, 5), (-1, -1, -1)),
                                       ((23, 4, 2), (1, 1, -1)))):
        a = g.standard_normal(shape) + 1j * g.standard_normal(shape)
        d[f"in{t}"] = a
        d[f"dir{t}"] = np.array(idir)
        d[f"out{t}"] = R.ccfft3d(a, idir)
    np.savez_compressed(out, **d)


PARTS = {
    "free": lambda R: part_free(R, os.path.join(HERE, "upstream_free.npz")),
    "legacy": lambda R: part_legacy(R, os.path.join(HERE, "upstream_legacy.npz")),
    "legacy_pad": lambda R: part_legacy_pad(R, os.path.join(HERE, "upstream_legacy_pad.npz")),
    "cathode1": lambda R: part_cathode(R, os.path.join(HERE, "upstream_cathode1.npz"), 1),
    "cathode2": lambda R: part_cathode(R, os.path.join(HERE, "upstream_cathode2.npz"), 2),
    "grn": lambda R: part_grn(R, os.path.join(HERE, "upstream_grn.npz")),
    "pipe": lambda R: part_pipe(R, os.path.join(HERE, "upstream_pipe.npz")),
    "interp": lambda R: part_interp(R, os.path.join(HERE, "upstream_interp.npz")),
    "fft": lambda R: part_fft(R, os.path.join(HERE, "upstream_fft.npz")),
}


def main(argv):
    if len(argv) > 1:
        from oracle.oracle import Reference
        R = Reference()
        for p in argv[1:]:
            PARTS[p](R)
            print("wrote", p)
        return 0
    for p in PARTS:
        rc = subprocess.call([sys.executable, os.path.abspath(__file__), p])
        if rc:
            return rc
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
