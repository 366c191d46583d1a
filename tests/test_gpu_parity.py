"""GPU parity of every C-ABI entry point against the oracle (the C
restatement in oracle/osc_oracle.c, itself pinned to the compiled upstream
library by tests/test_oracle_pins.py), at sizes the oracle finishes quickly.

Tolerances (written per test):
  * mesh rules, cell indices (through deposition), interpolation: bit-exact;
  * deposited rho: <= 1e-14 of max (summation order only);
  * IGF-based quantities: <= 4 x the conditioning floor + 2e-14, where the
    floor is the spread of the oracle's own result when each IGF primitive
    value is perturbed by 1 ulp (oracle.floor).  Device log/atan differ from
    glibc by <= 1 ulp, and the upstream 8-corner differencing amplifies that;
    no FP64 evaluation can agree more closely than the floor.  The floor is
    1.4e-13 of max|E| at 8x6x10 and 2.4e-12 at 32^3.
  * the solve path zeroes the never-extracted raw last planes of the
    offset-grid Green function, so it is compared with the oracle run the
    same way (oracle.set_zero_unused_planes); tests/test_oracle_pins.py shows
    that this changes the oracle by no more than its floor and brings it
    closer to the exact convolution.
"""
import os

import numpy as np
import pytest

import openspacecharge_amd as osc
from oracle.oracle import gaussian_bunch, uniform_ellipsoid

pytestmark = pytest.mark.gpu


def relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))


def relcomp(a, b):
    """max over components of max|a-b|/max|b| (field arrays (...,3))."""
    return max(relmax(a[..., c], b[..., c]) for c in range(a.shape[-1]))


def f64(shape):
    return np.zeros(shape, dtype=np.float64, order="F")


def floor_tol(oracle, fn, zero_raw=True):
    """(oracle result, tolerance) with the tolerance 4 x conditioning floor."""
    oracle.set_zero_unused_planes(zero_raw)
    try:
        base, fl = oracle.floor(fn)
    finally:
        oracle.set_zero_unused_planes(False)
    return base, 4.0 * fl + 2e-14


# ---------------------------------------------------------------------------
# FFT
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(16, 24, 32), (64, 8, 12), (5, 7, 9), (30, 10, 48), (128, 4, 2)])
@pytest.mark.parametrize("idir", [(1, 1, 1), (-1, -1, -1), (1, -1, 1), (-1, 1, -1)])
def test_ccfft3d_vs_numpy(ctx, shape, idir):
    rng = np.random.default_rng(sum(shape))
    a = np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    b = np.zeros_like(a, order="F")
    osc.ccfft3d(a, b, idir, ctx=ctx)
    ref = a.copy()
    for ax, s in enumerate(idir):
        ref = np.fft.ifft(ref, axis=ax) * shape[ax] if s > 0 else np.fft.fft(ref, axis=ax)
    assert relmax(b, ref) < 1e-14 * 8


def test_ccfft3d_identity_and_bad_sign(ctx):
    a = np.asfortranarray(np.arange(24, dtype=np.complex128).reshape(2, 3, 4))
    b = np.zeros_like(a, order="F")
    osc.ccfft3d(a, b, (0, 0, 0), ctx=ctx)
    assert np.array_equal(a, b)
    with pytest.raises(osc.OpenSCStop):
        osc.ccfft3d(a, b, (1, 0, 1), ctx=ctx)


# ---------------------------------------------------------------------------
# Green functions
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("icomp", [0, 1, 2, 3])
def test_get_cgrn_freespace(ctx, oracle, icomp):
    shape, delta, gam, off = (16, 12, 20), (1e-4, 2e-4, 0.5e-4), 9.78, (1e-5, -2e-5, 3e-4)
    g = np.zeros(shape, dtype=np.complex128, order="F")
    osc.osc_get_cgrn_freespace(g, delta, gam, icomp, off, ctx=ctx)
    ref, tol = floor_tol(oracle, lambda: oracle.get_cgrn_freespace(shape, delta, gam, icomp, off),
                         zero_raw=False)
    # differences come only from the device log/atan/sqrt (<= 1 ulp) through
    # the 8-corner cancellation
    assert relmax(g, ref) < tol
    assert np.all(g.imag == 0)


@pytest.mark.parametrize("icomp,igf", [(0, 0), (0, 1), (1, 1), (2, 1), (3, 1)])
def test_legacy_green_spectra(ctx, oracle, icomp, igf):
    nlo, nhi, npad, dl, gam = (1, 1, 1), (6, 5, 8), (1, 1, 1), (1e-4, 1.3e-4, 2e-5), 3.0
    g = osc.osc_getgrnfree(nlo, nhi, npad, dl, gam, icomp, igf, ctx=ctx)
    ref, tol = floor_tol(oracle, lambda: oracle.getgrnfree(nlo, nhi, npad, dl, gam, icomp, igf))
    assert relmax(g, ref) < tol
    g = osc.osc_getgrnimageshift(nlo, nhi, npad, dl, gam, 1.1e-3, icomp, igf, ctx=ctx)
    ref, tol = floor_tol(oracle, lambda: oracle.getgrnimageshift(nlo, nhi, npad, dl, gam, 1.1e-3, icomp, igf))
    assert relmax(g, ref) < tol
    um = (-3e-4, -3e-4, 2e-4)
    g = osc.osc_getgrnimageconvcorr(nlo, nhi, npad, um, dl, gam, icomp, igf, ctx=ctx)
    ref, tol = floor_tol(oracle, lambda: oracle.getgrnimageconvcorr(nlo, nhi, npad, um, dl, gam, icomp, igf))
    assert relmax(g, ref) < tol


# ---------------------------------------------------------------------------
# deposition
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("nhi", [(16, 12, 20), (9, 9, 9), (32, 32, 32)])
@pytest.mark.parametrize("use_q", [False, True])
def test_deposit_particles(ctx, oracle, nhi, use_q):
    x, y, z = gaussian_bunch(30000, (1e-3, 1e-3, 1e-4), seed=7)
    q = np.random.default_rng(3).uniform(0.5, 1.5, x.size) * 1e-14 if use_q else None
    m = osc.Mesh3D(nhi=nhi)
    nd = osc.deposit_particles(x, y, z, m, qa=q, total_charge=None if use_q else 1e-9, ctx=ctx)
    ref = oracle.deposit_particles(x, y, z, (1, 1, 1), nhi, q=q, total_charge=None if use_q else 1e-9)
    assert nd == 0 == ref["dropped"]
    # mesh rule bit-exact (open_spacecharge_mod.f90:176-189)
    assert np.array_equal(m.min, ref["min"]) and np.array_equal(m.max, ref["max"])
    assert np.array_equal(m.delta, ref["delta"])
    # summation order differs (cell-sorted vs particle-sequential): round-off only
    assert relmax(m.rho, ref["rho"]) < 1e-14
    assert abs(m.charge - ref["charge"]) <= 1e-13 * abs(ref["charge"])


def test_deposit_deterministic_and_dropped(ctx, oracle):
    x, y, z = gaussian_bunch(50000, (1e-3, 2e-3, 5e-4), seed=11)
    m = osc.Mesh3D(nhi=(16, 16, 16), min=np.array([-2e-3, -4e-3, -1e-3]),
                   max=np.array([2e-3, 4e-3, 1e-3]), delta=np.array([4e-3 / 15, 8e-3 / 15, 2e-3 / 15]))
    nd = osc.deposit_particles(x, y, z, m, total_charge=1e-9, resize_mesh=False, ctx=ctx)
    r1 = np.array(m.rho)
    nd2 = osc.deposit_particles(x, y, z, m, total_charge=1e-9, resize_mesh=False, ctx=ctx)
    assert nd == nd2 and nd > 0
    assert np.array_equal(r1, m.rho)  # bit-identical run to run
    ref = oracle.deposit_particles(x, y, z, (1, 1, 1), (16, 16, 16), total_charge=1e-9, resize=False,
                                   mn=m.min, mx=m.max, delta=m.delta)
    assert ref["dropped"] == nd
    assert relmax(m.rho, ref["rho"]) < 1e-14


def test_deposit_heavy_clustering(ctx, oracle):
    # a pancake bunch puts thousands of particles in a cell (heavy-cell path)
    x, y, z = gaussian_bunch(200000, (1e-3, 1e-3, 1e-6), seed=5)
    m = osc.Mesh3D(nhi=(8, 8, 4))
    osc.deposit_particles(x, y, z, m, total_charge=1e-9, ctx=ctx)
    ref = oracle.deposit_particles(x, y, z, (1, 1, 1), (8, 8, 4), total_charge=1e-9)
    assert relmax(m.rho, ref["rho"]) < 1e-13
    assert abs(float(np.sum(m.rho)) - 1e-9) < 1e-20


def test_deposit_no_charge(ctx, oracle, capsys):
    x, y, z = gaussian_bunch(100, (1, 1, 1), seed=1)
    m = osc.Mesh3D(nhi=(8, 8, 8))
    osc.deposit_particles(x, y, z, m, ctx=ctx)
    assert "must specify qa or total_charge" in capsys.readouterr().out
    # the mesh is resized before the missing charge is found (open_spacecharge_mod.f90:176-189, 249-252)
    ref = oracle.deposit_particles(x, y, z, (1, 1, 1), (8, 8, 8), total_charge=1.0)
    assert np.array_equal(m.min, ref["min"]) and np.array_equal(m.max, ref["max"])
    assert np.array_equal(m.delta, ref["delta"])


def test_deposit_empty_bunch(ctx):
    # minval/maxval of no particles are +huge/-huge upstream; the resize
    # (open_spacecharge_mod.f90:176-189) then overflows to min = +inf,
    # max = -inf, delta = -inf, which the device bounding box (starting from
    # +-inf) reproduces; no particle is deposited
    e = np.zeros(0)
    m = osc.Mesh3D(nhi=(8, 8, 8))
    nd = osc.deposit_particles(e, e, e, m, total_charge=1e-9, ctx=ctx)
    assert nd == 0
    assert np.all(m.min == np.inf) and np.all(m.max == -np.inf) and np.all(m.delta == -np.inf)
    assert m.rho.shape == (8, 8, 8) and not np.any(m.rho)


def test_harness_mesh_and_deposit(ctx, oracle):
    x, y, z = gaussian_bunch(20000, (1e-3, 0.6e-3, 3e-3), seed=3, center=(0, 0, 5e-3))
    for iset in (0, 1):
        dl, um = osc.get_mesh_quantities(x, y, z, (1, 1, 1), (16, 16, 32), iset, ctx=ctx)
        dlr, umr = oracle.get_mesh_quantities(x, y, z, (1, 1, 1), (16, 16, 32), iset)
        assert np.array_equal(dl, dlr) and np.array_equal(um, umr)
    rho = f64((16, 16, 32))
    nf = osc.depose_rho_scalar(x, y, z, rho, 1e-9 / x.size, (1, 1, 1), dl, um, ctx=ctx)
    ref, nfr = oracle.depose_rho_scalar(x, y, z, (1, 1, 1), (16, 16, 32), 1e-9 / x.size, dl, um)
    assert nf == nfr
    assert relmax(rho, ref) < 1e-14


# ---------------------------------------------------------------------------
# free space / cathode (offset-grid IGF)
# ---------------------------------------------------------------------------
def _bunch_mesh(ctx, oracle, nhi, sigma, seed, center=(0, 0, 0), n=20000):
    x, y, z = gaussian_bunch(n, sigma, seed=seed, center=center)
    m = osc.Mesh3D(nhi=nhi, gamma=9.78476)
    osc.deposit_particles(x, y, z, m, total_charge=1e-9, ctx=ctx)
    return m, (x, y, z)


@pytest.mark.parametrize("nhi", [(8, 6, 10), (16, 16, 16), (12, 18, 24), (32, 32, 32)])
@pytest.mark.parametrize("at_cathode", [False, True])
def test_space_charge_3d(ctx, oracle, nhi, at_cathode):
    m, _ = _bunch_mesh(ctx, oracle, nhi, (1e-3, 1e-3, 5e-4), 21, center=(0, 0, 3e-3))
    osc.space_charge_3d(m, at_cathode=at_cathode, calc_bfield=True, ctx=ctx)
    eb, tol = floor_tol(oracle, lambda: np.concatenate(oracle.space_charge_3d(
        m.rho, m.nlo, m.nhi, m.min, m.max, m.delta, m.gamma, at_cathode=at_cathode,
        calc_bfield=True), axis=3)[..., :5])
    assert relcomp(m.efield, eb[..., :3]) < tol
    assert relcomp(m.bfield[..., :2], eb[..., 3:5]) < tol
    assert np.all(m.bfield[..., 2] == 0)
    if nhi == (8, 6, 10):
        assert relcomp(m.efield, eb[..., :3]) < 1e-12


def test_space_charge_3d_offset_no_b(ctx, oracle):
    m, _ = _bunch_mesh(ctx, oracle, (10, 10, 12), (1e-3, 2e-3, 1e-3), 4)
    m.bfield[...] = 7.0
    osc.space_charge_3d(m, offset=(1e-3, 0, 2e-2), calc_bfield=False, ctx=ctx)
    # far from the bunch the IGF differencing loses ~8 digits (floor ~3e-8)
    e, tol = floor_tol(oracle, lambda: oracle.space_charge_3d(
        m.rho, m.nlo, m.nhi, m.min, m.max, m.delta, m.gamma, offset=(1e-3, 0, 2e-2),
        calc_bfield=False)[0])
    assert relcomp(m.efield, e) < tol
    assert np.all(m.bfield == 7.0)  # untouched when calc_bfield is false


def test_freespace_solver2_phi(ctx, oracle):
    m, _ = _bunch_mesh(ctx, oracle, (12, 10, 14), (1e-3, 1e-3, 1e-3), 9)
    e, p = f64(m.shape + (3,)), f64(m.shape)
    osc.osc_freespace_solver2(m.rho, m.gamma, m.delta, efield=e, phi=p, offset=(0, 0, 0), ctx=ctx)
    er, tol = floor_tol(oracle, lambda: oracle.freespace_solver2(m.rho, m.gamma, m.delta, (0, 0, 0))[0])
    pr, tolp = floor_tol(oracle, lambda: oracle.freespace_solver2(m.rho, m.gamma, m.delta, (0, 0, 0))[1])
    assert relcomp(e, er) < tol
    assert relmax(p, pr) < tolp


# ---------------------------------------------------------------------------
# legacy solvers
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("direct,igf", [(True, True), (False, True), (False, False)])
@pytest.mark.parametrize("npad", [(1, 1, 1), (2, 1, 3)])
def test_space_charge_freespace(ctx, oracle, direct, igf, npad):
    m, _ = _bunch_mesh(ctx, oracle, (8, 10, 12), (1e-3, 1e-3, 5e-4), 13)
    m.npad = npad
    osc.space_charge_freespace(m, direct, igf, ctx=ctx)
    run = lambda: oracle.space_charge_freespace(m.rho, m.nlo, m.nhi, npad, m.delta, m.gamma, direct, igf)
    p, tolp = floor_tol(oracle, lambda: run()[1])
    e, tol = floor_tol(oracle, lambda: run()[2])
    b = run()[3]
    if not direct:
        assert relmax(m.phi, p) < tolp
    assert relcomp(m.efield, e) < tol
    assert relcomp(m.bfield[..., :2], b[..., :2]) < tol


def test_freespace_direct_without_igf(ctx, capsys):
    m = osc.Mesh3D(nhi=(8, 8, 8), gamma=2.0, delta=np.ones(3) * 1e-4)
    m.allocate()
    m.efield[...] = 3.0
    osc.space_charge_freespace(m, True, False, ctx=ctx)
    assert "Direct field calc must use integrated Green function" in capsys.readouterr().out
    assert np.all(m.efield == 3.0)


@pytest.mark.parametrize("method", [1, 2])
@pytest.mark.parametrize("direct", [True, False])
def test_space_charge_cathodeimages(ctx, oracle, method, direct):
    m, _ = _bunch_mesh(ctx, oracle, (8, 10, 12), (1e-3, 1e-3, 5e-4), 17, center=(0, 0, 2.5e-3))
    rho0 = np.array(m.rho)
    osc.space_charge_cathodeimages(m, direct, True, method, ctx=ctx)
    run = lambda: oracle.space_charge_cathodeimages(rho0, m.nlo, m.nhi, m.npad, m.min, m.delta,
                                                    m.gamma, direct, True, method)
    rc, p, e, b, rho_after = run()
    _, tolp = floor_tol(oracle, lambda: run()[1])
    _, tol = floor_tol(oracle, lambda: run()[2])
    assert np.array_equal(m.rho, rho_after)  # method 2 flips rho in z, as upstream
    if not direct:
        assert relmax(m.phi, p) < tolp
    assert relcomp(m.efield, e) < tol
    assert relcomp(m.bfield[..., :2], b[..., :2]) < tol


# ---------------------------------------------------------------------------
# rectangular pipe
# ---------------------------------------------------------------------------
def _pipe_case(nhi=(8, 8, 12)):
    a, b, sz = 10e-3, 6e-3, 3e-3
    umin = np.array([-a / 2, -b / 2, -6 * sz])
    dl = np.array([a / (nhi[0] - 1), b / (nhi[1] - 1), 12 * sz / (nhi[2] - 1)])
    x, y, z = gaussian_bunch(20000, (1e-3, 0.6e-3, 3e-3), seed=3)
    return a, b, umin, -umin, dl, (x, y, z)


def test_space_charge_rectpipe(ctx, oracle, tmp_path):
    nlo, nhi, npad = (1, 1, 1), (8, 8, 12), (1, 1, 1)
    a, b, umin, umax, dl, (x, y, z) = _pipe_case(nhi)
    gam = 9.78476
    rho = f64(nhi)
    osc.depose_rho_scalar(x, y, z, rho, 1e-9 / x.size, nlo, dl, umin, ctx=ctx)
    osc.osc_alloc_rectpipe_arrays(nlo, nhi, npad, ctx=ctx)
    osc.osc_getgrnpipe(gam, a, b, dl, umin, npad, ctx=ctx)
    m = osc.Mesh3D(nlo=nlo, nhi=nhi, npad=npad, min=umin, max=umax, delta=dl, gamma=gam, rho=rho)
    osc.space_charge_rectpipe(m, a, b, False, True, ctx=ctx)
    gs = oracle.getgrnpipe(nlo, nhi, npad, gam, a, b, dl, umin)
    rc, p, e, bb = oracle.space_charge_rectpipe(rho, nlo, nhi, npad, dl, gam, gs, False)
    # the 5x5 mode sum is well conditioned: FFT round-off only
    assert relmax(m.phi, p) < 1e-13
    assert relcomp(m.efield, e) < 1e-12
    # file round trip in the upstream format
    path = str(tmp_path / "rectpipegrn.dat")
    osc.osc_write_rectpipe_grn(a, b, dl, umin, umax, nlo, nhi, gam, path=path, ctx=ctx)
    raw = open(path, "rb").read()
    per = oracle.legacy_period(nlo, nhi, npad)
    ne = int(np.prod(per))
    assert len(raw) == 4 + 132 + 4 + 4 + 4 * 16 * ne + 4
    body = np.frombuffer(raw[144:144 + 64 * ne], np.complex128)
    for t in range(4):
        assert relmax(body[t * ne:(t + 1) * ne].reshape(per, order="F"), gs[t]) < 1e-12
    ctx2 = osc.Context(0)
    osc.osc_alloc_rectpipe_arrays(nlo, nhi, npad, ctx=ctx2)
    hdr = osc.osc_read_rectpipe_grn(path, ctx=ctx2)
    assert hdr["period"] == list(per)
    m2 = osc.Mesh3D(nlo=nlo, nhi=nhi, npad=npad, min=umin, max=umax, delta=dl, gamma=gam, rho=rho)
    osc.space_charge_rectpipe(m2, a, b, False, True, ctx=ctx2)
    assert relmax(m2.phi, m.phi) < 1e-14
    ctx2.close()


# ---------------------------------------------------------------------------
# gather
# ---------------------------------------------------------------------------
def test_interpolate_field_bit_exact(ctx, oracle):
    nhi = (12, 10, 14)
    rng = np.random.default_rng(2)
    m = osc.Mesh3D(nhi=nhi, min=np.array([-1e-3, -2e-3, 0.5e-3]), delta=np.array([2e-4, 4e-4, 1e-4]))
    m.allocate()
    m.efield[...] = rng.standard_normal(m.efield.shape)
    m.bfield[...] = rng.standard_normal(m.bfield.shape)
    x = rng.uniform(-1.2e-3, 1.5e-3, 5000)
    y = rng.uniform(-2.2e-3, 2.0e-3, 5000)
    z = rng.uniform(0.4e-3, 2.0e-3, 5000)
    (E, B), cl = osc.interpolate_field(x, y, z, m, E=True, B=True, ctx=ctx, return_clamped=True)
    Er, Br, clr = oracle.interpolate_field(m.nlo, m.nhi, m.min, m.delta, m.efield, m.bfield, x, y, z)
    assert np.array_equal(E, Er.T) and np.array_equal(B, Br.T)
    assert np.array_equal(cl, clr) and clr.sum() > 0


def test_interpolate_field_lower_bound_not_one(ctx, oracle):
    """Fields allocated (nlo:nhi) with nlo = (3, 2, 5): read at element
    ip - nlo (open_spacecharge_mod.f90:369-376), bit-exact with the oracle,
    clamped points included."""
    nlo, nhi = (3, 2, 5), (14, 12, 18)
    rng = np.random.default_rng(12)
    m = osc.Mesh3D(nlo=nlo, nhi=nhi, min=np.array([-1e-3, -2e-3, 0.5e-3]),
                   delta=np.array([2e-4, 4e-4, 1e-4]))
    m.allocate()
    m.efield[...] = rng.standard_normal(m.efield.shape)
    m.bfield[...] = rng.standard_normal(m.bfield.shape)
    x = rng.uniform(-0.8e-3, 1.8e-3, 5000)
    y = rng.uniform(-2.2e-3, 2.6e-3, 5000)
    z = rng.uniform(0.4e-3, 2.3e-3, 5000)
    (E, B), cl = osc.interpolate_field(x, y, z, m, E=True, B=True, ctx=ctx, return_clamped=True)
    Er, Br, clr = oracle.interpolate_field(m.nlo, m.nhi, m.min, m.delta, m.efield, m.bfield, x, y, z)
    assert np.array_equal(E, Er.T) and np.array_equal(B, Br.T)
    assert np.array_equal(cl, clr) and clr.sum() > 0


def test_interpolate_after_deposit_takes_kept_order_only_for_same_points(oracle):
    """A deposit of >= 65536 points keeps its cell order; interpolate_field at
    the same points takes it (content hash), at changed points it sorts anew.
    Either way the fields are bit-exact with the oracle; the stage counts show
    which way was taken (gather.sort: the hash alone, or hash + sort)."""
    c = osc.Context(0)
    x, y, z = gaussian_bunch(70000, (1e-3, 1e-3, 2e-3), 5)
    m = osc.Mesh3D(nhi=(16, 16, 16), gamma=9.78476)
    osc.deposit_particles(x, y, z, m, total_charge=1e-9, ctx=c)
    osc.space_charge_3d(m, calc_bfield=True, ctx=c)

    def check(px, py, pz, scopes):
        c.profile_read()
        c.set_profiling(True)
        (E, B), cl = osc.interpolate_field(px, py, pz, m, E=True, B=True, ctx=c,
                                           return_clamped=True)
        st = c.profile_read()
        c.set_profiling(False)
        Er, Br, clr = oracle.interpolate_field(m.nlo, m.nhi, m.min, m.delta, m.efield, m.bfield,
                                               px, py, pz)
        assert np.array_equal(E, Er.T) and np.array_equal(B, Br.T) and np.array_equal(cl, clr)
        assert st["gather.sort"][1] == scopes, st

    check(x, y, z, 1)  # the deposit's points: its order
    x2 = x.copy()
    x2[123] += 1e-4
    check(x2, y, z, 2)  # one point moved: sorted anew
    xs = x.copy()
    xs[[5, 6]] = xs[[6, 5]]
    check(xs, y, z, 2)  # two x coordinates swapped
    check(x, y, z, 1)  # the deposit's points again
    c.close()


# ---------------------------------------------------------------------------
# FP32 convolution path
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("nhi", [(16, 16, 16), (32, 32, 64)])
def test_fp32_convolution_error(oracle, nhi):
    c32 = osc.Context(0, precision=osc.OSC_PREC_FP32)
    x, y, z = uniform_ellipsoid(50000, (1e-3, 1e-3, 3e-3), seed=2)
    m = osc.Mesh3D(nhi=nhi, gamma=9.78476)
    osc.deposit_particles(x, y, z, m, total_charge=1e-9, ctx=c32)
    osc.space_charge_3d(m, calc_bfield=True, ctx=c32)
    oracle.set_zero_unused_planes(True)
    e, _ = oracle.space_charge_3d(m.rho, m.nlo, m.nhi, m.min, m.max, m.delta, m.gamma, calc_bfield=True)
    oracle.set_zero_unused_planes(False)
    err = relcomp(m.efield, e)
    assert err < 1e-5, err
    c32.close()


# ---------------------------------------------------------------------------
# fused tracking step == the three separate calls
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("at_cathode", [False, True])
@pytest.mark.parametrize("use_q", [False, True])
def test_space_charge_step_equals_separate_calls(ctx, at_cathode, use_q):
    x, y, z = gaussian_bunch(40000, (1e-3, 1e-3, 5e-4), seed=31, center=(0, 0, 3e-3))
    q = np.random.default_rng(8).uniform(0.5, 1.5, x.size) * 2.5e-14 if use_q else None
    tc = None if use_q else 1e-9
    m1 = osc.Mesh3D(nhi=(16, 16, 24), gamma=4.0)
    osc.deposit_particles(x, y, z, m1, qa=q, total_charge=tc, ctx=ctx)
    osc.space_charge_3d(m1, at_cathode=at_cathode, calc_bfield=True, ctx=ctx)
    E1, B1 = osc.interpolate_field(x, y, z, m1, E=True, B=True, ctx=ctx)
    m2 = osc.Mesh3D(nhi=(16, 16, 24), gamma=4.0)
    E2, B2 = osc.space_charge_step(x, y, z, m2, qa=q, total_charge=tc, at_cathode=at_cathode,
                                   calc_bfield=True, ctx=ctx)
    assert np.array_equal(m1.rho, m2.rho)
    assert np.array_equal(m1.efield, m2.efield) and np.array_equal(m1.bfield, m2.bfield)
    assert np.array_equal(E1, E2) and np.array_equal(B1, B2)


def test_interpolate_few_points_planar_path(ctx, oracle):
    # fewer points than nodes/8 take the planar-field kernel
    rng = np.random.default_rng(4)
    m = osc.Mesh3D(nhi=(24, 20, 28), min=np.array([-1e-3, -2e-3, 0.5e-3]),
                   delta=np.array([1e-4, 2e-4, 5e-5]))
    m.allocate()
    m.efield[...] = rng.standard_normal(m.efield.shape)
    m.bfield[...] = rng.standard_normal(m.bfield.shape)
    x, y, z = (rng.uniform(-1.1e-3, 1.4e-3, 37), rng.uniform(-2.1e-3, 2.0e-3, 37),
               rng.uniform(0.4e-3, 1.9e-3, 37))
    E, B = osc.interpolate_field(x, y, z, m, E=True, B=True, ctx=ctx)
    Er, Br, _ = oracle.interpolate_field(m.nlo, m.nhi, m.min, m.delta, m.efield, m.bfield, x, y, z)
    assert np.array_equal(E, Er.T) and np.array_equal(B, Br.T)
    e0 = osc.interpolate_field(x[0], y[0], z[0], m, ctx=ctx)  # scalar call returns E(3)
    assert e0.shape == (3,) and np.array_equal(e0, Er[:, 0])
