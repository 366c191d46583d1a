"""Every FFT length the library has a compile-time plan for, checked against
an independent result on the GPU:

* ``osc_ccfft3d`` (the strided and contiguous line kernels, FP64) against
  numpy's FFT for each planned length L on each axis, with all four per-axis
  sign mixes of ``ccfft3d`` (code/fft_interface_mod.f90:96-167: axis d uses
  exp(+2 pi i jk/n) when idir(d) = +1, unnormalised);
* the solve path (real-row x passes, strided y passes, the fused z
  convolution, the multi-term z inverse of the cathode images, the legacy and
  pipe spectra) on thin meshes whose doubled length along one axis is L and
  16 along the others, against the oracle, in FP64 within 4 x the IGF
  conditioning floor + 2e-14 and in FP32 within 1e-5 of max|E| per component.

Lengths without a plan (prime factors <= 23) go through the runtime-radix
engine; two of those are checked the same way.
"""
import numpy as np
import pytest

import openspacecharge_amd as osc
from oracle.oracle import gaussian_bunch

pytestmark = pytest.mark.gpu

# the doubled lengths 2n of every mesh edge n = 2^a 3^b in [8, 512]
PLANNED = sorted(2 * n for n in {2 ** a * 3 ** b for a in range(10) for b in range(7)}
                 if 8 <= n <= 512)
GENERIC = [10, 14, 46, 100, 200]
GAMMA = 9.78476


def relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))


def relcomp(a, b, comps=3):
    return max(relmax(a[..., c], b[..., c]) for c in range(comps))


def numpy_ccfft3d(a, idir):
    ref = a.copy()
    for ax, s in enumerate(idir):
        ref = np.fft.ifft(ref, axis=ax) * a.shape[ax] if s > 0 else np.fft.fft(ref, axis=ax)
    return ref


@pytest.mark.parametrize("L", PLANNED + GENERIC)
def test_ccfft3d_every_length(ctx, L):
    print(f"ccfft3d L={L}", flush=True)
    rng = np.random.default_rng(L)
    for shape in ((L, 3, 2), (3, L, 2), (2, 3, L)):
        a = np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
        for idir in ((1, 1, 1), (-1, -1, -1), (1, -1, 1), (-1, 1, -1)):
            b = np.zeros_like(a, order="F")
            osc.ccfft3d(a, b, idir, ctx=ctx)
            err = relmax(b, numpy_ccfft3d(a, idir))
            # unnormalised transforms of unit-variance data: round-off grows
            # like log2(L) ulps of the largest output
            assert err < 2e-15 * (np.log2(L) + 4), (shape, idir, err)


@pytest.fixture(scope="module")
def ctx32():
    c = osc.Context(0, precision=osc.OSC_PREC_FP32)
    yield c
    c.close()


def _thin_mesh(ctx, n, axis, cathode):
    nhi = [8, 8, 8]
    nhi[axis] = n
    sig = [1e-3, 1e-3, 1e-3]
    sig[axis] = 1e-3 * max(1.0, n / 16)
    center = (0.0, 0.0, 5 * sig[2]) if cathode else (0.0, 0.0, 0.0)
    x, y, z = gaussian_bunch(20000, sig, seed=n * 3 + axis, center=center)
    m = osc.Mesh3D(nhi=tuple(nhi), gamma=GAMMA)
    osc.deposit_particles(x, y, z, m, total_charge=1e-9, ctx=ctx)
    return m


def _oracle_sc3d(oracle, m, at_cathode):
    """The oracle's space_charge_3d with the never-extracted raw Green planes
    zeroed, as on the GPU, and its conditioning floor (one perturbed run)."""
    run = lambda: np.concatenate(oracle.space_charge_3d(
        m.rho, m.nlo, m.nhi, m.min, m.max, m.delta, m.gamma, at_cathode=at_cathode,
        calc_bfield=True), axis=3)[..., :5]
    oracle.set_zero_unused_planes(True)
    try:
        base, fl = oracle.floor(run, seeds=(12345,))
    finally:
        oracle.set_zero_unused_planes(False)
    return base, 4.0 * fl + 2e-14


@pytest.mark.parametrize("n", [L // 2 for L in PLANNED] + [L // 2 for L in GENERIC[:2]])
def test_space_charge_3d_every_length(ctx, ctx32, oracle, n):
    for axis in range(3):
        for at_cathode in (False, True):
            print(f"space_charge_3d n={n} axis={axis} cathode={at_cathode}", flush=True)
            m = _thin_mesh(ctx, n, axis, at_cathode)
            rho = np.array(m.rho, order="F")
            osc.space_charge_3d(m, at_cathode=at_cathode, calc_bfield=True, ctx=ctx)
            ref, tol = _oracle_sc3d(oracle, m, at_cathode)
            err = relcomp(m.efield, ref[..., :3])
            assert err < tol, (axis, at_cathode, err, tol)
            assert relcomp(m.bfield[..., :2], ref[..., 3:5], 2) < tol
            m.rho = rho
            osc.space_charge_3d(m, at_cathode=at_cathode, calc_bfield=True, ctx=ctx32)
            err32 = relcomp(m.efield, ref[..., :3])
            assert err32 < 1e-5, (axis, at_cathode, err32)


@pytest.mark.parametrize("n", [L // 2 for L in PLANNED] + [L // 2 for L in GENERIC[:2]])
def test_legacy_and_pipe_every_length(ctx, oracle, n):
    """Legacy free space (direct IGF field) and the rectangular pipe, whose
    period 2n - 1 + npad is 2n at npad = 1."""
    one = (1, 1, 1)
    for axis in range(3):
        print(f"legacy/pipe n={n} axis={axis}", flush=True)
        m = _thin_mesh(ctx, n, axis, False)
        osc.space_charge_freespace(m, True, True, ctx=ctx)
        run = lambda: oracle.space_charge_freespace(m.rho, m.nlo, m.nhi, m.npad, m.delta, m.gamma,
                                                    True, True)[2]
        e, fl = oracle.floor(run, seeds=(12345,))
        assert relcomp(m.efield, e) < 4.0 * fl + 2e-14, axis
        # pipe: the mesh spans the pipe cross-section and +-6 sigma_z
        nhi = [8, 8, 8]
        nhi[axis] = n
        a, b, sz = 10e-3, 6e-3, 3e-3
        umin = np.array([-a / 2, -b / 2, -6 * sz])
        dl = -2 * umin / (np.array(nhi) - 1)
        x, y, z = gaussian_bunch(20000, (1e-3, 0.6e-3, sz), seed=5 + axis)
        rho = np.zeros(nhi, dtype=np.float64, order="F")
        osc.depose_rho_scalar(x, y, z, rho, 1e-9 / x.size, one, dl, umin, ctx=ctx)
        osc.osc_alloc_rectpipe_arrays(one, nhi, one, ctx=ctx)
        osc.osc_getgrnpipe(GAMMA, a, b, dl, umin, one, ctx=ctx)
        mp = osc.Mesh3D(nhi=tuple(nhi), min=umin, max=-umin, delta=dl, gamma=GAMMA, rho=rho)
        osc.space_charge_rectpipe(mp, a, b, False, True, ctx=ctx)
        gs = oracle.getgrnpipe(one, nhi, one, GAMMA, a, b, dl, umin)
        _, p, e, _ = oracle.space_charge_rectpipe(rho, one, nhi, one, dl, GAMMA, gs, False)
        # the 5x5 mode sum is well conditioned: FFT round-off only
        assert relmax(mp.phi, p) < 1e-13, axis
        assert relcomp(mp.efield, e) < 1e-12, axis
