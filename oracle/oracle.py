"""ctypes bindings for the two CPU checkers built by ``oracle/Makefile``.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg, never by the product package.

* :class:`Oracle` -- this repository's C restatement (``liboscoracle.so``).
* :class:`Reference` -- the upstream Fortran library compiled unchanged with
  amdflang plus ``ref_capi.f90`` (``_ref/libopensc_ref.so``).

Both expose the same numpy-level calls. Grids are Fortran-ordered numpy arrays
of shape (nx, ny, nz) (fields (nx, ny, nz, 3)).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboscoracle.so")
REF_SO = os.path.join(HERE, "_ref", "libopensc_ref.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _i3(v):
    return (C.c_int * 3)(*[int(t) for t in v])


def _d3(v):
    return (C.c_double * 3)(*[float(t) for t in v])


def _f64(a):
    return np.require(np.asarray(a, dtype=np.float64), requirements=["F", "A", "W"])


def nsize(nlo, nhi):
    return tuple(int(h) - int(l) + 1 for l, h in zip(nlo, nhi))


def _relcomp(a, b):
    """max over trailing components of max|a-b| / max|b| (whole array if 1-D/3-D)."""
    a, b = np.asarray(a), np.asarray(b)
    if a.ndim == 4:
        return max(_relcomp(a[..., c], b[..., c]) for c in range(a.shape[-1]))
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))


class Oracle:
    """The C restatement (oracle/osc_oracle.c)."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle oracle)")
        self.lib = C.CDLL(path)
        L = self.lib
        L.oc_ccfft3d.restype = C.c_int
        L.oc_deposit_particles.restype = C.c_int
        L.oc_depose_rho_scalar.restype = C.c_int64
        L.oc_space_charge_freespace.restype = C.c_int
        L.oc_space_charge_cathodeimages.restype = C.c_int
        L.oc_space_charge_rectpipe.restype = C.c_int
        for f in ("oc_xlafun2", "oc_lafun2", "oc_xlafun", "oc_ylafun", "oc_zlafun", "oc_lafun"):
            getattr(L, f).restype = C.c_double
            getattr(L, f).argtypes = [C.c_double] * 3
        L.oc_igf.restype = C.c_double
        L.oc_igf.argtypes = [C.c_int, C.c_int] + [C.c_double] * 7
        L.oc_rfun.restype = C.c_double
        L.oc_rfun.argtypes = [C.c_double] * 7 + [C.c_int, C.c_int]
        L.oc_set_perturb.argtypes = [C.c_uint64]
        L.oc_set_zero_unused_planes.argtypes = [C.c_int]

    # -- test knobs ----------------------------------------------------------
    def set_perturb(self, seed: int) -> None:
        """Perturb every IGF primitive result by -1/0/+1 ulp (seed 0 = off)."""
        self.lib.oc_set_perturb(C.c_uint64(seed))

    def set_zero_unused_planes(self, on: bool) -> None:
        self.lib.oc_set_zero_unused_planes(int(bool(on)))

    def floor(self, fn, seeds=(12345, 67890)):
        """Conditioning floor of fn(): max spread of its (array) result when
        the IGF primitives are perturbed by 1 ulp, over a few seeds.  Returns
        (unperturbed result, floor as a fraction of max|result| per component)."""
        base = fn()
        worst = 0.0
        for s in seeds:
            self.set_perturb(s)
            try:
                pert = fn()
            finally:
                self.set_perturb(0)
            worst = max(worst, _relcomp(pert, base))
        return base, worst

    # -- FFT ---------------------------------------------------------------
    def ccfft3d(self, a, idir):
        a = np.require(np.asarray(a, dtype=np.complex128), requirements=["F", "A", "W", "O"]).copy(order="F")
        n1, n2, n3 = a.shape
        rc = self.lib.oc_ccfft3d(a.ctypes.data_as(_dp), n1, n2, n3, _i3(idir))
        if rc:
            raise ValueError("FFT length has a prime factor > 23")
        return a

    # -- deposition ----------------------------------------------------------
    def deposit_particles(self, x, y, z, nlo, nhi, q=None, total_charge=None, resize=True,
                          mn=(0, 0, 0), mx=(0, 0, 0), delta=(1, 1, 1)):
        x, y, z = _f64(x), _f64(y), _f64(z)
        mn_, mx_, dl_ = _d3(mn), _d3(mx), _d3(delta)
        charge = C.c_double(0)
        rho = np.zeros(nsize(nlo, nhi), order="F")
        nd = C.c_int64(0)
        qa = _f64(q) if q is not None else None
        rc = self.lib.oc_deposit_particles(
            C.c_int64(x.size), _d(x), _d(y), _d(z), _d(qa), int(total_charge is not None),
            C.c_double(total_charge or 0.0), int(bool(resize)), _i3(nlo), _i3(nhi), mn_, mx_, dl_,
            C.byref(charge), _d(rho), C.byref(nd))
        return dict(rc=rc, rho=rho, charge=charge.value, min=np.array(mn_[:]), max=np.array(mx_[:]),
                    delta=np.array(dl_[:]), dropped=nd.value)

    def get_mesh_quantities(self, x, y, z, nlo, nhi, isetzmin=0):
        x, y, z = _f64(x), _f64(y), _f64(z)
        dl, um = (C.c_double * 3)(), (C.c_double * 3)()
        self.lib.oc_get_mesh_quantities(C.c_int64(x.size), _d(x), _d(y), _d(z), _i3(nlo), _i3(nhi),
                                        int(isetzmin), dl, um)
        return np.array(dl[:]), np.array(um[:])

    def depose_rho_scalar(self, x, y, z, nlo, nhi, chrgpermacro, delta, umin):
        x, y, z = _f64(x), _f64(y), _f64(z)
        rho = np.zeros(nsize(nlo, nhi), order="F")
        nf = self.lib.oc_depose_rho_scalar(C.c_int64(x.size), _d(x), _d(y), _d(z), _i3(nlo), _i3(nhi),
                                           C.c_double(chrgpermacro), _d3(delta), _d3(umin), _d(rho))
        return rho, int(nf)

    # -- Green functions -------------------------------------------------------
    def get_cgrn_freespace(self, shape, delta, gamma, icomp, offset=None):
        g = np.zeros(shape, dtype=np.complex128, order="F")
        self.lib.oc_get_cgrn_freespace(g.ctypes.data_as(_dp), *[int(s) for s in shape], _d3(delta),
                                       C.c_double(gamma), int(icomp),
                                       _d3(offset) if offset is not None else None)
        return g

    @staticmethod
    def legacy_period(nlo, nhi, npad):
        return tuple(2 * (h - l) + p + 1 for l, h, p in zip(nlo, nhi, npad))

    def getgrnfree(self, nlo, nhi, npad, delta, gamma, icomp, igfflag=1):
        g = np.zeros(self.legacy_period(nlo, nhi, npad), dtype=np.complex128, order="F")
        self.lib.oc_getgrnfree(g.ctypes.data_as(_dp), _i3(nlo), _i3(nhi), _i3(npad), _d3(delta),
                               C.c_double(gamma), int(icomp), int(igfflag))
        return g

    def getgrnimageshift(self, nlo, nhi, npad, delta, gamma, zshft, icomp, igfflag=1):
        g = np.zeros(self.legacy_period(nlo, nhi, npad), dtype=np.complex128, order="F")
        self.lib.oc_getgrnimageshift(g.ctypes.data_as(_dp), _i3(nlo), _i3(nhi), _i3(npad), _d3(delta),
                                     C.c_double(gamma), C.c_double(zshft), int(icomp), int(igfflag))
        return g

    def getgrnimageconvcorr(self, nlo, nhi, npad, umin, delta, gamma, icomp, igfflag=1):
        g = np.zeros(self.legacy_period(nlo, nhi, npad), dtype=np.complex128, order="F")
        self.lib.oc_getgrnimageconvcorr(g.ctypes.data_as(_dp), _i3(nlo), _i3(nhi), _i3(npad),
                                        _d3(umin), _d3(delta), C.c_double(gamma), int(icomp),
                                        int(igfflag))
        return g

    def getgrnpipe(self, nlo, nhi, npad, gamma, a, b, delta, umin):
        per = self.legacy_period(nlo, nhi, npad)
        gs = [np.zeros(per, dtype=np.complex128, order="F") for _ in range(4)]
        arr = (_dp * 4)(*[g.ctypes.data_as(_dp) for g in gs])
        self.lib.oc_getgrnpipe(arr, _i3(nlo), _i3(nhi), _i3(npad), C.c_double(gamma), C.c_double(a),
                               C.c_double(b), _d3(delta), _d3(umin))
        return gs

    # -- solvers ---------------------------------------------------------------
    def freespace_solver2(self, rho, gamma, delta, offset=None, want_phi=True, want_e=True):
        rho = _f64(rho)
        nx, ny, nz = rho.shape
        e = np.zeros((nx, ny, nz, 3), order="F") if want_e else None
        phi = np.zeros((nx, ny, nz), order="F") if want_phi else None
        self.lib.oc_freespace_solver2(nx, ny, nz, _d(rho), C.c_double(gamma), _d3(delta), _d(e),
                                      _d(phi), _d3(offset) if offset is not None else None)
        return e, phi

    def space_charge_3d(self, rho, nlo, nhi, mn, mx, delta, gamma, offset=(0, 0, 0),
                        at_cathode=False, calc_bfield=True):
        rho = _f64(rho)
        sh = rho.shape + (3,)
        e, b = np.zeros(sh, order="F"), np.zeros(sh, order="F")
        self.lib.oc_space_charge_3d(_i3(nlo), _i3(nhi), _d3(mn), _d3(mx), _d3(delta), C.c_double(gamma),
                                    _d(rho), _d3(offset), int(bool(at_cathode)),
                                    int(bool(calc_bfield)), _d(e), _d(b))
        return e, b

    def space_charge_freespace(self, rho, nlo, nhi, npad, delta, gamma, direct=True, igf=True):
        rho = _f64(rho)
        phi = np.zeros(rho.shape, order="F")
        e, b = np.zeros(rho.shape + (3,), order="F"), np.zeros(rho.shape + (3,), order="F")
        rc = self.lib.oc_space_charge_freespace(_i3(nlo), _i3(nhi), _i3(npad), _d3(delta),
                                                C.c_double(gamma), _d(rho), int(bool(direct)),
                                                int(bool(igf)), _d(phi), _d(e), _d(b))
        return rc, phi, e, b

    def space_charge_cathodeimages(self, rho, nlo, nhi, npad, umin, delta, gamma, direct=True,
                                   igf=True, method=1):
        rho = _f64(rho).copy(order="F")
        phi = np.zeros(rho.shape, order="F")
        e, b = np.zeros(rho.shape + (3,), order="F"), np.zeros(rho.shape + (3,), order="F")
        rc = self.lib.oc_space_charge_cathodeimages(_i3(nlo), _i3(nhi), _i3(npad), _d3(umin),
                                                    _d3(delta), C.c_double(gamma), _d(rho),
                                                    int(bool(direct)), int(bool(igf)), int(method),
                                                    _d(phi), _d(e), _d(b))
        return rc, phi, e, b, rho

    def space_charge_rectpipe(self, rho, nlo, nhi, npad, delta, gamma, greens, direct=False,
                              efield=None):
        rho = _f64(rho)
        phi = np.zeros(rho.shape, order="F")
        e = np.zeros(rho.shape + (3,), order="F") if efield is None else _f64(efield).copy(order="F")
        b = np.zeros(rho.shape + (3,), order="F")
        arr = (_dp * 4)(*[np.asarray(g).ctypes.data_as(_dp) for g in greens])
        rc = self.lib.oc_space_charge_rectpipe(_i3(nlo), _i3(nhi), _i3(npad), _d3(delta),
                                               C.c_double(gamma), _d(rho), arr, int(bool(direct)),
                                               _d(phi), _d(e), _d(b))
        return rc, phi, e, b

    def write_rectpipe_grn(self, path, a, b, delta, umin, umax, nlo, nhi, npad, gamma, greens):
        per = self.legacy_period(nlo, nhi, npad)
        gs = [np.require(g, dtype=np.complex128, requirements=["F"]) for g in greens]
        arr = (_dp * 4)(*[g.ctypes.data_as(_dp) for g in gs])
        return self.lib.oc_write_rectpipe_grn(os.fsencode(path), C.c_double(a), C.c_double(b),
                                              _d3(delta), _d3(umin), _d3(umax), _i3(nlo), _i3(nhi),
                                              _i3(per), C.c_double(gamma), arr)

    def read_rectpipe_grn(self, path, nlo, nhi, npad):
        per = self.legacy_period(nlo, nhi, npad)
        gs = [np.zeros(per, dtype=np.complex128, order="F") for _ in range(4)]
        arr = (_dp * 4)(*[g.ctypes.data_as(_dp) for g in gs])
        hr, hi = (C.c_double * 12)(), (C.c_int * 9)()
        rc = self.lib.oc_read_rectpipe_grn(os.fsencode(path), hr, hi, arr, C.c_int64(int(np.prod(per))))
        return rc, np.array(hr[:]), np.array(hi[:]), gs

    # -- gather ------------------------------------------------------------------
    def interpolate_field(self, nlo, nhi, mn, delta, efield, bfield, x, y, z):
        x, y, z = _f64(x), _f64(y), _f64(z)
        ef, bf = _f64(efield), _f64(bfield) if bfield is not None else None
        n = x.size
        E = np.zeros((3, n), order="F")
        B = np.zeros((3, n), order="F") if bf is not None else None
        cl = np.zeros(n, dtype=np.int8)
        self.lib.oc_interpolate_field(_i3(nlo), _i3(nhi), _d3(mn), _d3(delta), _d(ef), _d(bf),
                                      C.c_int64(n), _d(x), _d(y), _d(z), _d(E), _d(B),
                                      cl.ctypes.data_as(C.POINTER(C.c_int8)))
        return E, B, cl


class Reference:
    """The upstream Fortran library (oracle/_ref/libopensc_ref.so)."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle ref)")
        self.lib = C.CDLL(path)

    def ccfft3d(self, a, idir):
        a = np.asarray(a, dtype=np.complex128).copy(order="F")
        n1, n2, n3 = a.shape
        self.lib.ref_ccfft3d(n1, n2, n3, a.ctypes.data_as(_dp), _i3(idir))
        return a

    def deposit_particles(self, x, y, z, nlo, nhi, q=None, total_charge=None, resize=True,
                          mn=(0, 0, 0), mx=(0, 0, 0), delta=(1, 1, 1), npad=(1, 1, 1)):
        x, y, z = _f64(x), _f64(y), _f64(z)
        mn_, mx_, dl_ = _d3(mn), _d3(mx), _d3(delta)
        charge = C.c_double(0)
        rho = np.zeros(nsize(nlo, nhi), order="F")
        qmode = 1 if q is not None else (2 if total_charge is not None else 0)
        qa = _f64(q) if q is not None else np.zeros(x.size)
        self.lib.ref_deposit_particles(C.c_int64(x.size), _d(x), _d(y), _d(z), _d(qa), int(qmode),
                                       C.c_double(total_charge or 0.0), int(bool(resize)), _i3(nlo),
                                       _i3(nhi), _i3(npad), mn_, mx_, dl_, C.c_double(1.0),
                                       C.byref(charge), _d(rho))
        return dict(rho=rho, charge=charge.value, min=np.array(mn_[:]), max=np.array(mx_[:]),
                    delta=np.array(dl_[:]))

    def get_mesh_quantities(self, x, y, z, nlo, nhi, isetzmin=0):
        x, y, z = _f64(x), _f64(y), _f64(z)
        dl, um = (C.c_double * 3)(), (C.c_double * 3)()
        self.lib.ref_get_mesh_quantities(int(x.size), _d(x), _d(y), _d(z), _i3(nlo), _i3(nhi),
                                         int(isetzmin), dl, um)
        return np.array(dl[:]), np.array(um[:])

    def depose_rho_scalar(self, x, y, z, nlo, nhi, chrgpermacro, delta, umin):
        x, y, z = _f64(x), _f64(y), _f64(z)
        rho = np.zeros(nsize(nlo, nhi), order="F")
        nf = C.c_int(0)
        self.lib.ref_depose_rho_scalar(int(x.size), _d(x), _d(y), _d(z), _i3(nlo), _i3(nhi),
                                       C.c_double(chrgpermacro), _d3(delta), _d3(umin), _d(rho),
                                       C.byref(nf))
        return rho, nf.value

    def get_cgrn_freespace(self, shape, delta, gamma, icomp, offset=(0, 0, 0)):
        g = np.zeros(shape, dtype=np.complex128, order="F")
        self.lib.ref_get_cgrn_freespace(*[int(s) for s in shape], g.ctypes.data_as(_dp), _d3(delta),
                                        C.c_double(gamma), int(icomp), _d3(offset))
        return g

    def getgrnfree(self, nlo, nhi, npad, delta, gamma, icomp, igfflag=1):
        g = np.zeros(Oracle.legacy_period(nlo, nhi, npad), dtype=np.complex128, order="F")
        self.lib.ref_getgrnfree(_i3(nlo), _i3(nhi), _i3(npad), g.ctypes.data_as(_dp), _d3(delta),
                                C.c_double(gamma), int(icomp), int(igfflag))
        return g

    def getgrnimageshift(self, nlo, nhi, npad, delta, gamma, zshft, icomp, igfflag=1):
        g = np.zeros(Oracle.legacy_period(nlo, nhi, npad), dtype=np.complex128, order="F")
        self.lib.ref_getgrnimageshift(_i3(nlo), _i3(nhi), _i3(npad), g.ctypes.data_as(_dp),
                                      _d3(delta), C.c_double(gamma), C.c_double(zshft), int(icomp),
                                      int(igfflag))
        return g

    def getgrnimageconvcorr(self, nlo, nhi, npad, umin, delta, gamma, icomp, igfflag=1):
        g = np.zeros(Oracle.legacy_period(nlo, nhi, npad), dtype=np.complex128, order="F")
        self.lib.ref_getgrnimageconvcorr(_i3(nlo), _i3(nhi), _i3(npad), g.ctypes.data_as(_dp),
                                         _d3(umin), _d3(delta), C.c_double(gamma), int(icomp),
                                         int(igfflag))
        return g

    def freespace_solver2(self, rho, gamma, delta, offset=(0, 0, 0)):
        rho = _f64(rho)
        nx, ny, nz = rho.shape
        e = np.zeros((nx, ny, nz, 3), order="F")
        phi = np.zeros((nx, ny, nz), order="F")
        self.lib.ref_freespace_solver2(nx, ny, nz, _d(rho), C.c_double(gamma), _d3(delta), _d3(offset),
                                       _d(e), _d(phi))
        return e, phi

    def space_charge_3d(self, rho, nlo, nhi, mn, mx, delta, gamma, offset=(0, 0, 0),
                        at_cathode=False, calc_bfield=True, npad=(1, 1, 1)):
        rho = _f64(rho)
        sh = rho.shape + (3,)
        e, b = np.zeros(sh, order="F"), np.zeros(sh, order="F")
        self.lib.ref_space_charge_3d(_i3(nlo), _i3(nhi), _i3(npad), _d3(mn), _d3(mx), _d3(delta),
                                     C.c_double(gamma), _d(rho), _d3(offset), int(bool(at_cathode)),
                                     int(bool(calc_bfield)), _d(e), _d(b))
        return e, b

    def _mesh_call(self, fn, rho, nlo, nhi, npad, mn, mx, delta, gamma, *flags, extra=()):
        rho = _f64(rho).copy(order="F")
        phi = np.zeros(rho.shape, order="F")
        e, b = np.zeros(rho.shape + (3,), order="F"), np.zeros(rho.shape + (3,), order="F")
        fn(_i3(nlo), _i3(nhi), _i3(npad), _d3(mn), _d3(mx), _d3(delta), C.c_double(gamma), _d(rho),
           *extra, *[int(f) for f in flags], _d(phi), _d(e), _d(b))
        return phi, e, b, rho

    def space_charge_freespace(self, rho, nlo, nhi, npad, mn, mx, delta, gamma, direct=True, igf=True):
        phi, e, b, _ = self._mesh_call(self.lib.ref_space_charge_freespace, rho, nlo, nhi, npad, mn,
                                       mx, delta, gamma, direct, igf)
        return phi, e, b

    def space_charge_cathodeimages(self, rho, nlo, nhi, npad, mn, mx, delta, gamma, direct=True,
                                   igf=True, method=1):
        return self._mesh_call(self.lib.ref_space_charge_cathodeimages, rho, nlo, nhi, npad, mn, mx,
                               delta, gamma, direct, igf, method)

    def pipe_green(self, nlo, nhi, npad, gamma, a, b, delta, umin, umax, write_file=True):
        """Builds the upstream pipe Green arrays (module globals) and writes
        ./rectpipegrn.dat when write_file. Module globals are never freed
        upstream, so one process can only build one pipe grid size."""
        self.lib.ref_pipe_green(_i3(nlo), _i3(nhi), _i3(npad), C.c_double(gamma), C.c_double(a),
                                C.c_double(b), _d3(delta), _d3(umin), _d3(umax), int(bool(write_file)))

    def pipe_read(self, nlo, nhi, npad):
        self.lib.ref_pipe_read(_i3(nlo), _i3(nhi), _i3(npad))

    def space_charge_rectpipe(self, rho, nlo, nhi, npad, mn, mx, delta, gamma, a, b, direct=False,
                              igf=True):
        phi, e, bb, _ = self._mesh_call(self.lib.ref_space_charge_rectpipe, rho, nlo, nhi, npad, mn,
                                        mx, delta, gamma, direct, igf,
                                        extra=(C.c_double(a), C.c_double(b)))
        return phi, e, bb

    def interpolate_field(self, nlo, nhi, mn, mx, delta, efield, bfield, x, y, z):
        x, y, z = _f64(x), _f64(y), _f64(z)
        n = x.size
        E, B = np.zeros((3, n), order="F"), np.zeros((3, n), order="F")
        ef = _f64(efield)
        bf = _f64(bfield) if bfield is not None else np.zeros_like(ef)
        self.lib.ref_interpolate_field(_i3(nlo), _i3(nhi), _d3(mn), _d3(mx), _d3(delta), _d(ef), _d(bf),
                                       C.c_int64(n), _d(x), _d(y), _d(z), _d(E), _d(B))
        return E, B


def gaussian_bunch(n, sigma, seed, cutoff=4.0, center=(0.0, 0.0, 0.0)):
    """Seeded truncated Gaussian (rejection at r > cutoff in sigma units, the
    rule of test_mod.f90:59), FP64. Deterministic for a given (n, seed)."""
    rng = np.random.Generator(np.random.Philox(seed))
    out = np.empty((0, 3))
    while out.shape[0] < n:
        r = rng.standard_normal((max(n - out.shape[0], 16) * 2, 3))
        r = r[(r * r).sum(1) <= cutoff * cutoff]
        out = np.vstack([out, r])
    out = out[:n] * np.asarray(sigma, dtype=np.float64) + np.asarray(center, dtype=np.float64)
    return np.ascontiguousarray(out[:, 0]), np.ascontiguousarray(out[:, 1]), np.ascontiguousarray(out[:, 2])


def uniform_ellipsoid(n, semi_axes, seed, center=(0.0, 0.0, 0.0)):
    """Seeded uniform-density ellipsoid by rejection from the unit cube
    (test_mod.f90:42-48 rule)."""
    rng = np.random.Generator(np.random.Philox(seed))
    out = np.empty((0, 3))
    while out.shape[0] < n:
        r = rng.uniform(-1.0, 1.0, (max(n - out.shape[0], 16) * 2, 3))
        r = r[(r * r).sum(1) <= 1.0]
        out = np.vstack([out, r])
    out = out[:n] * np.asarray(semi_axes, dtype=np.float64) + np.asarray(center, dtype=np.float64)
    return np.ascontiguousarray(out[:, 0]), np.ascontiguousarray(out[:, 1]), np.ascontiguousarray(out[:, 2])
