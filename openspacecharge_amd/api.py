"""Python mirror of the upstream Fortran module ``open_spacecharge_mod``
(code/open_spacecharge_mod.f90) and the core routines it re-exports
(code/open_spacecharge_core_mod.f90), running on the HIP library.

Names, argument meaning and defaults follow the upstream routines.  Grids are
Fortran-ordered float64 arrays of shape (nx, ny, nz) (fields (nx, ny, nz, 3)):
numpy arrays in host memory, or torch CUDA tensors with x fastest (strides
(1, nx, nx*ny[, nx*ny*nz])) in device memory.  Upstream error behaviour is
kept: "print + return" cases print the upstream message and return,
"stop" cases raise :class:`OpenSCStop`.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Any, Optional

import numpy as np

from . import _capi as K

_default_ctx: Optional[K.Context] = None


class OpenSCStop(RuntimeError):
    """An upstream ``stop`` (the Fortran code terminates the program)."""


def context(device: int = 0) -> K.Context:
    """The module-level context (the counterpart of the upstream module state)."""
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = K.Context(device)
    return _default_ctx


def set_context(ctx: K.Context) -> None:
    global _default_ctx
    _default_ctx = ctx


# ---------------------------------------------------------------------------
# array plumbing
# ---------------------------------------------------------------------------
def _is_torch(a) -> bool:
    return type(a).__module__.startswith("torch")


def _ptr(a, n_expected: int | None = None, writable: bool = False) -> int | None:
    """Device or host address of a float64 (or int8) array, checking layout."""
    if a is None:
        return None
    if _is_torch(a):
        import torch
        if a.dtype not in (torch.float64, torch.int8, torch.complex128):
            raise TypeError("torch arrays must be float64")
        sh, st = list(a.shape), list(a.stride())
        exp = 1
        for s, t in zip(sh, st):
            if s > 1 and t != exp:
                raise ValueError("torch arrays must be Fortran-contiguous (x fastest)")
            exp *= s
        if n_expected is not None and a.numel() != n_expected:
            raise ValueError(f"array has {a.numel()} elements, expected {n_expected}")
        return a.data_ptr()
    if not isinstance(a, np.ndarray):
        raise TypeError("arrays must be numpy arrays or torch tensors")
    if not (a.flags.f_contiguous or a.ndim <= 1):
        raise ValueError("numpy arrays must be Fortran-contiguous (order='F')")
    if writable and not a.flags.writeable:
        raise ValueError("output array is read-only")
    if n_expected is not None and a.size != n_expected:
        raise ValueError(f"array has {a.size} elements, expected {n_expected}")
    return a.ctypes.data


def _vec(a) -> tuple[Any, int | None]:
    """1-D particle coordinate array -> (keepalive, pointer)."""
    if a is None:
        return None, None
    if _is_torch(a):
        import torch
        a = a.to(torch.float64).contiguous()
        return a, a.data_ptr()
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
    return a, a.ctypes.data


def _zeros_like_layout(ref, shape):
    if ref is not None and _is_torch(ref):
        import torch
        return torch.zeros(tuple(reversed(shape)), dtype=torch.float64,
                           device=ref.device).permute(*reversed(range(len(shape))))
    return np.zeros(shape, dtype=np.float64, order="F")


# ---------------------------------------------------------------------------
# mesh3d_struct (open_spacecharge_mod.f90:14-27)
# ---------------------------------------------------------------------------
@dataclass
class Mesh3D:
    nlo: tuple = (1, 1, 1)
    nhi: tuple = (64, 64, 64)
    npad: tuple = (1, 1, 1)
    min: np.ndarray = field(default_factory=lambda: np.zeros(3))
    max: np.ndarray = field(default_factory=lambda: np.zeros(3))
    delta: np.ndarray = field(default_factory=lambda: np.ones(3))
    gamma: float = 1.0
    charge: float = 0.0
    rho: Any = None
    phi: Any = None
    efield: Any = None
    bfield: Any = None
    device: str = "host"   # "host" (numpy) or "cuda" (torch tensors)

    @property
    def shape(self) -> tuple:
        return tuple(int(h) - int(l) + 1 for l, h in zip(self.nlo, self.nhi))

    @property
    def n(self) -> int:
        s = self.shape
        return s[0] * s[1] * s[2]

    def to_c(self) -> K.OscMesh:
        m = K.OscMesh()
        for i in range(3):
            m.nlo[i], m.nhi[i], m.npad[i] = int(self.nlo[i]), int(self.nhi[i]), int(self.npad[i])
            m.min[i], m.max[i], m.delta[i] = float(self.min[i]), float(self.max[i]), float(self.delta[i])
        m.gamma, m.charge = float(self.gamma), float(self.charge)
        return m

    def from_c(self, m: K.OscMesh) -> None:
        self.min = np.array(m.min[:])
        self.max = np.array(m.max[:])
        self.delta = np.array(m.delta[:])
        self.charge = float(m.charge)

    def allocate(self, like=None) -> None:
        """(Re)allocates rho, phi, efield, bfield when absent or of the wrong
        shape (open_spacecharge_mod.f90:193-219)."""
        sh = self.shape
        if like is None and self.device == "cuda":
            import torch
            like = torch.empty(1, dtype=torch.float64, device="cuda")
        if self.rho is None or tuple(self.rho.shape) != sh:
            self.rho = self.phi = self.efield = self.bfield = None
        if self.rho is None:
            self.rho = _zeros_like_layout(like, sh)
        like = self.rho
        if self.phi is None or tuple(self.phi.shape) != sh:
            self.phi = _zeros_like_layout(like, sh)
        if self.efield is None or tuple(self.efield.shape) != sh + (3,):
            self.efield = _zeros_like_layout(like, sh + (3,))
        if self.bfield is None or tuple(self.bfield.shape) != sh + (3,):
            self.bfield = _zeros_like_layout(like, sh + (3,))


def print_mesh3d(mesh3d: Mesh3D) -> None:
    """print_mesh3d (open_spacecharge_mod.f90:36-52)."""
    print(" ------------------------")
    print(" Mesh: ")
    print(" nlo: ", *mesh3d.nlo)
    print(" nhi: ", *mesh3d.nhi)
    print(" min: ", *mesh3d.min)
    print(" max: ", *mesh3d.max)
    print(" delta: ", *mesh3d.delta)
    print(" gamma: ", mesh3d.gamma)
    print(" charge: ", mesh3d.charge)
    for name in ("rho", "phi", "efield", "bfield"):
        if getattr(mesh3d, name) is not None:
            print(f" {name} allocated")
    print(" ------------------------")


# ---------------------------------------------------------------------------
# deposition
# ---------------------------------------------------------------------------
def deposit_particles(xa, ya, za, mesh3d: Mesh3D, qa=None, total_charge=None, resize_mesh=True,
                      ctx: Optional[K.Context] = None, count_dropped=True):
    """deposit_particles (open_spacecharge_mod.f90:156-288). Returns the number
    of particles skipped because their CIC stencil leaves the mesh; with
    count_dropped=False, None (upstream returns no count), so that a call on
    device arrays in an asynchronous context does not wait for the deposit."""
    ctx = ctx or context()
    kx, px = _vec(xa)
    ky, py = _vec(ya)
    kz, pz = _vec(za)
    kq, pq = _vec(qa)
    n = int(kx.shape[0]) if kx is not None else 0
    mesh3d.allocate(like=kx if _is_torch(kx) else None)
    m = mesh3d.to_c()
    tc = C.c_double(float(total_charge)) if total_charge is not None else None
    nd = C.c_int64(0)
    rc = ctx.lib.osc_deposit_particles(ctx.h, n, px, py, pz, pq, C.byref(tc) if tc is not None else None,
                                       int(bool(resize_mesh)), C.byref(m), _ptr(mesh3d.rho, mesh3d.n),
                                       C.byref(nd) if count_dropped else None)
    mesh3d.from_c(m)
    if rc == K.OSC_ERR_NO_CHARGE:
        print(" Error, must specify qa or total_charge")
        return 0
    ctx.check(rc)
    return int(nd.value) if count_dropped else None


def space_charge_step(xa, ya, za, mesh3d: Mesh3D, qa=None, total_charge=None, resize_mesh=True,
                      offset=None, at_cathode=False, calc_bfield=False, want_B=None,
                      ctx: Optional[K.Context] = None):
    """One tracking step at the same particles: deposit_particles ->
    space_charge_3d -> interpolate_field for every particle, in one call that
    reuses the deposit's cell-sorted order for the gather.  Results equal the
    three separate calls.  Returns E (n,3) (and B (n,3) when calc_bfield)."""
    ctx = ctx or context()
    kx, px = _vec(xa)
    ky, py = _vec(ya)
    kz, pz = _vec(za)
    kq, pq = _vec(qa)
    n = int(kx.shape[0])
    want_B = calc_bfield if want_B is None else want_B
    mesh3d.allocate(like=kx if _is_torch(kx) else None)
    dev = _is_torch(kx)
    if dev:
        import torch
        E = torch.empty((n, 3), dtype=torch.float64, device=kx.device)
        B = torch.empty((n, 3), dtype=torch.float64, device=kx.device) if want_B else None
        pE, pB = E.data_ptr(), (B.data_ptr() if B is not None else None)
    else:
        E = np.empty((n, 3))
        B = np.empty((n, 3)) if want_B else None
        pE, pB = E.ctypes.data, (B.ctypes.data if B is not None else None)
    m = mesh3d.to_c()
    tc = C.c_double(float(total_charge)) if total_charge is not None else None
    nt = mesh3d.n
    rc = ctx.lib.osc_space_charge_step(
        ctx.h, n, px, py, pz, pq, C.byref(tc) if tc is not None else None, int(bool(resize_mesh)),
        K.d3(offset) if offset is not None else None, int(bool(at_cathode)), int(bool(calc_bfield)),
        C.byref(m), _ptr(mesh3d.rho, nt, True), _ptr(mesh3d.efield, 3 * nt, True),
        _ptr(mesh3d.bfield, 3 * nt, True) if calc_bfield else None, pE, pB, None, None)
    mesh3d.from_c(m)
    if rc == K.OSC_ERR_NO_CHARGE:
        print(" Error, must specify qa or total_charge")
        return None
    ctx.check(rc)
    return (E, B) if want_B else E


def get_mesh_quantities(xa, ya, za, nlo, nhi, isetzmin=0, ctx: Optional[K.Context] = None):
    """get_mesh_quantities (code/test_mod.f90:288-358): (delta, umin)."""
    ctx = ctx or context()
    kx, px = _vec(xa)
    ky, py = _vec(ya)
    kz, pz = _vec(za)
    dl, um = (C.c_double * 3)(), (C.c_double * 3)()
    ctx.check(ctx.lib.osc_get_mesh_quantities(ctx.h, int(kx.shape[0]), px, py, pz, K.i3(nlo), K.i3(nhi),
                                              int(isetzmin), dl, um))
    return np.array(dl[:]), np.array(um[:])


def depose_rho_scalar(xa, ya, za, rho, chrgpermacro, nlo, delta, umin,
                      ctx: Optional[K.Context] = None) -> int:
    """depose_rho_scalar (code/test_mod.f90:410-474); fills rho, returns ifail."""
    ctx = ctx or context()
    kx, px = _vec(xa)
    ky, py = _vec(ya)
    kz, pz = _vec(za)
    nhi = [int(nlo[i]) + int(rho.shape[i]) - 1 for i in range(3)]
    nf = C.c_int64(0)
    ctx.check(ctx.lib.osc_depose_rho_scalar(ctx.h, int(kx.shape[0]), px, py, pz, K.i3(nlo), K.i3(nhi),
                                            float(chrgpermacro), K.d3(delta), K.d3(umin),
                                            _ptr(rho, writable=True), C.byref(nf)))
    return int(nf.value)


# ---------------------------------------------------------------------------
# solvers
# ---------------------------------------------------------------------------
def space_charge_3d(mesh3d: Mesh3D, offset=None, at_cathode=False, calc_bfield=False,
                    ctx: Optional[K.Context] = None) -> None:
    """space_charge_3d (open_spacecharge_mod.f90:431-489)."""
    ctx = ctx or context()
    mesh3d.allocate(like=mesh3d.rho)
    m = mesh3d.to_c()
    off = K.d3(offset) if offset is not None else None
    ctx.check(ctx.lib.osc_space_charge_3d(ctx.h, C.byref(m), _ptr(mesh3d.rho, mesh3d.n), off,
                                          int(bool(at_cathode)), int(bool(calc_bfield)),
                                          _ptr(mesh3d.efield, 3 * mesh3d.n, True),
                                          _ptr(mesh3d.bfield, 3 * mesh3d.n, True) if calc_bfield else None))


def osc_freespace_solver2(rho, gamma, delta, efield=None, phi=None, offset=None,
                          ctx: Optional[K.Context] = None) -> None:
    """osc_freespace_solver2 (open_spacecharge_mod.f90:535-597): fills the
    given efield (nx,ny,nz,3) and/or phi (nx,ny,nz) arrays."""
    ctx = ctx or context()
    nx, ny, nz = (int(s) for s in rho.shape)
    n = nx * ny * nz
    ctx.check(ctx.lib.osc_freespace_solver2(ctx.h, nx, ny, nz, _ptr(rho, n), float(gamma), K.d3(delta),
                                            _ptr(efield, 3 * n, True), _ptr(phi, n, True),
                                            K.d3(offset) if offset is not None else None))


def osc_get_cgrn_freespace(cgrn, delta, gamma, icomp, offset=None, ctx: Optional[K.Context] = None):
    """osc_get_cgrn_freespace (open_spacecharge_mod.f90:633-707): fills the
    complex array cgrn (isize,jsize,ksize) with the real-space Green function."""
    ctx = ctx or context()
    isz, jsz, ksz = (int(s) for s in cgrn.shape)
    ctx.check(ctx.lib.osc_get_cgrn_freespace(ctx.h, _ptr(cgrn, writable=True), isz, jsz, ksz, K.d3(delta),
                                             float(gamma), int(icomp),
                                             K.d3(offset) if offset is not None else None))
    return cgrn


def ccfft3d(a, b, idir, ctx: Optional[K.Context] = None):
    """ccfft3d (fft_interface_mod.f90:96-143): b = transform of a (may alias)."""
    ctx = ctx or context()
    n1, n2, n3 = (int(s) for s in a.shape)
    rc = ctx.lib.osc_ccfft3d(ctx.h, _ptr(a), _ptr(b, writable=True), K.i3(idir), n1, n2, n3)
    if rc == K.OSC_ERR_FFT_LENGTH:
        print(" Error return from FFT package due to transform length.")
        print(" Try increasing the padding by 1 in each dimension and re-run")
        raise OpenSCStop(ctx.last_error())
    ctx.check(rc)
    return b


def _legacy_call(fn, ctx, mesh3d, *args):
    rc = fn(ctx.h, *args)
    if rc == K.OSC_ERR_DIRECT_NEEDS_IGF:
        print(" Error: Direct field calc must use integrated Green function")
        print(" Aborting...")
        return rc
    if rc == K.OSC_ERR_FFT_LENGTH:
        print(" Error return from FFT package due to transform length.")
        print(" Try increasing the padding by 1 in each dimension and re-run")
        raise OpenSCStop(ctx.last_error())
    ctx.check(rc)
    return rc


def space_charge_freespace(mesh3d: Mesh3D, direct_field_calc=True, integrated_green_function=True,
                           ctx: Optional[K.Context] = None):
    """space_charge_freespace (open_spacecharge_mod.f90:60-75)."""
    ctx = ctx or context()
    mesh3d.allocate(like=mesh3d.rho)
    m = mesh3d.to_c()
    n = mesh3d.n
    return _legacy_call(ctx.lib.osc_space_charge_freespace, ctx, mesh3d, C.byref(m), _ptr(mesh3d.rho, n),
                        int(bool(direct_field_calc)), int(bool(integrated_green_function)),
                        _ptr(mesh3d.phi, n, True), _ptr(mesh3d.efield, 3 * n, True),
                        _ptr(mesh3d.bfield, 3 * n, True))


def space_charge_cathodeimages(mesh3d: Mesh3D, direct_field_calc=True, integrated_green_function=True,
                               image_method=1, ctx: Optional[K.Context] = None):
    """space_charge_cathodeimages (open_spacecharge_mod.f90:83-103);
    image_method=2 flips mesh3d.rho in z in place, as upstream."""
    ctx = ctx or context()
    mesh3d.allocate(like=mesh3d.rho)
    m = mesh3d.to_c()
    n = mesh3d.n
    return _legacy_call(ctx.lib.osc_space_charge_cathodeimages, ctx, mesh3d, C.byref(m),
                        _ptr(mesh3d.rho, n, True), int(bool(direct_field_calc)),
                        int(bool(integrated_green_function)), int(image_method or 1),
                        _ptr(mesh3d.phi, n, True), _ptr(mesh3d.efield, 3 * n, True),
                        _ptr(mesh3d.bfield, 3 * n, True))


def osc_getgrnfree(nlo, nhi, npad, delta, gamma, icomp, igfflag=1, ctx: Optional[K.Context] = None):
    """osc_getgrnfree (core:454-497): the transformed legacy Green function in
    the upstream convention, complex array of the convolution period."""
    ctx = ctx or context()
    per = tuple(2 * (int(h) - int(l)) + int(p) + 1 for l, h, p in zip(nlo, nhi, npad))
    g = np.zeros(per, dtype=np.complex128, order="F")
    ctx.check(ctx.lib.osc_getgrnfree(ctx.h, g.ctypes.data, K.i3(nlo), K.i3(nhi), K.i3(npad), K.d3(delta),
                                     float(gamma), int(icomp), int(igfflag)))
    return g


def osc_getgrnimageshift(nlo, nhi, npad, delta, gamma, zshft, icomp, igfflag=1,
                         ctx: Optional[K.Context] = None):
    """osc_getgrnimageshift (core:1135-1178)."""
    ctx = ctx or context()
    per = tuple(2 * (int(h) - int(l)) + int(p) + 1 for l, h, p in zip(nlo, nhi, npad))
    g = np.zeros(per, dtype=np.complex128, order="F")
    ctx.check(ctx.lib.osc_getgrnimageshift(ctx.h, g.ctypes.data, K.i3(nlo), K.i3(nhi), K.i3(npad),
                                           K.d3(delta), float(gamma), float(zshft), int(icomp),
                                           int(igfflag)))
    return g


def osc_getgrnimageconvcorr(nlo, nhi, npad, umin, delta, gamma, icomp, igfflag=1,
                            ctx: Optional[K.Context] = None):
    """osc_getgrnimageconvcorr (core:1185-1230)."""
    ctx = ctx or context()
    per = tuple(2 * (int(h) - int(l)) + int(p) + 1 for l, h, p in zip(nlo, nhi, npad))
    g = np.zeros(per, dtype=np.complex128, order="F")
    ctx.check(ctx.lib.osc_getgrnimageconvcorr(ctx.h, g.ctypes.data, K.i3(nlo), K.i3(nhi), K.i3(npad),
                                              K.d3(umin), K.d3(delta), float(gamma), int(icomp),
                                              int(igfflag)))
    return g


# ---- rectangular pipe ---------------------------------------------------------
def osc_alloc_rectpipe_arrays(nlo, nhi, npad, ctx: Optional[K.Context] = None):
    """osc_alloc_rectpipe_arrays (core:876-898)."""
    ctx = ctx or context()
    ctx.check(ctx.lib.osc_alloc_rectpipe_arrays(ctx.h, K.i3(nlo), K.i3(nhi), K.i3(npad)))


def osc_getgrnpipe(gam, a, b, delta, umin, npad, ctx: Optional[K.Context] = None):
    """osc_getgrnpipe (core:632-715)."""
    ctx = ctx or context()
    ctx.check(ctx.lib.osc_getgrnpipe(ctx.h, float(gam), float(a), float(b), K.d3(delta), K.d3(umin),
                                     K.i3(npad)))


def osc_write_rectpipe_grn(apipe, bpipe, delta, umin, umax, nlo, nhi, gamma,
                           path="rectpipegrn.dat", ctx: Optional[K.Context] = None):
    """osc_write_rectpipe_grn (core:851-870)."""
    ctx = ctx or context()
    ctx.check(ctx.lib.osc_write_rectpipe_grn(ctx.h, str(path).encode(), float(apipe), float(bpipe),
                                             K.d3(delta), K.d3(umin), K.d3(umax), K.i3(nlo), K.i3(nhi),
                                             float(gamma)))
    print(" done writing rectpipe transformed Green functions")


def osc_read_rectpipe_grn(path="rectpipegrn.dat", ctx: Optional[K.Context] = None):
    """osc_read_rectpipe_grn (core:825-844); returns the header."""
    ctx = ctx or context()
    hr, hi = (C.c_double * 12)(), (C.c_int32 * 9)()
    ctx.check(ctx.lib.osc_read_rectpipe_grn(ctx.h, str(path).encode(), hr, hi))
    r, i = list(hr), list(hi)
    print(" Done reading rectpipe transformed Green functions with these parameters:")
    print(" apipe,bpipe=", r[0], r[1])
    print(" dx,dy,dz=", *r[2:5])
    print(" xmin,ymin,zmin=", *r[5:8])
    print(" xmax,ymax,zmax=", *r[8:11])
    print(" ilo,jlo,klo=", *i[0:3])
    print(" ihi,jhi,khi=", *i[3:6])
    print(" convolution iperiod,jperiod,kperiod=", *i[6:9])
    print(" gamma=", r[11])
    return dict(apipe=r[0], bpipe=r[1], delta=r[2:5], min=r[5:8], max=r[8:11], nlo=i[0:3],
                nhi=i[3:6], period=i[6:9], gamma=r[11])


def space_charge_rectpipe(mesh3d: Mesh3D, apipe, bpipe, direct_field_calc=True,
                          integrated_green_function=True, ctx: Optional[K.Context] = None):
    """space_charge_rectpipe (open_spacecharge_mod.f90:110-127); needs
    direct_field_calc=False as upstream (core:608-610)."""
    ctx = ctx or context()
    mesh3d.allocate(like=mesh3d.rho)
    m = mesh3d.to_c()
    n = mesh3d.n
    rc = ctx.lib.osc_space_charge_rectpipe(ctx.h, C.byref(m), float(apipe), float(bpipe),
                                           int(bool(direct_field_calc)), int(bool(integrated_green_function)),
                                           _ptr(mesh3d.rho, n), _ptr(mesh3d.phi, n, True),
                                           _ptr(mesh3d.efield, 3 * n, True), _ptr(mesh3d.bfield, 3 * n, True))
    if rc == K.OSC_ERR_PIPE_DIRECT:
        print(" **********ERROR********* rectpipe implementation presently requires directfieldcalc=0")
        return rc
    if rc == K.OSC_ERR_NO_GREEN:
        print(" something wrong. rectpipe green functions should have been computed during initialization")
        raise OpenSCStop(ctx.last_error())
    ctx.check(rc)
    return rc


# ---------------------------------------------------------------------------
# gather
# ---------------------------------------------------------------------------
def interpolate_field(x, y, z, mesh3d: Mesh3D, E=True, B=False, ctx: Optional[K.Context] = None,
                      return_clamped=False):
    """interpolate_field (open_spacecharge_mod.f90:311-390) for one point or
    an array of points.  Returns E (and/or B) of shape (3,) for scalars or
    (n, 3) for arrays (the per-point E(3) of upstream, stacked)."""
    ctx = ctx or context()
    scalar = np.ndim(x) == 0 and not _is_torch(x)
    kx, px = _vec(np.atleast_1d(x) if scalar else x)
    ky, py = _vec(np.atleast_1d(y) if scalar else y)
    kz, pz = _vec(np.atleast_1d(z) if scalar else z)
    n = int(kx.shape[0])
    dev = _is_torch(kx)

    def out(k):
        if dev:
            import torch
            return torch.empty((n, k), dtype=torch.float64 if k else torch.int8, device=kx.device)
        return np.empty((n, k), dtype=np.float64)

    Eo = out(3) if E else None
    Bo = out(3) if B else None
    cl, pcl = None, None
    if return_clamped:
        if dev:
            import torch
            cl = torch.empty(n, dtype=torch.int8, device=kx.device)
            pcl = cl.data_ptr()
        else:
            cl = np.empty(n, dtype=np.int8)
            pcl = cl.ctypes.data
    m = mesh3d.to_c()
    nt = mesh3d.n
    pE = (Eo.data_ptr() if dev else Eo.ctypes.data) if Eo is not None else None
    pB = (Bo.data_ptr() if dev else Bo.ctypes.data) if Bo is not None else None
    ctx.check(ctx.lib.osc_interpolate_field(ctx.h, C.byref(m), _ptr(mesh3d.efield, 3 * nt) if E else None,
                                            _ptr(mesh3d.bfield, 3 * nt) if B else None, n, px, py, pz, pE,
                                            pB, pcl, None))
    if scalar:
        Eo = Eo[0] if Eo is not None else None
        Bo = Bo[0] if Bo is not None else None
    res = tuple(v for v in (Eo, Bo) if v is not None)
    res = res[0] if len(res) == 1 else res
    if return_clamped:
        return res, cl
    return res


# ---------------------------------------------------------------------------
# request loop (osc_server_*): independent solves overlapping on one GPU
# ---------------------------------------------------------------------------
_SOLVERS = {"free": K.OSC_SOLVE_FREE, "cathode": K.OSC_SOLVE_CATHODE, "pipe": K.OSC_SOLVE_PIPE}


class Server:
    """A stream of solves (free space, cathode images, rectangular pipe), each
    deposit -> solve -> gather at its particles.  ``nworkers`` contexts on
    their own HIP streams run requests concurrently; an idle worker starts the
    oldest queued request that fits the memory budget, preferring one whose
    geometry it served recently (its cached Green spectra).
    Particle and output arrays are torch tensors on the GPU and are kept
    alive by the server until ``wait`` returns for the request."""

    def __init__(self, nworkers: int = 4, device: int = 0, max_bytes: int = 0,
                 green_cache: bool = True):
        self.lib = K.load()
        h = C.c_void_p()
        rc = self.lib.osc_server_create(int(device), int(nworkers), int(max_bytes),
                                        int(bool(green_cache)), C.byref(h))
        if rc != K.OSC_OK:
            raise K.OSCError(rc, "osc_server_create failed (no usable GPU?)")
        self.h = h
        self._keep = {}

    def submit(self, x, y, z, E, B=None, *, nhi, gamma, total_charge, solver="free",
               precision=K.OSC_PREC_FP64, mesh_min=None, mesh_max=None, apipe=0.0,
               bpipe=0.0) -> int:
        """Queues one request; E (and B) are (n, 3) float64 tensors receiving
        the fields at the particles.  mesh_min/mesh_max give a fixed mesh
        (required for the pipe); otherwise the mesh follows the bunch."""
        for a in (x, y, z, E) + ((B,) if B is not None else ()):
            if not (_is_torch(a) and a.is_cuda and a.is_contiguous()):
                raise ValueError("server requests take contiguous CUDA tensors")
        r = K.OscRequest()
        r.solver = _SOLVERS[solver]
        r.precision = int(precision)
        for i in range(3):
            r.nhi[i] = int(nhi[i])
        r.resize_mesh = 0 if mesh_min is not None else 1
        r.n = int(x.shape[0])
        r.x, r.y, r.z = x.data_ptr(), y.data_ptr(), z.data_ptr()
        r.total_charge = float(total_charge)
        r.gamma = float(gamma)
        if mesh_min is not None:
            for i in range(3):
                r.min[i], r.max[i] = float(mesh_min[i]), float(mesh_max[i])
        r.apipe, r.bpipe = float(apipe), float(bpipe)
        r.E = E.data_ptr()
        r.B = B.data_ptr() if B is not None else None
        t = C.c_int64(0)
        rc = self.lib.osc_server_submit(self.h, C.byref(r), C.byref(t))
        if rc != K.OSC_OK:
            raise K.OSCError(rc, "osc_server_submit: bad request")
        self._keep[t.value] = (x, y, z, E, B)
        return t.value

    def wait(self, ticket: int) -> dict:
        """Blocks until the request is done; returns its times (ms) and counts."""
        tm = (C.c_double * 3)()
        cn = (C.c_int64 * 2)()
        rc = self.lib.osc_server_wait(self.h, int(ticket), tm, cn)
        self._keep.pop(ticket, None)
        if rc != K.OSC_OK:
            msg = self.lib.osc_server_last_error(self.h)
            raise K.OSCError(rc, msg.decode() if msg else "")
        return dict(queued_ms=tm[0], run_ms=tm[1], total_ms=tm[2], n_dropped=cn[0],
                    n_clamped=cn[1])

    def drain(self) -> None:
        self.lib.osc_server_drain(self.h)

    def clear_caches(self) -> None:
        """Waits until no request is in flight, then returns every worker's
        cached Green spectra and pipe Green functions to its buffer pool
        (osc_server_clear_caches); no device memory is freed."""
        rc = self.lib.osc_server_clear_caches(self.h)
        if rc != K.OSC_OK:
            raise K.OSCError(rc, "osc_server_clear_caches failed")

    STATS = ("done", "max_running", "affine_starts", "bypasses", "green_hits", "green_misses",
             "held_bytes", "budget_bytes", "trims", "trimmed_bytes", "allocs", "alloc_us")

    def stats(self) -> dict:
        """Counters since creation (osc_server_stats)."""
        v = (C.c_int64 * K.OSC_SERVER_NSTATS)()
        rc = self.lib.osc_server_stats(self.h, v, K.OSC_SERVER_NSTATS)
        if rc != K.OSC_OK:
            raise K.OSCError(rc, "osc_server_stats failed")
        return dict(zip(self.STATS, list(v)))

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.osc_server_destroy(self.h)
            self.h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
