"""ctypes binding of the C-ABI in include/opensc_hip.h (libopensc_hip.so).

The library is loaded from the package directory; importing this module
never falls back to anything else: if the library is missing every call
raises :class:`LibraryNotBuilt`.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libopensc_hip.so")

OSC_OK = 0
OSC_ERR_ARG = 1
OSC_ERR_NO_CHARGE = 2
OSC_ERR_FFT_LENGTH = 3
OSC_ERR_HIP = 4
OSC_ERR_NOMEM = 5
OSC_ERR_DIRECT_NEEDS_IGF = 6
OSC_ERR_PIPE_DIRECT = 7
OSC_ERR_NO_GREEN = 8
OSC_ERR_IO = 9
OSC_ERR_NO_DEVICE = 10

OSC_PREC_FP64 = 0
OSC_PREC_FP32 = 1
# most mesh nodes osc_interpolate_field_noalloc's fixed scratch holds
OSC_NOALLOC_MAX_NODES = 1048576
OSC_SERVER_NSTATS = 12


class LibraryNotBuilt(RuntimeError):
    pass


class OSCError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[osc error {code}] {msg}")
        self.code = code
        self.msg = msg


OSC_SOLVE_FREE = 0
OSC_SOLVE_CATHODE = 1
OSC_SOLVE_PIPE = 2


class OscRequest(C.Structure):
    """osc_request (one solve of the request loop)."""
    _fields_ = [
        ("solver", C.c_int32),
        ("precision", C.c_int32),
        ("nhi", C.c_int32 * 3),
        ("resize_mesh", C.c_int32),
        ("n", C.c_int64),
        ("x", C.c_void_p),
        ("y", C.c_void_p),
        ("z", C.c_void_p),
        ("total_charge", C.c_double),
        ("gamma", C.c_double),
        ("min", C.c_double * 3),
        ("max", C.c_double * 3),
        ("apipe", C.c_double),
        ("bpipe", C.c_double),
        ("E", C.c_void_p),
        ("B", C.c_void_p),
    ]


class OscMesh(C.Structure):
    """osc_mesh (mesh3d_struct without arrays)."""
    _fields_ = [
        ("nlo", C.c_int32 * 3),
        ("nhi", C.c_int32 * 3),
        ("npad", C.c_int32 * 3),
        ("min", C.c_double * 3),
        ("max", C.c_double * 3),
        ("delta", C.c_double * 3),
        ("gamma", C.c_double),
        ("charge", C.c_double),
    ]


_vp = C.c_void_p
_i64 = C.c_int64
_i32p = C.POINTER(C.c_int32)
_dp = C.POINTER(C.c_double)
_ctxp = C.c_void_p
_meshp = C.POINTER(OscMesh)

# name -> (restype, argtypes); every symbol declared in include/opensc_hip.h
SIGNATURES = {
    "osc_abi_version": (C.c_int, []),
    "osc_fft_length_supported": (C.c_int, [C.c_int]),
    "osc_fft_length_planned": (C.c_int, [C.c_int]),
    "osc_fft_tuning": (C.c_int, [C.c_int]),
    "osc_fft_tune_stats": (C.c_int, [C.POINTER(_i64)]),
    "osc_ctx_create": (C.c_int, [C.c_int, C.POINTER(_ctxp)]),
    "osc_ctx_destroy": (C.c_int, [_ctxp]),
    "osc_ctx_set_stream": (C.c_int, [_ctxp, _vp]),
    "osc_ctx_set_async": (C.c_int, [_ctxp, C.c_int]),
    "osc_ctx_set_precision": (C.c_int, [_ctxp, C.c_int]),
    "osc_ctx_set_green_cache": (C.c_int, [_ctxp, C.c_int, _i64]),
    "osc_ctx_set_graphs": (C.c_int, [_ctxp, C.c_int]),
    "osc_ctx_graph_stats": (C.c_int, [_ctxp, C.POINTER(_i64)]),
    "osc_ctx_clear_caches": (C.c_int, [_ctxp]),
    "osc_ctx_synchronize": (C.c_int, [_ctxp]),
    "osc_ctx_last_error": (C.c_char_p, [_ctxp]),
    "osc_ctx_set_profiling": (C.c_int, [_ctxp, C.c_int]),
    "osc_ctx_profile_read": (C.c_int, [_ctxp, _dp, C.POINTER(_i64)]),
    "osc_stage_name": (C.c_char_p, [C.c_int]),
    "osc_deposit_particles": (C.c_int, [_ctxp, _i64, _vp, _vp, _vp, _vp, _dp, C.c_int, _meshp,
                                        _vp, C.POINTER(_i64)]),
    "osc_space_charge_step": (C.c_int, [_ctxp, _i64, _vp, _vp, _vp, _vp, _dp, C.c_int, _dp, C.c_int,
                                        C.c_int, _meshp, _vp, _vp, _vp, _vp, _vp, C.POINTER(_i64),
                                        C.POINTER(_i64)]),
    "osc_get_mesh_quantities": (C.c_int, [_ctxp, _i64, _vp, _vp, _vp, _i32p, _i32p, C.c_int, _dp,
                                          _dp]),
    "osc_depose_rho_scalar": (C.c_int, [_ctxp, _i64, _vp, _vp, _vp, _i32p, _i32p, C.c_double, _dp,
                                        _dp, _vp, C.POINTER(_i64)]),
    "osc_space_charge_3d": (C.c_int, [_ctxp, _meshp, _vp, _dp, C.c_int, C.c_int, _vp, _vp]),
    "osc_freespace_solver2": (C.c_int, [_ctxp, C.c_int, C.c_int, C.c_int, _vp, C.c_double, _dp, _vp,
                                        _vp, _dp]),
    "osc_get_cgrn_freespace": (C.c_int, [_ctxp, _vp, C.c_int, C.c_int, C.c_int, _dp, C.c_double,
                                         C.c_int, _dp]),
    "osc_ccfft3d": (C.c_int, [_ctxp, _vp, _vp, _i32p, C.c_int, C.c_int, C.c_int]),
    "osc_space_charge_freespace": (C.c_int, [_ctxp, _meshp, _vp, C.c_int, C.c_int, _vp, _vp, _vp]),
    "osc_space_charge_cathodeimages": (C.c_int, [_ctxp, _meshp, _vp, C.c_int, C.c_int, C.c_int,
                                                 _vp, _vp, _vp]),
    "osc_getgrnfree": (C.c_int, [_ctxp, _vp, _i32p, _i32p, _i32p, _dp, C.c_double, C.c_int,
                                 C.c_int]),
    "osc_getgrnimageshift": (C.c_int, [_ctxp, _vp, _i32p, _i32p, _i32p, _dp, C.c_double,
                                       C.c_double, C.c_int, C.c_int]),
    "osc_getgrnimageconvcorr": (C.c_int, [_ctxp, _vp, _i32p, _i32p, _i32p, _dp, _dp, C.c_double,
                                          C.c_int, C.c_int]),
    "osc_alloc_rectpipe_arrays": (C.c_int, [_ctxp, _i32p, _i32p, _i32p]),
    "osc_getgrnpipe": (C.c_int, [_ctxp, C.c_double, C.c_double, C.c_double, _dp, _dp, _i32p]),
    "osc_write_rectpipe_grn": (C.c_int, [_ctxp, C.c_char_p, C.c_double, C.c_double, _dp, _dp, _dp,
                                         _i32p, _i32p, C.c_double]),
    "osc_read_rectpipe_grn": (C.c_int, [_ctxp, C.c_char_p, _dp, _i32p]),
    "osc_space_charge_rectpipe": (C.c_int, [_ctxp, _meshp, C.c_double, C.c_double, C.c_int,
                                            C.c_int, _vp, _vp, _vp, _vp]),
    "osc_interpolate_field": (C.c_int, [_ctxp, _meshp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp,
                                        _vp, C.POINTER(_i64)]),
    "osc_interpolate_field_noalloc": (C.c_int, [_ctxp, _meshp, _vp, _vp, _i64, _vp, _vp, _vp, _vp,
                                                _vp, _vp, C.POINTER(_i64)]),
    "osc_server_create": (C.c_int, [C.c_int, C.c_int, _i64, C.c_int, C.POINTER(_vp)]),
    "osc_server_submit": (C.c_int, [_vp, C.c_void_p, C.POINTER(_i64)]),
    "osc_server_wait": (C.c_int, [_vp, _i64, _dp, C.POINTER(_i64)]),
    "osc_server_drain": (C.c_int, [_vp]),
    "osc_server_clear_caches": (C.c_int, [_vp]),
    "osc_server_stats": (C.c_int, [_vp, C.POINTER(_i64), C.c_int]),
    "osc_server_last_error": (C.c_char_p, [_vp]),
    "osc_server_destroy": (C.c_int, [_vp]),
}

_lib = None
_lock = threading.Lock()


def load(path: str = LIB_PATH) -> C.CDLL:
    """Loads libopensc_hip.so (once) and declares every entry point."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise LibraryNotBuilt(
                f"{path} is not built; run `python -m openspacecharge_amd.build` (or "
                "__graft_entry__.build()). There is no CPU fallback.")
        # torch ships its own HIP and HSA runtime libraries under the same
        # sonames as ROCm's; the first loaded serves the whole process, and
        # torch cannot initialise the GPU on a runtime it was not built with.
        # Loading torch first keeps one runtime for both (and lets torch
        # tensors be passed as device arrays).
        try:
            import torch  # noqa: F401
        except Exception:
            pass
        lib = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
        return lib


def i3(v) -> C.Array:
    return (C.c_int32 * 3)(*[int(t) for t in v])


def d3(v) -> C.Array:
    return (C.c_double * 3)(*[float(t) for t in v])


class Context:
    """One osc_ctx: a GPU, a stream, pooled buffers and Green-function caches."""

    def __init__(self, device: int = 0, precision: int = OSC_PREC_FP64):
        self.lib = load()
        h = _ctxp()
        rc = self.lib.osc_ctx_create(int(device), C.byref(h))
        if rc != OSC_OK:
            raise OSCError(rc, "osc_ctx_create failed (no usable GPU?)")
        self.h = h
        self.device = device
        if precision != OSC_PREC_FP64:
            self.set_precision(precision)

    def check(self, rc: int) -> int:
        if rc != OSC_OK:
            raise OSCError(rc, self.last_error())
        return rc

    def last_error(self) -> str:
        s = self.lib.osc_ctx_last_error(self.h)
        return s.decode() if s else ""

    def set_precision(self, prec: int):
        self.check(self.lib.osc_ctx_set_precision(self.h, int(prec)))

    def set_stream(self, stream_handle: int | None):
        self.check(self.lib.osc_ctx_set_stream(self.h, _vp(stream_handle or 0)))

    def set_async(self, on: bool):
        self.check(self.lib.osc_ctx_set_async(self.h, int(bool(on))))

    def set_green_cache(self, enable: bool, max_bytes: int = 0):
        self.check(self.lib.osc_ctx_set_green_cache(self.h, int(bool(enable)), int(max_bytes)))

    def clear_caches(self):
        self.check(self.lib.osc_ctx_clear_caches(self.h))

    def set_graphs(self, enable: bool):
        """Captured-solve replay of fused steps (osc_ctx_set_graphs)."""
        self.check(self.lib.osc_ctx_set_graphs(self.h, int(bool(enable))))

    def graph_stats(self) -> dict:
        v = (C.c_int64 * 3)()
        self.check(self.lib.osc_ctx_graph_stats(self.h, v))
        return dict(held=v[0], captured=v[1], replays=v[2])

    def synchronize(self):
        self.check(self.lib.osc_ctx_synchronize(self.h))

    NSTAGE = 20

    def set_profiling(self, on: bool):
        self.check(self.lib.osc_ctx_set_profiling(self.h, int(bool(on))))

    def profile_read(self) -> dict:
        """{stage name: (total ms, launches)} since the last read (synchronises)."""
        ms = (C.c_double * self.NSTAGE)()
        n = (C.c_int64 * self.NSTAGE)()
        self.check(self.lib.osc_ctx_profile_read(self.h, ms, n))
        return {self.lib.osc_stage_name(i).decode(): (ms[i], n[i]) for i in range(self.NSTAGE)
                if n[i] > 0}

    def close(self):
        if getattr(self, "h", None):
            self.lib.osc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
