"""CPU checks of the C-ABI library (openspacecharge_amd/libopensc_hip.so):
it loads without a GPU, exports every entry point include/opensc_hip.h
declares, the ctypes table and constants agree with the header, the osc_mesh
layout is the C compiler's, and the host-only calls behave.  No compute call
is made here (there is no GPU on the CPU runner); the Python API must fail
loudly, never fall back to a CPU path."""
import ctypes as C
import os
import re
import shutil
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "opensc_hip.h")

from openspacecharge_amd import _capi  # noqa: E402


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_capi.LIB_PATH):
        from openspacecharge_amd.build import build_hip
        build_hip()
    return _capi.load()


def header_text():
    with open(HEADER) as f:
        return f.read()


def header_functions():
    """Names of the functions the header declares (declarations start at column 0)."""
    return set(re.findall(r"^[A-Za-z_][\w \t\*]*?\b(osc_\w+)\s*\(", header_text(), flags=re.M))


def gpu_present():
    return os.path.exists("/dev/kfd")


def test_header_declares_the_surface():
    fns = header_functions()
    # the upstream routines each entry point replaces (open_spacecharge_mod.f90 and
    # open_spacecharge_core_mod.f90) plus the context, profiling and fused-step calls
    for name in ("osc_deposit_particles", "osc_space_charge_3d", "osc_freespace_solver2",
                 "osc_get_cgrn_freespace", "osc_ccfft3d", "osc_space_charge_freespace",
                 "osc_space_charge_cathodeimages", "osc_getgrnfree", "osc_getgrnimageshift",
                 "osc_getgrnimageconvcorr", "osc_alloc_rectpipe_arrays", "osc_getgrnpipe",
                 "osc_write_rectpipe_grn", "osc_read_rectpipe_grn", "osc_space_charge_rectpipe",
                 "osc_interpolate_field", "osc_get_mesh_quantities", "osc_depose_rho_scalar",
                 "osc_space_charge_step", "osc_ctx_create", "osc_ctx_destroy"):
        assert name in fns, name
    assert len(fns) >= 33


def test_library_exports_every_declared_symbol(lib):
    missing = [f for f in sorted(header_functions()) if not hasattr(lib, f)]
    assert not missing, missing
    nm = shutil.which("nm")
    if nm:
        out = subprocess.run([nm, "-D", "--defined-only", _capi.LIB_PATH], capture_output=True,
                             text=True, check=True).stdout
        exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
        assert header_functions() <= exported
        # nothing else is exported under the osc_ prefix
        assert {s for s in exported if s.startswith("osc_")} == header_functions()


def test_ctypes_table_matches_header():
    assert set(_capi.SIGNATURES) == header_functions()


def test_constants_match_header(lib):
    txt = header_text()
    enum = dict((k, int(v)) for k, v in re.findall(r"\b(OSC_(?:OK|ERR_\w+|PREC_\w+))\s*=\s*(\d+)", txt))
    assert enum, "no enum parsed"
    for k, v in enum.items():
        assert getattr(_capi, k) == v, k
    ver = int(re.search(r"#define OSC_ABI_VERSION (\d+)", txt).group(1))
    assert lib.osc_abi_version() == ver
    nst = int(re.search(r"#define OSC_NSTAGE (\d+)", txt).group(1))
    assert _capi.Context.NSTAGE == nst
    mx = int(re.search(r"#define OSC_NOALLOC_MAX_NODES (\d+)", txt).group(1))
    assert _capi.OSC_NOALLOC_MAX_NODES == mx
    ns = int(re.search(r"#define OSC_SERVER_NSTATS (\d+)", txt).group(1))
    assert _capi.OSC_SERVER_NSTATS == ns


def test_mesh_struct_layout(tmp_path):
    """ctypes OscMesh has the size and field offsets the C compiler gives osc_mesh."""
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stddef.h>\n#include <stdio.h>\n#include "opensc_hip.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu %zu %zu\\n\", sizeof(osc_mesh),"
        " offsetof(osc_mesh,nlo), offsetof(osc_mesh,nhi), offsetof(osc_mesh,npad),"
        " offsetof(osc_mesh,min), offsetof(osc_mesh,max), offsetof(osc_mesh,delta),"
        " offsetof(osc_mesh,gamma), offsetof(osc_mesh,charge)); return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)],
                   check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                           check=True).stdout.split()]
    M = _capi.OscMesh
    mine = [C.sizeof(M)] + [getattr(M, f).offset for f in
                            ("nlo", "nhi", "npad", "min", "max", "delta", "gamma", "charge")]
    assert mine == vals


def test_request_struct_layout(tmp_path):
    """ctypes OscRequest matches osc_request."""
    fields = [f for f, _ in _capi.OscRequest._fields_]
    src = tmp_path / "req.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "opensc_hip.h"\n'
                   "int main(void){printf(\"%zu\\n\", sizeof(osc_request));" +
                   "".join(f'printf("%zu\\n", offsetof(osc_request,{f}));' for f in fields) +
                   "return 0;}\n")
    exe = tmp_path / "req"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)],
                   check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                           check=True).stdout.split()]
    R = _capi.OscRequest
    assert [C.sizeof(R)] + [getattr(R, f).offset for f in fields] == vals


def prime_factors(n):
    out, p = set(), 2
    while p * p <= n:
        while n % p == 0:
            out.add(p)
            n //= p
        p += 1
    if n > 1:
        out.add(n)
    return out


def test_fft_length_rule(lib):
    """Every supported length satisfies upstream's rule (prime factors <= 23,
    code/fft_interface_mod.f90:158-162); every doubled grid the workloads use
    (2 x 2^a 3^b up to 2 x 256, and 2 x (16..300)) is supported."""
    for n in range(1, 1100):
        if lib.osc_fft_length_supported(n):
            assert max(prime_factors(n) | {1}) <= 23, n
    for n in range(16, 301):
        assert lib.osc_fft_length_supported(2 * n) == (max(prime_factors(n) | {1}) <= 23), n
    for a in range(0, 9):
        for b in range(0, 6):
            n = 2 ** a * 3 ** b
            if 32 <= n <= 256:
                assert lib.osc_fft_length_supported(2 * n), n
    assert not lib.osc_fft_length_supported(0) and not lib.osc_fft_length_supported(-4)
    assert not lib.osc_fft_length_supported(29) and not lib.osc_fft_length_supported(2 * 31)


def test_every_doubled_2a3b_length_is_planned(lib):
    """Each doubled length 2n of a mesh edge n = 2^a 3^b in [8, 512] has a
    compile-time plan (the runtime-radix engine is an order of magnitude
    slower); the GPU plan tests (tests/test_gpu_plans.py) check exactly this
    list, so a plan added without a test fails here."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import test_gpu_plans
    want = sorted(2 * 2 ** a * 3 ** b for a in range(10) for b in range(7) if 8 <= 2 ** a * 3 ** b <= 512)
    planned = [n for n in range(1, 2049) if lib.osc_fft_length_planned(n)]
    assert planned == want == test_gpu_plans.PLANNED
    for n in test_gpu_plans.GENERIC:
        assert lib.osc_fft_length_supported(n) and not lib.osc_fft_length_planned(n)


def test_stage_names(lib):
    names = [lib.osc_stage_name(i).decode() for i in range(_capi.Context.NSTAGE)]
    assert len(set(names)) == len(names) and all(names)
    assert lib.osc_stage_name(-1) == b"?" and lib.osc_stage_name(_capi.Context.NSTAGE) == b"?"


def test_null_context_is_rejected(lib):
    # every context-taking call rejects a null context before touching the GPU
    assert lib.osc_ctx_destroy(None) == _capi.OSC_OK
    assert lib.osc_ctx_synchronize(None) == _capi.OSC_ERR_ARG
    assert lib.osc_ctx_set_precision(None, 1) == _capi.OSC_ERR_ARG
    assert lib.osc_ctx_last_error(None) == b"null context"
    m = _capi.OscMesh()
    assert lib.osc_space_charge_3d(None, C.byref(m), None, None, 0, 0, None, None) == _capi.OSC_ERR_ARG
    assert lib.osc_ctx_create(0, None) == _capi.OSC_ERR_ARG


@pytest.mark.skipif(gpu_present(), reason="checks the no-GPU error path")
def test_no_device_fails_loudly(lib):
    h = C.c_void_p()
    assert lib.osc_ctx_create(0, C.byref(h)) == _capi.OSC_ERR_NO_DEVICE and not h.value
    with pytest.raises(_capi.OSCError):
        _capi.Context(0)
    # the Python API has no CPU fallback: a solve without a GPU raises
    import openspacecharge_amd as osc
    with pytest.raises(_capi.OSCError):
        osc.set_context(None)
        osc.osc_freespace_solver2(np.zeros((4, 4, 4), order="F"), 2.0, (1e-3, 1e-3, 1e-3),
                                  efield=np.zeros((4, 4, 4, 3), order="F"))


def test_missing_library_raises(tmp_path):
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from openspacecharge_amd import _capi\n"
            "try:\n    _capi.load(%r)\nexcept _capi.LibraryNotBuilt as e:\n    print('raised')\n"
            % (ROOT, str(tmp_path / "absent.so")))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, check=True)
    assert out.stdout.strip() == "raised"


def test_product_package_never_imports_the_oracle():
    """The oracle is test infrastructure: nothing under the package names it."""
    pkg = os.path.join(ROOT, "openspacecharge_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".inl", ".f90", ".cpp")):
                txt = open(os.path.join(dirpath, f), errors="replace").read()
                assert "oracle" not in txt.lower(), os.path.join(dirpath, f)
