"""The tuner of the strided FFT passes' tile width (osc_fft_tuning): the width
only changes how many lines share a workgroup, so every forced width must
give bit-identical transforms and solves, and the online tuner (the first
calls of each pass measure the candidates on real work) must settle on a
width and leave results unchanged."""
import ctypes as C

import numpy as np
import pytest

import openspacecharge_amd as osc
from openspacecharge_amd import _capi as K
from oracle.oracle import gaussian_bunch

pytestmark = pytest.mark.gpu


def _runs(ctx32):
    rng = np.random.default_rng(3)
    out = []
    for shape in ((64, 96, 48), (5, 7, 9), (256, 4, 6), (6, 4, 512)):
        a = np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
        b = np.zeros_like(a, order="F")
        osc.ccfft3d(a, b, (1, -1, 1))
        out.append(b)
    x, y, z = gaussian_bunch(50000, (1e-3, 1e-3, 2e-3), seed=9)
    for c in (None, ctx32):
        for cath in (False, True):
            m = osc.Mesh3D(nhi=(24, 36, 48), gamma=4.0)
            zz = z + (8e-3 if cath else 0.0)
            E, B = osc.space_charge_step(x, y, zz, m, total_charge=1e-9, at_cathode=cath,
                                         calc_bfield=True, ctx=c)
            out += [m.efield.copy(), E, B]
    return out


def test_forced_widths_are_bit_identical(ctx):
    lib = K.load()
    c32 = osc.Context(0, precision=osc.OSC_PREC_FP32)
    try:
        res = []
        for mode in (2, 3, 4):
            assert lib.osc_fft_tuning(mode) == K.OSC_OK
            res.append(_runs(c32))
        for other in res[1:]:
            for a, b in zip(res[0], other):
                assert np.array_equal(a, b)
        # the tuner picks widths on real work and changes nothing
        assert lib.osc_fft_tuning(1) == K.OSC_OK
        for _ in range(6):
            tuned = _runs(c32)
        for a, b in zip(res[1], tuned):
            assert np.array_equal(a, b)
        st = (C.c_int64 * 2)()
        assert lib.osc_fft_tune_stats(st) == K.OSC_OK
        print(f"tuner: {st[0]} decided, {st[1]} measuring")
        assert st[0] >= 4
        assert lib.osc_fft_tuning(7) == K.OSC_ERR_ARG
    finally:
        lib.osc_fft_tuning(1)
        c32.close()
