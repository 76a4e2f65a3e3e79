"""The allocation-free norms entry point (dgtd_op_norms_noalloc) through the
C ABI: its FP64 norm, energy and non-finite count equal those of
dgtd_op_norms bitwise (the same kernels and partial order, the partials in the
library's static scratch instead of the operator's buffer), for the resident
state and for a caller's device state, on two calls in a row on different
operators, the second one smaller."""
import ctypes

import numpy as np
import pytest

import dgtd_amd as D
from dgtd_amd._abi import check, lib
import meshgen as MG

pytestmark = pytest.mark.gpu


def hip():
    lib()  # libdgtd_hip.so brings the HIP runtime it was linked with
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


def noalloc_norms(h, op, d_state, d_out):
    check(lib().dgtd_op_norms_noalloc(op._h, d_state, d_out, None))
    assert h.hipDeviceSynchronize() == 0
    out = np.zeros(3)
    assert h.hipMemcpy(out.ctypes.data, d_out, 24, D2H) == 0
    return out


def test_noalloc_norms_equal_norms_on_two_calls():
    h = hip()
    d_out = ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(d_out), 24) == 0
    bufs = []
    try:
        for n, order, prec in ((4, 3, 32), (2, 2, 64)):  # the second call smaller
            vx, E, fbc = MG.cube_mesh(n, 1.0, "pec")
            op = D.HesthavenEvolution(vx, E, fbc, order, precision=prec, integrator="lserk4")
            pos = op.node_positions()
            q = MG.gaussian(pos, (0.4, 0.55, 0.5), 0.15, comps=range(6))
            op.set_state(q)
            op.advance(1e-3, 2)
            want = np.array([op.norms()[k] for k in ("norm2", "energy", "nonfinite")], dtype=float)
            got = noalloc_norms(h, op, None, d_out)
            np.testing.assert_array_equal(got, want)
            # a caller's device state in the operator's precision
            q2 = np.cos(3.0 * q) * (1.0 + np.arange(q.size) % 5)
            q2[7] = np.nan
            host = q2.astype(np.float32 if prec == 32 else np.float64)
            d_q = ctypes.c_void_p()
            assert h.hipMalloc(ctypes.byref(d_q), host.nbytes) == 0
            bufs.append(d_q)
            assert h.hipMemcpy(d_q, host.ctypes.data, host.nbytes, H2D) == 0
            got2 = noalloc_norms(h, op, d_q, d_out)
            op.set_state(host.astype(np.float64))
            want2 = np.array([op.norms()[k] for k in ("norm2", "energy", "nonfinite")], dtype=float)
            np.testing.assert_array_equal(got2, want2)
            assert got2[2] == 1.0
            op.close()
    finally:
        for b in bufs:
            h.hipFree(b)
        h.hipFree(d_out)


def test_noalloc_rejects_null_output():
    vx, E, fbc = MG.cube_mesh(2, 1.0, "pec")
    op = D.HesthavenEvolution(vx, E, fbc, 2, precision=32)
    assert lib().dgtd_op_norms_noalloc(op._h, None, None, None) == -1
    op.close()
