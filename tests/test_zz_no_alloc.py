"""tbl_div_check_noalloc: the divergence check on caller-owned device memory
with its block partials in the library's static device scratch (no
allocation).  Checked against tbl_div_check on two calls in a row, the second
on a smaller grid (the scratch keeps the first call's partials), and on a
mixed-precision context.  The test plays the caller: it owns the device
buffers, allocated through the HIP runtime the library itself is linked to."""
import ctypes as C

import numpy as np
import pytest

from tests.case_util import C1_GRID, RAGGED, gpu_from_case, make_case

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    from turbulentbl_amd import tbl
    tbl.load()                                   # the library's libamdhip64 is now loaded
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


def _noalloc(sv, Ut, dtype):
    hip = _hip()
    host = np.ascontiguousarray(np.asfortranarray(Ut).ravel(order="F").astype(dtype))
    out = np.full(2, -1.0)
    d_ut, d_out = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(d_ut), host.nbytes) == 0
    assert hip.hipMalloc(C.byref(d_out), out.nbytes) == 0
    try:
        assert hip.hipMemcpy(d_ut, host.ctypes.data, host.nbytes, H2D) == 0
        assert hip.hipMemcpy(d_out, out.ctypes.data, out.nbytes, H2D) == 0
        sv.div_check_noalloc(d_ut.value, d_out.value)
        sv.sync()
        assert hip.hipMemcpy(out.ctypes.data, d_out, out.nbytes, D2H) == 0
    finally:
        hip.hipFree(d_ut)
        hip.hipFree(d_out)
    return out


def test_noalloc_div_check_equals_div_check_on_two_calls():
    got = []
    for grid, seed in ((C1_GRID, 1), (RAGGED, 2)):      # second call on a smaller grid
        g, s = make_case(grid, seed=seed)
        with gpu_from_case(g, s) as sv:
            ref = sv.div_check()
            out = _noalloc(sv, s.Ut, np.float64)
        print("noalloc %s: %r vs tbl_div_check %r" % (grid, out.tolist(), ref), flush=True)
        assert out[0] == ref[0] and out[1] == ref[1]
        got.append(out)
    assert got[0][1] != got[1][1]


def test_noalloc_div_check_mixed_context():
    g, s = make_case(RAGGED, seed=5)
    with gpu_from_case(g, s, precision="mixed") as sv:
        ref = sv.div_check()
        out = _noalloc(sv, s.Ut, np.float32)
    assert out[0] == ref[0] and out[1] == ref[1]
