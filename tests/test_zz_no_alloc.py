"""The no-allocation fill (lgh_hip_randn_fill_noalloc) through the C ABI on
an MI355X: into device memory the test allocates itself, it writes exactly
the draws of lgh_hip_randn_fill (the reference's, bit for bit) on two calls
in a row, the second smaller (the shared scratch keeps the first call's
list), leaves the rows between nr and ld alone, and refuses a request above
its limit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipDeviceSynchronize.argtypes = []
    return h


def _fill_on_device(hip, lib, seed, row0, col0, nr, nc, prec, ld):
    dt = np.float64 if prec == 0 else np.float32
    host = np.full((nc, ld), np.nan, dtype=dt)          # column j at j * ld
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), host.nbytes) == 0
    try:
        assert hip.hipMemcpy(p, host.ctypes.data, host.nbytes, 1) == 0     # H2D
        rc = lib.lgh_hip_randn_fill_noalloc(seed, row0, col0, nr, nc, prec, p, ld)
        assert rc == 0, lib.lgh_hip_last_error()
        assert hip.hipMemcpy(host.ctypes.data, p, host.nbytes, 2) == 0     # D2H
    finally:
        hip.hipFree(p)
    return host


def test_noalloc_fill_equals_randn_fill(lgh, hip):
    lgh.init()
    lib = lgh.lib()
    cases = [(271828, 5, 3, 70001, 70, 70003, 0),     # 4.9 M draws, ~108 k listed
             (271828, 5, 3, 1000, 9, 1000, 0)]        # then a smaller one
    for seed, row0, col0, nr, nc, ld, prec in cases:
        got = _fill_on_device(hip, lib, seed, row0, col0, nr, nc, prec, ld)
        ref = lgh.randn_fill(seed, row0, col0, nr, nc, W=64, prec=prec)
        assert np.array_equal(got[:, :nr].T.astype(np.float64), ref), (nr, nc)
        assert np.isnan(got[:, nr:]).all()
    for seed, row0, col0, nr, nc, ld, prec in [(7, 0, 10, 50000, 33, 50000, 1),
                                               (7, 0, 10, 300, 3, 301, 1)]:
        got = _fill_on_device(hip, lib, seed, row0, col0, nr, nc, prec, ld)
        ref = lgh.randn_fill(seed, row0, col0, nr, nc, W=64, prec=prec)
        assert np.array_equal(got[:, :nr].T.astype(np.float64), ref), (nr, nc)


def test_noalloc_fill_refuses_above_limit(lgh):
    lgh.init()
    lib = lgh.lib()
    rc = lib.lgh_hip_randn_fill_noalloc(1, 0, 0, 1 << 13, (1 << 12) + 1, 0,
                                        C.c_void_p(16), 1 << 13)
    assert rc != 0
    assert b"33554432" in lib.lgh_hip_last_error()
