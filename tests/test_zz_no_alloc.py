"""ibfs_ctci_noalloc, the entry point that allocates nothing: the same
Poisson solve as ibfs_op_ctci (src/myfft.f90:97-109), on device buffers the
caller owns, with the library's intermediates in its static device array.
Two calls in a row, the second on a smaller grid (the array keeps the first
call's contents), each against ibfs_op_ctci on the same input to the FP64
transform bar; the buffers come from the HIP runtime the library itself
links (libamdhip64.so.7), through ctypes, so nothing here goes through
another runtime."""
import ctypes as C

import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu


def hip_runtime():
    rt = C.CDLL("libamdhip64.so.7")
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipDeviceSynchronize.argtypes = []
    return rt


H2D, D2H = 1, 2


def solver(m, n):
    from ibfs_amd import geometry, hip

    len_ = 4.0
    d = len_ / m
    xb, yb, bd = geometry.circle(d, 1.0, 0.0, 0.0)
    return hip.Solver(m, n, 2, 0.25 * d, 100.0, len_, 2.0, 2.0 * n / m, xb, yb, bd)


def noalloc_ctci(rt, m, n, x):
    from ibfs_amd import hip

    lib = hip.lib()
    lib.ibfs_ctci_noalloc.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    nbytes = x.size * 8
    din, dout = C.c_void_p(), C.c_void_p()
    assert rt.hipMalloc(C.byref(din), nbytes) == 0 and rt.hipMalloc(C.byref(dout), nbytes) == 0
    try:
        assert rt.hipMemcpy(din, x.ctypes.data, nbytes, H2D) == 0
        rc = lib.ibfs_ctci_noalloc(m, n, din, dout, None)
        assert rc == 0, lib.ibfs_last_error(None)
        assert rt.hipDeviceSynchronize() == 0
        out = np.empty_like(x)
        assert rt.hipMemcpy(out.ctypes.data, dout, nbytes, D2H) == 0
        return out
    finally:
        rt.hipFree(din)
        rt.hipFree(dout)


def test_noalloc_ctci_matches_op_ctci_on_two_calls():
    rt = hip_runtime()
    rng = np.random.default_rng(7)
    for m, n in ((512, 512), (200, 256)):  # the second call smaller than the first
        x = rng.standard_normal((m - 1) * (n - 1))
        ref = solver(m, n).ctci(x)
        got = noalloc_ctci(rt, m, n, x)
        print(f"{m}x{n}: relative difference {relerr(got, ref):.2e}")
        assert relerr(got, ref) < 1e-12


def test_noalloc_rejects_sizes_beyond_its_scratch():
    from ibfs_amd import hip

    lib = hip.lib()
    lib.ibfs_ctci_noalloc.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    # checked on the host before any device work: no buffers are touched
    assert lib.ibfs_ctci_noalloc(2048, 2048, C.c_void_p(8), C.c_void_p(8), None) == 1
    assert b"1024" in lib.ibfs_last_error(None)
    assert lib.ibfs_ctci_noalloc(96, 96, C.c_void_p(8), C.c_void_p(8), None) == 1
