"""cbz_sur_predict_device, the inference entry point that allocates
nothing (a fixed device array holds its staging), against
cbz_sur_predict: the same predictions bit for bit, on two calls in a row
(the second smaller), and a clear error beyond its row limit.

The device buffers come from the HIP runtime the library itself loaded
(dlopen by soname returns the library already in the process), so the test
does not depend on a second runtime being initialised beside it."""
import ctypes

import numpy as np
import pytest

from cubez_amd import World, scenes, surrogate

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


def _dev(hip, host):
    p = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(p), host.nbytes) == 0
    assert hip.hipMemcpy(p, host.ctypes.data, host.nbytes, H2D) == 0
    return p


def test_predict_device_equals_predict_without_allocating():
    params, norm = surrogate.load()
    rng = np.random.default_rng(11)
    with World(scenes.cubedrop(0, 8), precision="f32") as w:
        hip = _hip()
        w.sur_set_norm(norm)
        w.sur_set_params(params)
        fn = w._lib.cbz_sur_predict_device
        for rows in (5000, 777):
            # features around the training data's statistics
            X = (norm[:40] + rng.standard_normal((rows, 40)) * norm[40:80]).astype(np.float32)
            want = w.sur_predict(X)
            got = np.full((rows, 13), np.nan, dtype=np.float32)
            xd, yd = _dev(hip, X), _dev(hip, got)
            try:
                assert hip.hipDeviceSynchronize() == 0
                rc = fn(w._h, rows, xd, yd)
                assert rc == 0, w._lib.cbz_last_error(w._h)
                assert hip.hipMemcpy(got.ctypes.data, yd, got.nbytes, D2H) == 0
                assert np.array_equal(got, want), (rows, np.nanmax(np.abs(got - want)))
                if rows == 777:
                    # beyond the row limit: an argument error, nothing launched
                    rc = fn(w._h, 65537, xd, yd)
                    assert rc == -1
                    assert b"65536" in w._lib.cbz_last_error(w._h)
            finally:
                hip.hipFree(xd)
                hip.hipFree(yd)
