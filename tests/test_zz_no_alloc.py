"""gs_acceleration_noalloc (direct sums without device allocation) against
acceleration() (the library's host-buffer entry point) through the C-ABI.

The positions, masses and accelerations are device arrays that the test
allocates through the HIP runtime the library itself links (a second HIP
runtime in the process, such as torch's, would not see the device); the entry
point's own scratch is its static device array.  Two calls in a row, the
second with a smaller input, must give what acceleration() gives for the same
inputs, bit for bit (same kernels, same configuration)."""
import ctypes as C

import numpy as np
import pytest

from grav_sim_amd import _abi, ic
from grav_sim_amd import lib as L

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    L.load()  # the library's own libamdhip64 is then the one this name resolves to
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


def _device_call(x, m, method, eps, G=1.0):
    hip = _hip()
    x, m = np.ascontiguousarray(x), np.ascontiguousarray(m)
    n = len(m)
    ptrs = []
    try:
        for nbytes in (x.nbytes, m.nbytes, x.nbytes):
            p = C.c_void_p()
            assert hip.hipMalloc(C.byref(p), nbytes) == 0
            ptrs.append(p)
        dx, dm, da = ptrs
        assert hip.hipMemcpy(dx, x.ctypes.data, x.nbytes, H2D) == 0
        assert hip.hipMemcpy(dm, m.ctypes.data, m.nbytes, H2D) == 0
        nan = np.full_like(x, np.nan)
        assert hip.hipMemcpy(da, nan.ctypes.data, nan.nbytes, H2D) == 0
        assert hip.hipDeviceSynchronize() == 0
        s = _abi.System()
        s.num_particles = n
        s.x = C.cast(dx, _abi.PD)
        s.m = C.cast(dm, _abi.PD)
        s.G = G
        p = L.param(method, eps=eps)
        L.check(L.load().gs_acceleration_noalloc(da, C.byref(s), C.byref(p)), "gs_acceleration_noalloc")
        a = np.empty_like(x)
        assert hip.hipMemcpy(a.ctypes.data, da, a.nbytes, D2H) == 0
        return a
    finally:
        for p in ptrs:
            hip.hipFree(p)


@pytest.mark.parametrize("method,eps", [(16, 1e-3), (17, 1e-3), (1, 0.0), (2, 1e-2)])
def test_noalloc_equals_acceleration_two_calls_second_smaller(method, eps):
    for n, seed in ((1 << 17, 2), (3001, 13)):
        if method in (1, 2) and n > 4096:
            n = 4096
        x, v, m = ic.plummer(n, seed)
        if method in (2, 17):
            m = m.copy()
            m[n // 2:] = 0.0  # second half massless
        a_ref = L.acceleration(x, m, method, eps=eps)
        a = _device_call(x, m, method, eps)
        assert np.array_equal(a, a_ref), (method, n)


def test_noalloc_rejects_sizes_beyond_the_scratch():
    x, v, m = ic.uniform_cube(1 << 21, 3)
    with pytest.raises(RuntimeError, match="too large for the no-allocation scratch"):
        _device_call(x, m, 16, 1e-3)
