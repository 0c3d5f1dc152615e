"""pic_mglin_noalloc (no allocation on any call) against pic_mglin, bit for
bit, through the C ABI: two calls in a row on the same GPU, the second on a
smaller grid, so the second call runs over the first one's leftovers in the
shared scratch array.  Device buffers come from the HIP runtime the library
itself uses (loaded by libpic2d.so; ctypes reuses it)."""
import ctypes as C

import numpy as np
import pytest

import pic2d

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    pic2d.load_library()  # maps the runtime libpic2d.so links against
    rt = C.CDLL("libamdhip64.so.7")
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipDeviceSynchronize.argtypes = []
    return rt


def solve_noalloc(rt, f, ncycle, length):
    n = f.shape[0]
    a = np.ascontiguousarray(f, dtype=np.float64)
    d = C.c_void_p()
    assert rt.hipMalloc(C.byref(d), a.nbytes) == 0  # the caller's buffer, not the library's
    try:
        assert rt.hipMemcpy(d, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0
        lib = pic2d.load_library()
        rc = lib.pic_mglin_noalloc(d, n, ncycle, length, None)
        assert rc == 0, lib.pic_last_error(None)
        assert rt.hipDeviceSynchronize() == 0
        out = np.empty_like(a)
        assert rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), d, a.nbytes, D2H) == 0
    finally:
        rt.hipFree(d)
    return out


@pytest.mark.parametrize("sizes", [(1025, 257), (513, 65)])
def test_mglin_noalloc_matches_mglin_on_two_calls(sizes):
    rt = hip()
    L = 0.0184
    for k, n in enumerate(sizes):
        print(f"noalloc call {k}: n={n}", flush=True)
        rng = np.random.default_rng(11 * n + k)
        f = rng.standard_normal((n, n)) * 1e5
        for ncycle in (2, 1):
            got = solve_noalloc(rt, f, ncycle, L)
            want = pic2d.mglin(f, ncycle, L)
            assert np.array_equal(got, want), (n, ncycle, np.max(np.abs(got - want)))


def test_mglin_noalloc_capacity_error_on_gpu():
    rt = hip()
    d = C.c_void_p()
    assert rt.hipMalloc(C.byref(d), 8 * 2049 * 2049) == 0
    try:
        lib = pic2d.load_library()
        assert lib.pic_mglin_noalloc(d, 2049, 2, 1.0, None) == -6
    finally:
        rt.hipFree(d)
