"""The allocation-free entry point zmpm_calculate_dt_noalloc against
zmpm_calculate_dt (include/zmpm.h): the same reduction with its partial maxima
in the library's static device array and the result in caller device memory.
Two calls in a row, the second on a smaller input with fewer partial rows, must
each equal the allocating entry point's result exactly (maxima are exact)."""
import ctypes as C

import numpy as np
import pytest

from ziran2020_amd import _abi as A, scenes

pytestmark = pytest.mark.gpu


class DeviceBuffer:
    """8 doubles of device memory through the HIP runtime the library uses."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(64)) == 0
        nan = np.full(8, np.nan)
        assert self.hip.hipMemcpy(self.ptr, nan.ctypes.data_as(C.c_void_p), C.c_size_t(64), 1) == 0

    def read(self):
        host = np.zeros(8)
        assert self.hip.hipMemcpy(host.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(64), 2) == 0
        return host

    def free(self):
        self.hip.hipFree(self.ptr)


def noalloc(sim, max_dt, wgs, buf):
    return sim.L.zmpm_calculate_dt_noalloc(sim.h, max_dt, wgs, C.cast(buf.ptr, C.POINTER(C.c_double)))


def check_against_allocating(sim, max_dt, wgs):
    buf = DeviceBuffer()
    assert noalloc(sim, max_dt, wgs, buf) == A.OK
    sim.synchronize()
    got = buf.read()
    buf.free()
    want = sim.calculate_dt(max_dt)
    assert got[0] == want["dt"] and got[1] == want["max_speed"]
    assert np.array_equal(got[2:5], want["min_corner"]) and np.array_equal(got[5:8], want["max_corner"])
    return got


def test_noalloc_calculate_dt_matches_the_allocating_entry_point(gpu):
    rng = np.random.default_rng(31)
    big = scenes.elastic_cube(n_side_cells=12)
    big.v = rng.normal(scale=2.0, size=big.v.shape)
    small = scenes.elastic_cube(n_side_cells=3, seed=5)
    small.v = rng.normal(scale=0.1, size=small.v.shape)
    a = big.make(gpu, A.PREC_FP32)
    a.run(3, big.dt, 0.0)
    ga = check_against_allocating(a, 1e-2, 256)
    # second call: a smaller input and fewer partial rows than the first left
    # behind; a row of the first call read again would raise its speed and its
    # max corner to the first call's
    b = small.make(gpu, A.PREC_FP64)
    gb = check_against_allocating(b, 1e-2, 3)
    assert gb[1] < ga[1] and np.all(gb[5:8] < ga[5:8])
    # an empty handle: speed 0 and dt = max_dt, as the allocating entry point gives
    e = small.make(gpu, A.PREC_FP64)
    z3 = np.zeros((0, 3))
    e.upload(z3, z3, np.zeros((0, 9)), np.zeros((0, 9)), np.zeros(0), np.zeros(0), np.zeros(0, np.int32))
    ge = check_against_allocating(e, 5e-3, 7)
    assert ge[0] == 5e-3 and ge[1] == 0.0
    for s in (a, b, e):
        s.close()


def test_noalloc_calculate_dt_rejects_sizes_beyond_its_scratch(gpu):
    sim = scenes.elastic_cube(n_side_cells=2).make(gpu, A.PREC_FP32)
    buf = DeviceBuffer()
    for wgs in (0, -1, 1025, 1 << 20):
        assert noalloc(sim, 1e-2, wgs, buf) == A.E_ARG
        assert b"1024" in sim.L.zmpm_last_error(sim.h)
    assert noalloc(sim, 1e-2, 1024, buf) == A.OK
    assert noalloc(sim, -1.0, 8, buf) == A.E_ARG
    sim.synchronize()
    buf.free()
    sim.close()
