"""asg_device_math_noalloc, the no-allocation variant of asg_device_math
(inputs and outputs staged through a file-scope device array instead of
memory allocated per call), against asg_device_math itself through the C
ABI: bit for bit, on two calls in a row with the second smaller (the array
keeps the first call's contents), and sizes beyond the array refused before
any device work."""

import numpy as np
import pytest

import astrosimgpu_amd as A
from astrosimgpu_amd import _capi as capi

pytestmark = pytest.mark.gpu

MAX_VALUES = 1 << 20
EXP, LOG, COS = 0, 1, 3


def _call(name, fn, x):
    y = np.full_like(x, np.nan)
    st = getattr(A.lib(), name)(fn, x.ctypes.data_as(capi.DBLP), y.ctypes.data_as(capi.DBLP),
                                x.size)
    return st, y


def test_noalloc_equals_the_allocating_entry_point_on_two_calls():
    print("[gpu case] device math without allocation", flush=True)
    rng = np.random.default_rng(5)
    for n in (MAX_VALUES, 1000):
        for fn, lo, hi in ((EXP, -30.0, 30.0), (LOG, 1e-6, 1e3), (COS, 0.0, 6.3)):
            x = rng.uniform(lo, hi, n)
            st_a, ya = _call("asg_device_math", fn, x)
            st_b, yb = _call("asg_device_math_noalloc", fn, x)
            assert st_a == capi.ASG_OK and st_b == capi.ASG_OK
            assert np.array_equal(ya.view(np.uint64), yb.view(np.uint64)), (fn, n)


def test_noalloc_refuses_sizes_beyond_its_array():
    print("[gpu case] device math without allocation: size limit", flush=True)
    x = np.zeros(MAX_VALUES + 1)
    st, y = _call("asg_device_math_noalloc", EXP, x)
    assert st == capi.ASG_ERR_CAPACITY
    assert b"exceeds its scratch" in A.lib().asg_last_error()
    assert np.all(np.isnan(y))
