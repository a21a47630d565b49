"""The host ordering of FP32 rasters (csrc/raster_order.h) on the CPU: the
entries of several windows in the shapes the delivery pass appends them
(one ascending run per (block, step), runs of a window in any order) come out
sorted by (step, source), and so do inputs of any other shape (the radix
path and the comparison-sort fallback for keys too wide or ranges too
short)."""

import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "astrosimgpu_amd", "csrc")

DRIVER = r"""
#include "raster_order.h"
extern "C" void order(int64_t* step, uint32_t* src, size_t n) {
    std::vector<uint64_t> keys, tmp;
    asg::order_raster(step, src, n, keys, tmp);
}
"""


@pytest.fixture(scope="module")
def order(tmp_path_factory):
    d = tmp_path_factory.mktemp("raster_order")
    cpp = d / "driver.cpp"
    cpp.write_text(DRIVER)
    so = d / "driver.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", CSRC, str(cpp), "-o",
                    str(so)], check=True)
    lib = ctypes.CDLL(str(so))
    lib.order.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]

    def run(step, src):
        step = np.ascontiguousarray(step, np.int64).copy()
        src = np.ascontiguousarray(src, np.uint32).copy()
        lib.order(step.ctypes.data, src.ctypes.data, len(step))
        return step, src
    return run


def _delivery_shaped(rng, windows, W, n_neurons, rate):
    steps, srcs = [], []
    for w in range(windows):
        runs = []
        for k in range(W):
            t = w * W + k + 1000
            fired = np.flatnonzero(rng.random(n_neurons) < rate)
            for b in np.unique(fired >> 6):
                s = fired[(fired >> 6) == b]
                runs.append((np.full(len(s), t), s))
        rng.shuffle(runs)
        for st, s in runs:
            steps.append(st)
            srcs.append(s)
    return np.concatenate(steps), np.concatenate(srcs).astype(np.uint32)


@pytest.mark.parametrize("rate", [0.001, 0.02, 0.3])
def test_delivery_shaped_windows_come_out_sorted(order, rate):
    rng = np.random.default_rng(int(rate * 1000))
    step, src = _delivery_shaped(rng, windows=12, W=10, n_neurons=3000, rate=rate)
    got_s, got_r = order(step, src)
    idx = np.lexsort((src, step))
    assert np.array_equal(got_s, step[idx]) and np.array_equal(got_r, src[idx])


def test_any_input_comes_out_sorted(order):
    rng = np.random.default_rng(5)
    for n in (0, 1, 2, 7, 500):
        step = rng.integers(0, 20, n)
        src = rng.integers(0, 200, n).astype(np.uint32)
        got_s, got_r = order(step, src)
        idx = np.lexsort((src, step))
        assert np.array_equal(got_s, step[idx]) and np.array_equal(got_r, src[idx])
    # steps far apart and large sources (wide keys), and keys past 64 bits
    for lo, hi in ((10**12, 10**12 + 5000), (0, 2**40)):
        step = rng.integers(lo, hi, 3000)
        src = rng.integers(0, 2**31, 3000).astype(np.uint32)
        got_s, got_r = order(step, src)
        idx = np.lexsort((src, step))
        assert np.array_equal(got_s, step[idx]) and np.array_equal(got_r, src[idx])
