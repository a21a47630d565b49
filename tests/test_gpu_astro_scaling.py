"""The isolated astrocyte update at scale against the CPU oracle: same initial
state, per-cell parameters and Poisson spike counts from the same counter
streams; FP64 bit for bit, FP32 within FP32 accuracy."""

import numpy as np
import pytest

import oracle
from astrosimgpu_amd import _capi as capi
from astrosimgpu_amd import astro_bench as AB

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("randomized", [False, True])
def test_astro_scaling_fp64_equals_oracle(randomized):
    print(f"[gpu case] astro scaling fp64 randomized={randomized}", flush=True)
    constants, p = AB.setup(randomized=randomized)
    n, steps = 3000, 300
    _, (ca, ip3, h) = AB.run(n, steps, capi.ASG_FP64, constants, p)
    oca, oip3, oh = oracle.astro_scaling(n, steps, AB.SEED, constants,
                                         (p.Ca_tot, p.IP3_0, p.tau_IP3, p.delta_IP3), p.dt,
                                         p.lambda_, randomized, p.lower, p.upper, p.var)
    for got, want in ((ca, oca), (ip3, oip3), (h, oh)):
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    assert np.all((ca >= 0.0) & (h >= 0.0) & (h <= 1.0) & (ip3 >= 0.0))


def test_astro_scaling_fp32_tracks_fp64():
    print("[gpu case] astro scaling fp32", flush=True)
    constants, p = AB.setup()
    n, steps = 20000, 1000
    r64, s64 = AB.run(n, steps, capi.ASG_FP64, constants, p)
    r32, s32 = AB.run(n, steps, capi.ASG_FP32, constants, p)
    for a, b in zip(s64, s32):
        mx, mean = AB.rel_err(b, a)
        assert mean < 1e-5 and mx < 1e-3
    assert r32["bytes_per_launch"] == 28 * n and r64["bytes_per_launch"] == 52 * n
