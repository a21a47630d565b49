"""FP32 step, device-side recording and batched sweeps on the device."""

import numpy as np
import pytest

from conftest import config_path

import astrosimgpu_amd as A
import oracle
from astrosimgpu_amd import _capi as capi
from astrosimgpu_amd import sweep

pytestmark = pytest.mark.gpu


def _announce(what):
    print(f"[gpu case] {what}", flush=True)


def test_recording_equals_oracle_state_at_recorded_steps():
    _announce("recording")
    m = A.Model(config_path("quick"), seed=2)
    net = A.DeviceNetwork(m)
    net.set_recording(start_step=50, every=37, n_astro=12, capacity=8)
    o = oracle.OracleRun(m.desc, oracle.LIBM_AMATH)
    expect_steps, expect_ca, expect_ip3 = [], [], []
    for step in range(0, 300):
        o.step(1)
        if step >= 50 and (step - 50) % 37 == 0:
            expect_steps.append(step)
            expect_ca.append(o.state("Ca")[:12])
            expect_ip3.append(o.state("IP3")[:12])
    net.step(120)
    s1, ca1, ip1 = net.read_trace()
    net.step(180)
    s2, ca2, ip2 = net.read_trace()
    steps = np.concatenate([s1, s2])
    assert steps.tolist() == expect_steps
    assert np.array_equal(np.concatenate([ca1, ca2]), np.array(expect_ca))
    assert np.array_equal(np.concatenate([ip1, ip2]), np.array(expect_ip3))


def test_recording_ring_overflow_is_reported():
    _announce("recording overflow")
    m = A.Model(config_path("quick"), seed=2)
    net = A.DeviceNetwork(m)
    net.set_recording(0, 10, 5, 4)
    net.step(100)
    with pytest.raises(A.AsgError) as e:
        net.read_trace()
    assert e.value.status == capi.ASG_ERR_OVERFLOW


def test_fp32_network_tracks_fp64_statistics():
    """Sparse network, 1 s: the FP32 step gives a mean rate close to FP64's
    (identical Poisson drive; trajectories decorrelate, statistics agree)."""
    _announce("fp32 sparse")
    m = A.Model(config_path("paper_sparse"), seed=1)
    rates = {}
    for prec in (capi.ASG_FP64, capi.ASG_FP32):
        m.desc.contents.precision = prec
        net = A.DeviceNetwork(m)
        net.step(10000)
        s, _ = net.spikes()
        rates[prec] = len(s) * 1000.0 / (1000.0 * m.n_neurons)
        v = net.state("V")
        assert np.all(np.isfinite(v))
        net.close()
    assert rates[capi.ASG_FP64] > 1.0
    assert abs(rates[capi.ASG_FP32] - rates[capi.ASG_FP64]) / rates[capi.ASG_FP64] < 0.05


def test_fp32_short_window_trajectory_error():
    """Over 20 ms of the quick network, FP32 state stays within FP32-level
    relative error of FP64 (before chaos decorrelates spike times)."""
    _announce("fp32 short window")
    m = A.Model(config_path("quick"), seed=1)
    out = {}
    for prec in (capi.ASG_FP64, capi.ASG_FP32):
        m.desc.contents.precision = prec
        net = A.DeviceNetwork(m)
        net.step(200)
        out[prec] = {f: net.state(f) for f in ("V", "w", "Ca", "IP3", "h")}
        net.close()
    for f in ("V", "Ca", "IP3", "h"):
        a, b = out[capi.ASG_FP64][f], out[capi.ASG_FP32][f]
        rel = np.max(np.abs(a - b) / np.maximum(np.abs(a), 1e-3))
        assert rel < 1e-4, (f, rel)


def test_small_sweep_fp64_batch_equals_plain_and_fp32_is_close():
    _announce("small sweep")
    specs = sweep.sweep_specs(count=12, n_min=200, n_max=1500, gen_seed=11)
    models = [sweep.build(config_path("use_case"), s, 1000.0, 200.0) for s in specs]
    plain = sweep.run_plain(models[:4])
    b64 = sweep.run_batched(models, capi.ASG_FP64)
    for k in range(4):
        assert np.array_equal(plain.spikes[k], b64.spikes[k])
        assert np.array_equal(plain.traces[k][1], b64.traces[k][1])
    b32 = sweep.run_batched(models, capi.ASG_FP32)
    cmp = sweep.compare(models, b64, b32)
    assert cmp["networks"] == 12
    assert cmp["pooled_rate_rel_err"] < 0.05
    assert cmp["corr_abs_err_median"] < 0.1
