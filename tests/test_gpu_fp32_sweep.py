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
    assert cmp["pooled_rate_rel_err"] < 0.01
    assert cmp["corr_abs_err_median"] < 0.02


def test_sweep_256_networks_5s_fp32_meets_the_accuracy_bar():
    """The 256-network sweep (N log-uniform in [200, 5000], w_n2a U[0.1,
    0.4], in-degree 100), 0.5 s transient + 5 s recorded, batched FP32
    against batched FP64, network by network: rate error below 1% and
    correlation error below 0.02 in the median, pooled over the sweep, and
    in at least 99% of the networks -- the FP64 chaos floor (FP64 against
    itself with every initial V moved by 1e-6 mV) keeps 99.6% / 98.8% of
    them within these bounds on this sweep."""
    _announce("256-network sweep, 5 s, FP32 vs FP64")
    specs = sweep.sweep_specs(256)
    models = [sweep.build(config_path("use_case"), s, 5000.0, 500.0) for s in specs]
    b64 = sweep.run_batched(models, capi.ASG_FP64)
    b32 = sweep.run_batched(models, capi.ASG_FP32)
    cmp = sweep.compare(models, b64, b32, rate_tol=0.01, corr_tol=0.02)
    rows = cmp.pop("rows")
    print({k: v for k, v in cmp.items()}, flush=True)
    assert cmp["networks"] == 256
    assert cmp["rate_rel_err_median"] < 0.01
    assert cmp["pooled_rate_rel_err"] < 0.01
    assert cmp["corr_abs_err_median"] < 0.02
    assert cmp["frac_rate_within_tol"] >= 0.99, [r for r in rows if r["rate_rel_err"] >= 0.01]
    assert cmp["frac_corr_within_tol"] >= 0.99, [r for r in rows if r["corr_abs_err"] >= 0.02]
    assert cmp["regime_agreement"] >= 0.99


# Delays of 13, 21 and 17 steps (windows of 13), SIC every third step, a SIC
# threshold just above the initial Ca.
ODD_TIMING = {"synapse.d_e": 1.3, "synapse.d_i": 2.1, "synapse.d_a2n": 1.7,
              "synapse.sic_interval": 3, "astrocyte.SIC_th": 0.076}


def _fp32_model(seed, **assign):
    m = A.Model(config_path("use_case"), seed=seed, sim_time=300.0, pre_time=0.0,
                assignments=dict(ODD_TIMING, **assign))
    m.desc.contents.precision = capi.ASG_FP32
    return m


def _fp32_run(m, chunks):
    net = A.DeviceNetwork(m)
    steps = m.info["sim_steps"]
    done, k, parts = 0, 0, []
    while done < steps:
        c = min(chunks[k % len(chunks)], steps - done)
        net.step(c)
        done += c
        k += 1
        parts.append(net.spikes())
    out = ({f: net.state(f) for f in ("V", "w", "g_ex", "Ca", "IP3", "I_sic")},
           np.concatenate([p[0] for p in parts]), np.concatenate([p[1] for p in parts]))
    net.close()
    return out


def test_fp32_results_do_not_depend_on_call_lengths():
    """The FP32 step rounds its conductances to FP32 and carries V, w, Ca,
    IP3 and h in FP64 at every step, and the fixed-point sums are
    order-free, so calls of 7/23 steps (windows cut at every call) and one
    call give the same bits."""
    _announce("fp32 call lengths")
    m = _fp32_model(5)
    a_state, a_s, a_n = _fp32_run(m, [m.info["sim_steps"]])
    b_state, b_s, b_n = _fp32_run(m, [7, 23])
    assert len(a_s) > 0
    assert np.array_equal(a_s, b_s) and np.array_equal(a_n, b_n)
    for f in a_state:
        assert np.array_equal(a_state[f].view(np.uint64), b_state[f].view(np.uint64)), f


def test_fp32_network_alone_equals_batch_with_shorter_windows():
    """In a batch with a network of shorter delays (windows of 7 steps) an
    FP32 network gives the same bits as alone (windows of 13)."""
    _announce("fp32 alone vs shorter-window batch")
    m = _fp32_model(5)
    alone_state, alone_s, alone_n = _fp32_run(m, [m.info["sim_steps"]])
    short = _fp32_model(8, **{"synapse.d_e": 0.7})
    batch = A.Batch(4)
    sa = batch.add(m, m.info["sim_steps"])
    batch.add(short, short.info["sim_steps"])
    batch.step(m.info["sim_steps"])
    s, n = batch.spikes(sa)
    assert np.array_equal(s, alone_s) and np.array_equal(n, alone_n)
    for f in alone_state:
        assert np.array_equal(batch.state(sa, f).view(np.uint64),
                              alone_state[f].view(np.uint64)), f
    batch.close()


def test_fp32_delivery_list_longer_than_its_wavefronts():
    """Fourteen FP32 Sparse networks in one batch: in some windows more
    64-neuron blocks spike than there are delivery wavefronts (512), so a
    wavefront delivers several listed blocks. Each network's raster and
    state equal its solo run, where every listed block has a wavefront of
    its own."""
    _announce("fp32 delivery list beyond 512 blocks")
    steps, window = 2000, 10  # Sparse: shortest delay 1 ms = 10 steps
    models = []
    for seed in range(1, 15):
        m = A.Model(config_path("paper_sparse"), seed=seed)
        m.desc.contents.precision = capi.ASG_FP32
        models.append(m)
    batch = A.Batch(len(models))
    slots = [batch.add(m, steps) for m in models]
    batch.step(steps)
    drained = batch.spikes_all()
    together = [(drained[s], batch.state(s, "V"), batch.state(s, "Ca")) for s in slots]
    batch.close()
    listed = np.zeros(steps // window, np.int64)
    for (st, src), _, _ in together:
        blocks = np.unique(np.stack([st // window, src // 64], axis=1), axis=0)
        listed += np.bincount(blocks[:, 0], minlength=steps // window)
    assert listed.max() > 512, listed.max()
    for m, ((st, src), v, ca) in zip(models, together):
        net = A.DeviceNetwork(m)
        net.step(steps)
        s1, n1 = net.spikes()
        assert np.array_equal(st, s1) and np.array_equal(src, n1)
        assert np.array_equal(v.view(np.uint64), net.state("V").view(np.uint64))
        assert np.array_equal(ca.view(np.uint64), net.state("Ca").view(np.uint64))
        net.close()
