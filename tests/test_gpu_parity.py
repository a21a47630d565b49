"""Device step against the oracle and the reference build's rasters.

Bar: the spike raster is identical (integer (step, neuron) pairs) and every
state array is bitwise equal to the oracle's after the same number of steps
(that is stricter than the 1e-12 relative tolerance asked of FP64 state).
"""

import numpy as np
import pytest

from _cases import model_for, raster_hash, recorded
from conftest import config_path, golden, golden_exists

import astrosimgpu_amd as A
import oracle
from astrosimgpu_amd import _capi as capi

pytestmark = pytest.mark.gpu

STATE_FIELDS = ("V", "w", "g_ex", "dg_ex", "g_in", "dg_in", "I_sic", "Ca", "IP3", "h")


def _announce(what):
    print(f"[gpu case] {what}", flush=True)


@pytest.mark.parametrize("fn,lo,hi", [
    ("exp", -745.0, 709.0), ("exp", -20.0, 30.0), ("log", 1e-300, 1.0), ("log", 1.0, 3000.0),
    ("sin", 0.0, 6.3), ("cos", 0.0, 6.3), ("sqrt", 0.0, 1e6),
])
def test_device_math_is_host_amath_bit_for_bit(fn, lo, hi):
    _announce(f"device math {fn}")
    rng = np.random.default_rng(20261203)
    x = rng.uniform(lo, hi, 1 << 20)
    if fn == "log":
        x = np.concatenate([x, np.exp(rng.uniform(-700, 700, 1 << 18))])
    host = oracle.math(fn, x, oracle.LIBM_AMATH)
    dev = A.device_math(fn, x)
    assert np.array_equal(host.view(np.uint64), dev.view(np.uint64))


def test_fp32_spike_exp2_error_bound():
    """The FP32 path's 2^y of the AdEx spike current (y rounded to FP32,
    hardware exp): relative error within ln2 |y| 2^-24 (rounding y) plus two
    FP32 ulps, against 2^y in x87 extended precision, over the range the
    exponent takes (V from far below rest to V_peak); below 2^-126 it only
    has to stay that small."""
    _announce("device math exp2_f32")
    rng = np.random.default_rng(11)
    y = np.concatenate([rng.uniform(-60.0, 40.0, 1 << 20), np.linspace(-1.0, 1.0, 4097),
                        np.arange(-60.0, 41.0)])
    got = A.device_math("exp2_f32", y)
    exact = np.exp2(y.astype(np.longdouble))
    rel = np.abs((got.astype(np.longdouble) - exact) / exact).astype(np.float64)
    bound = np.log(2.0) * np.abs(y) * 2.0 ** -24 * 1.0001 + 2.0 ** -22
    assert np.all(rel <= bound), (y[np.argmax(rel / bound)], rel.max())
    tiny = A.device_math("exp2_f32", np.linspace(-150.0, -127.0, 1001))
    assert np.all((tiny >= 0.0) & (tiny <= 2.0 ** -126))


def test_device_division_by_reciprocal_is_ieee_division():
    """The FP64 step divides by its constants through their reciprocals
    (two fused corrections after the product): bit for bit the IEEE
    quotient numpy gives, on 2^20 random pairs over wide ranges, numerators
    of the model's magnitudes over divisors near the shipped constants,
    quotients at binade edges with divisors at both ends of their binade,
    and zero, infinite and NaN numerators."""
    _announce("device math div")
    rng = np.random.default_rng(20261204)
    n = 1 << 20
    b = rng.uniform(1.0, 2.0, n) * 2.0 ** rng.integers(-200, 201, n).astype(np.float64)
    a = rng.uniform(-2.0, 2.0, n) * 2.0 ** rng.integers(-300, 301, n).astype(np.float64)
    consts = np.array([2.0, 130.0, 104.0, 450.0, 264.0, 281.0, 144.0, 0.185, 1058.7589568400845])
    b2 = np.repeat(consts, 1 << 14) * rng.uniform(0.9, 1.1, consts.size << 14)
    a2 = rng.uniform(-200.0, 200.0, b2.size)
    ulp = 2.0 ** -52
    bb = np.concatenate([2.0 - ulp * np.arange(1, 65), 1.0 + ulp * np.arange(1, 65)])
    tgt = np.array([2.0 - 3 * ulp, 2.0 - ulp, 2.0, 1.0, 1.0 + ulp, 1.5])
    a3 = (bb[:, None] * tgt[None, :]).ravel()
    b3 = np.repeat(bb, tgt.size)
    a3 = np.concatenate([np.nextafter(a3, np.inf), a3, np.nextafter(a3, -np.inf)])
    b3 = np.tile(b3, 3)
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan] * 3)
    b4 = np.repeat(np.array([2.0, 130.0, 0.185]), 5)
    A_ = np.concatenate([a, a2, a3, -a3, special])
    B_ = np.concatenate([b, b2, b3, b3, b4])
    got = A.device_math("div", np.stack([A_, B_], axis=1).ravel())[0::2]
    want = A_ / B_
    same = (got.view(np.uint64) == want.view(np.uint64)) | (np.isnan(got) & np.isnan(want))
    assert same.all(), (A_[~same][:5], B_[~same][:5])


def test_device_exp_log_equal_c_library():
    rng = np.random.default_rng(7)
    x = rng.uniform(-30, 30, 1 << 20)
    assert np.array_equal(A.device_math("exp", x).view(np.uint64),
                          oracle.math("exp", x, oracle.LIBM_GLIBC).view(np.uint64))
    y = np.exp(rng.uniform(-60, 10, 1 << 20))
    assert np.array_equal(A.device_math("log", y).view(np.uint64),
                          oracle.math("log", y, oracle.LIBM_GLIBC).view(np.uint64))


def _run_device(m, steps, chunk):
    net = A.DeviceNetwork(m)
    done = 0
    parts_s, parts_n = [], []
    while done < steps:
        k = min(chunk, steps - done)
        net.step(k)
        done += k
        s, n = net.spikes()
        parts_s.append(s)
        parts_n.append(n)
    return net, np.concatenate(parts_s), np.concatenate(parts_n)


CASES = ["quick_s1", "quick_s7", "use_case_2s", "bursting_2s", "random_pool_1s", "sparse_200ms"]
# Reference branches no shipped config reaches: substeps = 2
# (neuron.cpp:183-229), the refractory clamp (neuron.cpp:193-197), STP
# facilitation (network.cpp:237-243), shared noise (neuron.cpp:117-124,
# astrocyte.cpp:247-249), autapses and unique astrocyte outputs
# (network.cpp:154-204).
VARIANTS = ["sparse_sub2", "quick_tref", "quick_fac", "use_case_shared",
            "quick_autapse_unique", "bursting_sub2_tref"]


@pytest.mark.parametrize("case", CASES + VARIANTS)
def test_device_equals_oracle_bitwise(case):
    _announce(case)
    m, g = model_for(case)
    steps = g["total_steps"]
    net, s, n = _run_device(m, steps, chunk=steps // 3 + 7)
    o = oracle.OracleRun(m.desc, oracle.LIBM_AMATH)
    o.step(steps)
    os_, on = o.spikes()
    assert np.array_equal(s, os_) and np.array_equal(n, on)
    for f in STATE_FIELDS:
        assert np.array_equal(net.state(f).view(np.uint64), o.state(f).view(np.uint64)), f
    c = net.counters()
    assert c["steps"] == steps and c["spikes"] == len(s)


@pytest.mark.parametrize("case", CASES + VARIANTS + ["sparse_1s_s2"])
def test_device_equals_reference_raster(case):
    _announce(case)
    m, g = model_for(case)
    _, s, n = _run_device(m, g["total_steps"], chunk=997)
    s, n = recorded(s, n, g["first_recorded_step"])
    assert np.array_equal(s, np.array(g["steps"], np.int64))
    assert np.array_equal(n, np.array(g["neurons"], np.uint32))


@pytest.mark.parametrize("case", ["sparse_10s_s1", "sparse_10s_s2", "sparse_10s_s3"])
def test_sparse_10s_raster_equals_reference_build(case):
    """The 10,000-neuron Sparse network, 1 s transient + 10 s recorded:
    the device raster hashes to the reference build's."""
    if not golden_exists(case):
        pytest.skip("full-length fixture not generated")
    _announce(case)
    m, g = model_for(case)
    _, s, n = _run_device(m, g["total_steps"], chunk=10000)
    s, n = recorded(s, n, g["first_recorded_step"])
    assert len(s) == g["n_spikes"]
    counts = np.bincount((s - g["first_recorded_step"]) // 1000,
                         minlength=len(g["counts_per_1000"]))
    assert counts.tolist() == g["counts_per_1000"]
    assert raster_hash(s, n) == g["raster_sha256"]


@pytest.mark.parametrize("case", ["use_case_10s", "use_case_10s_pre20s"])
def test_use_case_10s_raster_equals_reference_build(case):
    """The use_case network (Gaussian noise on every cell, one Box-Muller
    cosine per cell per noise interval) for 10 s simulated, without and with
    its configured 20 s transient: the device raster hashes to the reference
    build's, so the device cosine equals the C library's on every draw."""
    _announce(case)
    m, g = model_for(case)
    _, s, n = _run_device(m, g["total_steps"], chunk=25000)
    s, n = recorded(s, n, g["first_recorded_step"])
    assert len(s) == g["n_spikes"]
    counts = np.bincount((s - g["first_recorded_step"]) // 1000,
                         minlength=len(g["counts_per_1000"]))
    assert counts.tolist() == g["counts_per_1000"]
    assert raster_hash(s, n) == g["raster_sha256"]


def test_windowed_and_unwindowed_stepping_agree():
    """Chunk boundaries that cut windows (odd step counts) change nothing."""
    _announce("chunking")
    m = A.Model(config_path("quick"), seed=3)
    a, sa, na = _run_device(m, 3000, chunk=3000)
    b, sb, nb = _run_device(m, 3000, chunk=7)
    c, sc, nc = _run_device(m, 300, chunk=1)
    assert np.array_equal(sc, sa[sa < 300]) and np.array_equal(nc, na[sa < 300])
    assert np.array_equal(sa, sb) and np.array_equal(na, nb)
    for f in STATE_FIELDS:
        assert np.array_equal(a.state(f), b.state(f)), f


def test_batch_members_equal_solo_runs():
    """Networks of different size and seed stepped in one batch give the
    same bits as each run alone."""
    _announce("batch")
    specs = [("quick", 1), ("quick", 9), ("use_case", 2), ("random_pool", 3)]
    models = [A.Model(config_path(c), seed=s) for c, s in specs]
    steps = [1500, 2200, 1800, 900]
    batch = A.Batch(8)
    slots = [batch.add(m, t) for m, t in zip(models, steps)]
    batch.step(700)
    batch.step(3000)
    for m, slot, t in zip(models, slots, steps):
        assert batch.net_step(slot) == t
        bs, bn = batch.spikes(slot)
        o = oracle.OracleRun(m.desc, oracle.LIBM_AMATH)
        o.step(t)
        os_, on = o.spikes()
        assert np.array_equal(bs, os_) and np.array_equal(bn, on)
        for f in ("V", "Ca", "IP3", "I_sic"):
            assert np.array_equal(batch.state(slot, f), o.state(f)), f


def test_mixed_precision_batch_with_partial_networks_equals_solo_runs():
    """FP64 and FP32 networks in one batch, among them an FP32 network of
    astrocytes only (no drive, delivery or spike list; its window closes
    every window) and an FP64 network without astrocytes: every member
    gives the bits of its solo run (the FP64 members also the oracle's),
    over calls that end mid-window. Each precision keeps its own delivery
    list and compaction set."""
    _announce("mixed batch with partial networks")
    specs = [("quick", 3, capi.ASG_FP64, {}), ("use_case", 4, capi.ASG_FP32, {}),
             ("quick", 5, capi.ASG_FP32, {"N.N_E": 0, "N.N_I": 0}),
             ("quick", 6, capi.ASG_FP64, {"N.N_A": 0}), ("quick", 7, capi.ASG_FP32, {})]
    models = []
    for cfg, seed, prec, assign in specs:
        m = A.Model(config_path(cfg), seed=seed, assignments=assign or None)
        m.desc.contents.precision = prec
        models.append(m)
    total, chunks = 2400, (333, 1000, 1067)
    batch = A.Batch(8)
    slots = [batch.add(m, total) for m in models]
    for c in chunks:
        batch.step(c)
    drained = batch.spikes_all()
    for (cfg, seed, prec, _), m, slot in zip(specs, models, slots):
        net = A.DeviceNetwork(m)
        net.step(total)
        s1, n1 = net.spikes()
        bs, bn = drained[slot]
        assert np.array_equal(bs, s1) and np.array_equal(bn, n1), (cfg, seed)
        fields = (("V",) if m.n_neurons else ()) + (("Ca", "IP3") if m.n_astro else ())
        for f in fields:
            assert np.array_equal(batch.state(slot, f).view(np.uint64),
                                  net.state(f).view(np.uint64)), (cfg, seed, f)
        net.close()
        if prec == capi.ASG_FP64:
            o = oracle.OracleRun(m.desc, oracle.LIBM_AMATH)
            o.step(total)
            os_, on = o.spikes()
            assert np.array_equal(bs, os_) and np.array_equal(bn, on), (cfg, seed)
    batch.close()


# Delays that are multiples neither of each other nor of the window (exc
# 13, inh 21, a2n 17 steps: windows of 13), the SIC exchanged every third
# step, and a threshold just above the initial Ca so that exchanges turn
# live during the run.
ODD_TIMING = {"synapse.d_e": 1.3, "synapse.d_i": 2.1, "synapse.d_a2n": 1.7,
              "synapse.sic_interval": 3, "astrocyte.SIC_th": 0.076}


def test_odd_delays_and_sic_interval_equal_oracle():
    """Calls of 7 and 23 steps cut windows short at every call, so ring rows,
    arrival cadences and the SIC liveness ring are exercised at every
    alignment; the device still equals the oracle bit for bit."""
    _announce("odd timing")
    m = A.Model(config_path("use_case"), seed=5, sim_time=300.0, pre_time=0.0,
                assignments=ODD_TIMING)
    steps = m.info["sim_steps"]
    net = A.DeviceNetwork(m)
    done, k, parts_s, parts_n = 0, 0, [], []
    while done < steps:
        c = min((7, 23)[k % 2], steps - done)
        net.step(c)
        done += c
        k += 1
        s, n = net.spikes()
        parts_s.append(s)
        parts_n.append(n)
    s, n = np.concatenate(parts_s), np.concatenate(parts_n)
    o = oracle.OracleRun(m.desc, oracle.LIBM_AMATH)
    o.step(steps)
    os_, on = o.spikes()
    assert len(s) > 0 and np.count_nonzero(o.state("I_sic")) > 0
    assert np.array_equal(s, os_) and np.array_equal(n, on)
    for f in STATE_FIELDS:
        assert np.array_equal(net.state(f).view(np.uint64), o.state(f).view(np.uint64)), f


@pytest.mark.parametrize("noise,substeps", [(0.0, 1), (0.003, 1), (0.003, 2)])
def test_astro_dropin_matches_engine_cell_math(noise, substeps):
    """asg_astro_* (the astrocyte_cuda.hpp replacement) against the oracle's
    astrocyte update: an astrocyte-only network driven by its Poisson input
    (and, where set, an independent Gaussian noise current held for 5
    steps), replayed through the drop-in with the same per-step inputs, at
    one and two RK4 substeps per step."""
    _announce(f"astro drop-in noise {noise} substeps {substeps}")
    text = """{ "N": {"N_A": 3000, "N_E": 0, "N_I": 0},
      "simulation": {"substeps": %d},
      "astrocyte": {"IP3": 0.3, "Ca_tot": 1.8958264782990153, "IP3_0": 0.010925583871830699,
                    "tau_IP3": 1058.7589568400845, "delta_IP3": 0.04911592657233997},
      "randomize_astro": {"enabled": true, "lower": 0.9, "upper": 1.05, "var": 0.1},
      "input_astro": {"poiss_rate": 300.0, "poiss_weight": 1.0, "gauss_noise_std": %r,
                      "noise_dt": 0.5} }""" % (substeps, noise)
    m = A.Model(text=text)
    d = m.desc.contents
    n = d.n_astro
    steps = 400
    o = oracle.OracleRun(m.desc, oracle.LIBM_AMATH)
    import ctypes as C
    from astrosimgpu_amd import _capi as capi
    lib = A.lib()
    arr = lambda ptr: np.ctypeslib.as_array(ptr, shape=(n,)).copy()
    Ca = np.full(n, d.astro.Ca_init)
    IP3 = np.full(n, d.astro.IP3_init)
    hh = np.full(n, d.astro.h_init)
    Ca_tot, IP3_0, tau, delta = arr(d.Ca_tot), arr(d.IP3_0), arr(d.tau_IP3), arr(d.delta_IP3)
    P = lambda a: a.ctypes.data_as(capi.DBLP)
    h = lib.asg_astro_create(n, P(Ca), P(IP3), P(hh), P(Ca_tot), P(IP3_0), P(tau), P(delta))
    lam = d.input_astro.poiss_rate * d.dt * 1e-3
    seed = d.seed ^ 0xC2B2AE3D27D4EB4F
    noise_seed = d.seed ^ 0x9E3779B97F4A7C15
    spn = max(1, int(round(d.input_astro.noise_dt / d.dt)))
    for step in range(steps):
        ev = _poisson_counts(seed, step, n, lam)
        ip3_in = d.input_astro.poiss_weight * ev.astype(np.float64)
        ip3_in[ev == 0] = 0.0
        lib.asg_astro_push_input(h, P(ip3_in))
        lib.asg_astro_update(h, C.byref(d.astro.c), d.dt / substeps, substeps, noise, 1, 0.0,
                             noise_seed, step // spn)
    lib.asg_astro_destroy(h, P(Ca), P(IP3), P(hh))
    o.step(steps)
    assert np.array_equal(Ca, o.state("Ca"))
    assert np.array_equal(IP3, o.state("IP3"))
    assert np.array_equal(hh, o.state("h"))


def test_astro_dropin_empty_and_null_handles_are_no_ops():
    """n == 0 gives a handle with no buffers; every call on it, and on a null
    handle, returns without touching its arguments (astrocyte_cuda.cu:90-93,
    108-110, 125-127)."""
    _announce("astro drop-in n = 0")
    import ctypes as C
    lib = A.lib()
    h = lib.asg_astro_create(0, None, None, None, None, None, None, None)
    assert h
    lib.asg_astro_push_input(h, None)
    lib.asg_astro_update(h, None, 0.1, 1, 0.0, 1, 0.0, 0, 0)
    lib.asg_astro_pull_calcium(h, None)
    lib.asg_astro_destroy(h, None, None, None)
    lib.asg_astro_push_input(None, None)
    lib.asg_astro_update(None, None, 0.1, 1, 0.0, 1, 0.0, 0, 0)
    lib.asg_astro_pull_calcium(None, None)
    lib.asg_astro_destroy(None, None, None, None)


def _poisson_counts(seed, step, n, lam):
    """Knuth Poisson of fresh counter generators (reference rng.hpp:107-124),
    vectorised in numpy with the splitmix64 finaliser (rng.hpp:26-33)."""
    M = np.uint64(0xFFFFFFFFFFFFFFFF)
    stream = np.uint64(step) * np.uint64(1000003) + np.arange(n, dtype=np.uint64)
    limit = oracle.math("exp", np.array([-lam]), oracle.LIBM_AMATH)[0]
    counts = np.zeros(n, np.int64)
    prod = None
    alive = np.ones(n, bool)
    c = 0
    with np.errstate(over="ignore"):
        while alive.any():
            z = (np.uint64(seed) + stream * np.uint64(0x9E3779B97F4A7C15)
                 + np.uint64(c) * np.uint64(0xBF58476D1CE4E5B9)) & M
            z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & M
            z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & M
            z = z ^ (z >> np.uint64(31))
            u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
            if prod is None:
                prod = u
            else:
                prod = np.where(alive, prod * u, prod)
            if c > 0:
                counts += alive
            alive = alive & (prod > limit)
            c += 1
    return counts


def test_capped_delivery_masks_equal_oracle():
    """20,000 neurons: full delivery masks would take about 700 MB, so the
    plan caps the chunk count (64 chunks, 4096 spikes per step and type);
    the device still equals the oracle bit for bit."""
    _announce("capped masks 20k")
    m = A.Model(config_path("paper_sparse"), seed=21, sim_time=200.0, pre_time=0.0,
                assignments={"N.N_E": 16000, "N.N_I": 4000, "N.N_A": 10000,
                             "connectivity.p_primary": 0.005,
                             "connectivity.generator": "per_row"})
    assert m.footprint() < 400 << 20
    steps = m.info["sim_steps"]
    net, s, n = _run_device(m, steps, chunk=700)
    o = oracle.OracleRun(m.desc, oracle.LIBM_AMATH)
    o.step(steps)
    os_, on = o.spikes()
    assert len(s) > 0
    assert np.array_equal(s, os_) and np.array_equal(n, on)
    for f in ("V", "w", "I_sic", "Ca", "IP3"):
        assert np.array_equal(net.state(f).view(np.uint64), o.state(f).view(np.uint64)), f
