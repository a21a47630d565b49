"""The CPU oracle against the reference build (golden rasters) and against
closed forms. These pin the oracle before it is used to check the device."""

import numpy as np
import pytest

from _cases import model_for, recorded, stats_match
from conftest import config_path, golden_exists

import astrosimgpu_amd as A
import oracle

SHORT = ["quick_s1", "quick_s7", "use_case_2s", "bursting_2s", "random_pool_1s", "sparse_200ms"]
# Branches no shipped config reaches: substeps = 2, refractory clamp, STP
# facilitation, shared noise, autapses + unique astrocyte outputs.
VARIANTS = ["sparse_sub2", "quick_tref", "quick_fac", "use_case_shared",
            "quick_autapse_unique", "bursting_sub2_tref"]


@pytest.mark.parametrize("case", SHORT + VARIANTS)
def test_network_stats_equal_reference(case):
    m, g = model_for(case)
    assert stats_match(m, g), (m.stats(), g["stats"])


@pytest.mark.parametrize("case", SHORT + VARIANTS)
def test_oracle_with_c_library_equals_reference_raster(case):
    """libm = the host C library, as the reference build: bit-identical raster."""
    m, g = model_for(case)
    o = oracle.OracleRun(m.desc, oracle.LIBM_GLIBC)
    o.step(g["total_steps"])
    s, n = recorded(*o.spikes(), g["first_recorded_step"])
    assert len(s) == g["n_spikes"]
    assert np.array_equal(s, np.array(g["steps"], np.int64))
    assert np.array_equal(n, np.array(g["neurons"], np.uint32))


@pytest.mark.parametrize("case", SHORT + VARIANTS)
def test_oracle_with_amath_equals_reference_raster(case):
    """libm = amath (what the device runs): exp, log and cos are the C
    library's bit for bit, so the rasters are the reference build's."""
    m, g = model_for(case)
    o = oracle.OracleRun(m.desc, oracle.LIBM_AMATH)
    o.step(g["total_steps"])
    s, n = recorded(*o.spikes(), g["first_recorded_step"])
    assert np.array_equal(s, np.array(g["steps"], np.int64))
    assert np.array_equal(n, np.array(g["neurons"], np.uint32))


@pytest.mark.parametrize("case", ["use_case_10s", "use_case_10s_pre20s"])
def test_oracle_amath_use_case_10s_hash_equals_reference(case):
    """The use_case network (Gaussian noise current on every neuron and
    astrocyte, a Box-Muller cosine per cell per 1 ms) for 10 s, with no
    transient and with the configured 20 s transient (300,000 steps): the
    amath oracle's raster hashes to the reference build's."""
    from _cases import raster_hash
    m, g = model_for(case)
    o = oracle.OracleRun(m.desc, oracle.LIBM_AMATH)
    o.step(g["total_steps"])
    s, n = recorded(*o.spikes(), g["first_recorded_step"])
    assert len(s) == g["n_spikes"]
    assert raster_hash(s, n) == g["raster_sha256"]


def test_oracle_sparse_1s_seed2_equals_reference():
    m, g = model_for("sparse_1s_s2")
    o = oracle.OracleRun(m.desc, oracle.LIBM_AMATH)
    o.step(g["total_steps"])
    s, n = recorded(*o.spikes(), g["first_recorded_step"])
    assert np.array_equal(s, np.array(g["steps"], np.int64))
    assert np.array_equal(n, np.array(g["neurons"], np.uint32))


def _single_astro_text(ip3_init, ip3_0, tau):
    return f"""{{ "N": {{"N_A": 1, "N_E": 0, "N_I": 0}},
      "astrocyte": {{"IP3": {ip3_init}, "IP3_0": {ip3_0}, "tau_IP3": {tau}}} }}"""


def test_ip3_relaxation_is_the_rk4_polynomial():
    """Undriven IP3 obeys dIP3/dt = (IP3_0 - IP3)/tau, linear and independent
    of Ca and h, so classic RK4 multiplies the excess by exactly
    R = 1 + z + z^2/2 + z^3/6 + z^4/24 per step, z = -h/tau."""
    ip3_0, tau, h, steps = 0.16, 1000.0, 0.1, 2000
    m = A.Model(text=_single_astro_text(0.9, ip3_0, tau))
    o = oracle.OracleRun(m.desc, oracle.LIBM_AMATH)
    o.step(steps)
    z = -h / tau
    R = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
    expected = ip3_0 + (0.9 - ip3_0) * R ** steps
    got = o.state("IP3")[0]
    assert abs(got - expected) < 1e-12
    # and RK4 tracks the exact exponential decay to O(h^4)
    assert abs(got - (ip3_0 + (0.9 - ip3_0) * np.exp(-steps * h / tau))) < 1e-12


def test_astrocyte_rest_state_stays_in_bounds():
    m = A.Model(text=_single_astro_text(0.16, 0.16, 7142.0))
    o = oracle.OracleRun(m.desc, oracle.LIBM_AMATH)
    o.step(20000)
    ca, h3 = o.state("Ca")[0], o.state("h")[0]
    assert 0.0 <= ca <= 2.0 and 0.0 <= h3 <= 1.0
    assert abs(o.state("IP3")[0] - 0.16) < 1e-12


def _single_neuron_text(I_e, tau_syn=2.0):
    return f"""{{ "N": {{"N_A": 0, "N_E": 1, "N_I": 0}},
      "simulation": {{"substeps": 4}},
      "neuron_exc": {{"C_m": 130.0, "g_L": 18.0, "E_L": -58.0, "V_th": -50.0,
        "V_reset": -50.0, "b": 300.0, "a": 4.0, "tau_w": 450.0, "V_m": -58.0,
        "tau_syn_ex": {tau_syn}, "tau_syn_in": {tau_syn}, "I_e": {I_e}}} }}"""


def test_neuron_rest_and_drive():
    """tests/test_astrosimgpu.cpp:222-270 of the reference, on the oracle."""
    m = A.Model(text=_single_neuron_text(0.0))
    o = oracle.OracleRun(m.desc, oracle.LIBM_AMATH)
    o.step(5000)
    assert len(o.spikes()[0]) == 0
    assert abs(o.state("V")[0] - (-58.0)) < 1.0
    m2 = A.Model(text=_single_neuron_text(700.0))
    o2 = oracle.OracleRun(m2.desc, oracle.LIBM_AMATH)
    o2.step(20000)
    s, n = o2.spikes()
    assert len(s) > 0 and np.all(n == 0)


def _draw_grid(seed, n=1 << 22):
    """u on the 53-bit uniform grid of the counter draws (rng.hpp:36-39)."""
    rng = np.random.default_rng(seed)
    return (rng.integers(0, 1 << 53, n, dtype=np.int64).astype(np.float64)
            * (1.0 / 9007199254740992.0)), rng


@pytest.mark.parametrize("fn", ["log", "exp"])
def test_amath_equals_c_library_on_the_draw_grid(fn):
    """The step's own libm (amath, shared by host and device) against the
    host C library on the inputs the step gives it, bit for bit on 4M
    values: log(u) of Box-Muller (rng.hpp:48) for u on the draw grid, exp
    over the AdEx, alpha-decay, STP and Poisson-limit arguments."""
    u, rng = _draw_grid({"log": 2, "exp": 3}[fn])
    x = np.where(u <= 1e-300, 1e-300, u) if fn == "log" else rng.uniform(-60.0, 30.0, u.size)
    a = oracle.math(fn, x, oracle.LIBM_AMATH)
    g = oracle.math(fn, x, oracle.LIBM_GLIBC)
    assert np.array_equal(a.view(np.uint64), g.view(np.uint64)), int(np.sum(a != g))


@pytest.mark.parametrize("fn,lo,hi", [
    ("cos", -1e-7, 1e-7), ("cos", -1.0, 1.0), ("cos", -7.0, 7.0), ("cos", -1e5, 1e5),
    ("cos", -1e8, 1e8), ("sin", -1e-7, 1e-7), ("sin", -1.0, 1.0), ("sin", -7.0, 7.0),
    ("sin", -1e5, 1e5),
])
def test_amath_sin_cos_equal_c_library(fn, lo, hi):
    """amath's sin/cos (glibc's table-driven method restated) against the
    host C library, bit for bit on 1M arguments per range, every branch of
    the method included (first-order, table, cofunction, reduction)."""
    rng = np.random.default_rng({"cos": 11, "sin": 12}[fn] + int(abs(lo)) % 97)
    x = rng.uniform(lo, hi, 1 << 20)
    a = oracle.math(fn, x, oracle.LIBM_AMATH)
    g = oracle.math(fn, x, oracle.LIBM_GLIBC)
    assert np.array_equal(a.view(np.uint64), g.view(np.uint64)), int(np.sum(a != g))


def test_amath_cos_equals_c_library_on_box_muller_angles():
    """cos(2 pi u) of Box-Muller (rng.hpp:49), u on the 53-bit draw grid:
    equal to the C library on all of 4M draws. A plain mistake in the
    restatement (a dropped table low part, an unfused product, a wrong
    polynomial coefficient) breaks equality on a fraction of the draws."""
    u, _ = _draw_grid(1)
    x = 6.283185307179586476925286766559 * u
    a = oracle.math("cos", x, oracle.LIBM_AMATH)
    g = oracle.math("cos", x, oracle.LIBM_GLIBC)
    assert np.array_equal(a.view(np.uint64), g.view(np.uint64)), int(np.sum(a != g))
    # Closed forms the table fixes: cos(k/128) for table points is the
    # table's rounded value; cos(0) = 1, cos(pi) = -1.
    assert oracle.math("cos", np.array([0.0, np.pi]), oracle.LIBM_AMATH).tolist() == [1.0, -1.0]
