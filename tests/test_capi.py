"""The C ABI library loads, exports what include/asg.h declares, and the host
side (config loading, network construction, argument validation) behaves as
the reference's host code does. No device calls."""

import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import REPO, config_path

import astrosimgpu_amd as A
from astrosimgpu_amd import _capi as capi


def declared_functions():
    with open(os.path.join(REPO, "include", "asg.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(asg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    names = declared_functions()
    assert len(names) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", A.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (asg_[a-z0-9_]+)$", out, flags=re.M))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    lib = A.lib()
    for n in names:
        assert hasattr(lib, n)


def test_prototype_table_covers_header():
    assert set(declared_functions()) <= set(capi.PROTOTYPES)


def test_struct_layout_matches_c():
    # sizes computed by the C compiler for the same declarations
    assert C.sizeof(capi.NeuronParams) == 17 * 8
    assert C.sizeof(capi.InputParams) == 40
    assert C.sizeof(capi.Csr) == 40
    assert C.sizeof(capi.Stp) == 32


def test_use_case_stats_match_published_counts():
    # docs/gpu-port.md:13-15 of the reference: seed 1 synapse counts
    m = A.Model(config_path("use_case"))
    s = m.stats()
    assert (s["n_primary_exc"], s["n_primary_inh"], s["n_neuron_to_astro"]) == (39790, 9944, 8097)
    assert s["n_astro_to_neuron"] == s["n_neuron_to_astro"]
    assert m.info["ring_slots"] == 11
    assert m.info["pre_steps"] == 200000 and m.info["sim_steps"] == 3000000


def test_same_seed_same_network_and_different_seed_differs():
    a = A.Model(config_path("quick"), seed=7)
    b = A.Model(config_path("quick"), seed=7)
    c = A.Model(config_path("quick"), seed=8)
    for name in ("exc", "inh", "n2a", "a2n"):
        ra, ta, _, _ = a.csr(name)
        rb, tb, _, _ = b.csr(name)
        assert np.array_equal(ra, rb) and np.array_equal(ta, tb)
    assert a.stats() != c.stats()


def test_unique_third_out_collapses_pairs():
    base = A.Model(config_path("quick"))
    uniq = A.Model(config_path("quick"), assignments={"connectivity.unique_third_out": True})
    assert uniq.stats()["n_astro_to_neuron"] < base.stats()["n_astro_to_neuron"]
    assert uniq.stats()["n_neuron_to_astro"] == base.stats()["n_neuron_to_astro"]


def test_overrides_apply_after_file():
    m = A.Model(config_path("quick"), seed=5, sim_time=123.0, pre_time=4.0, n_astro=30)
    assert m.info["seed"] == 5 and m.info["sim_time"] == 123.0 and m.info["pre_sim_time"] == 4.0
    assert m.info["n_astro"] == 30


def _text(**edits):
    import json
    with open(config_path("quick")) as f:
        cfg = json.load(f)
    for path, value in edits.items():
        node = cfg
        keys = path.split("__")
        for k in keys[:-1]:
            node = node.setdefault(k, {})
        node[keys[-1]] = value
    return json.dumps(cfg)


@pytest.mark.parametrize("edit,message", [
    ({"simulation__bogus": 1}, "unrecognised keys"),
    ({"simulation__substeps": 0}, "substeps must be at least 1"),
    ({"simulation__dt": 0}, "dt must be positive"),
    ({"simulation__record_every": 0}, "record_every must be at least 1"),
    ({"synapse__sic_interval": 0}, "sic_interval must be at least 1"),
    ({"connectivity__pool_size": 0}, "pool_size must be at least 1"),
    ({"connectivity__pool_type": "ring"}, "unknown pool_type"),
])
def test_config_errors_are_reported(edit, message):
    with pytest.raises(A.AsgError) as e:
        A.Model(text=_text(**edit))
    assert e.value.status == capi.ASG_ERR_CONFIG
    assert message in str(e.value)


def test_malformed_json_and_missing_file():
    with pytest.raises(A.AsgError) as e:
        A.Model(text="{ oops }")
    assert "JSON parse error" in str(e.value)
    with pytest.raises(A.AsgError) as e:
        A.Model("/nonexistent/config.json")
    assert "cannot open configuration file" in str(e.value)


def test_comments_and_aliases():
    text = """// leading comment
    { "N": {"N_A": 4, "N_E": 8, "N_I": 2}, // trailing
      "astrocyte": {"IP3": 0.3, "Ca_init": 0.1},
      "input_exc": {"gauss_noise_var": 7.0},
      "input_inh": {"gauss_noise_var": 7.0, "gauss_noise_std": 3.0} }"""
    m = A.Model(text=text)
    d = m.desc.contents
    assert d.astro.IP3_init == 0.3
    assert d.input_exc.gauss_noise_std == 7.0
    assert d.input_inh.gauss_noise_std == 3.0
    # noise_dt <= 0 resolves to 10 * dt (parameters.cpp:169-173)
    assert d.input_exc.noise_dt == 10.0 * d.dt


def test_invalid_descriptions_rejected_before_the_device():
    m = A.Model(config_path("quick"))
    d = capi.NetDesc.from_buffer_copy(m.desc.contents)
    d.substeps = 0
    h = C.c_void_p()
    st = A.lib().asg_net_create(C.byref(d), C.byref(h))
    assert st == capi.ASG_ERR_INVALID
    assert b"substeps" in A.lib().asg_last_error()
    d = capi.NetDesc.from_buffer_copy(m.desc.contents)
    d.exc.delay_steps = 0
    assert A.lib().asg_net_create(C.byref(d), C.byref(h)) == capi.ASG_ERR_INVALID
    assert A.lib().asg_net_create(None, C.byref(h)) == capi.ASG_ERR_INVALID


def test_fp32_limits_rejected_before_the_device():
    """The FP32 path refuses what it does not run: astrocyte Poisson input
    above 30 events per step (the reference's rounded-normal regime, which
    the FP64 path runs) and arrays beyond its 32-bit byte offsets."""
    m = A.Model(config_path("quick"))
    h = C.c_void_p()
    d = capi.NetDesc.from_buffer_copy(m.desc.contents)
    d.precision = capi.ASG_FP32
    d.input_astro.poiss_rate = 30.0 / (d.dt * 1e-3) * 1.01
    assert A.lib().asg_net_create(C.byref(d), C.byref(h)) == capi.ASG_ERR_INVALID
    assert b"30 events per step" in A.lib().asg_last_error()
    d = capi.NetDesc.from_buffer_copy(m.desc.contents)
    d.precision = capi.ASG_FP32
    d.exc.delay_steps = 1 << 22  # rings of 4M rows: beyond 4 GiB
    assert A.lib().asg_net_create(C.byref(d), C.byref(h)) == capi.ASG_ERR_INVALID
    assert b"4 GiB" in A.lib().asg_last_error()


def test_prepare_runs_without_a_device_and_validates():
    """asg_net_prepare (the host half of batch admission) needs no device
    and refuses what asg_batch_add refuses."""
    m = A.Model(config_path("quick"))
    h = C.c_void_p()
    assert A.lib().asg_net_prepare(m.desc, C.byref(h)) == capi.ASG_OK and h.value
    A.lib().asg_prepared_free(h)
    m.prepare()
    m.prepare()  # again: the earlier handle is freed
    m.release_prepared()
    d = capi.NetDesc.from_buffer_copy(m.desc.contents)
    d.substeps = 0
    assert A.lib().asg_net_prepare(C.byref(d), C.byref(h)) == capi.ASG_ERR_INVALID
    assert b"substeps" in A.lib().asg_last_error()


def test_footprint_of_sparse_network_fits_one_gpu():
    m = A.Model(config_path("paper_sparse"))
    fp = m.footprint()
    assert 100e6 < fp < 2e9


def test_astro_scaling_bytes_formula():
    """Per cell: Ca, IP3, h read + written, int32 count, 4 parameters when
    randomized."""
    from astrosimgpu_amd._capi import lib
    assert lib().asg_astro_scaling_bytes(1000, 0, 0) == 1000 * (6 * 8 + 4)
    assert lib().asg_astro_scaling_bytes(1000, 1, 0) == 1000 * (6 * 4 + 4)
    assert lib().asg_astro_scaling_bytes(1000, 0, 1) == 1000 * (10 * 8 + 4)
    from astrosimgpu_amd import astro_bench as AB
    c, p = AB.setup(randomized=True)
    assert p.dt == 0.1 and p.lambda_ == 0.05 and p.randomized == 1
    assert (p.lower, p.upper, p.var) == (0.9, 1.05, 0.1)
    assert c.rate_SERCA == 0.0009


def test_cli_usage_and_errors_without_a_device(tmp_path):
    """astrosim-amd: -h prints the reference's options; an unknown option is
    exit 2 with the usage; a bad configuration is exit 1 with the loader's
    message -- all before any device call."""
    import subprocess
    cli = os.path.join(REPO, "astrosimgpu_amd", "astrosim-amd")
    r = subprocess.run([cli, "-h"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    for opt in ("--config", "--output", "--seed", "--sim-time", "--pre-time", "--astrocytes",
                "--threads", "--no-analysis"):
        assert opt in r.stdout
    r = subprocess.run([cli, "--bogus"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "unknown option: --bogus" in r.stderr
    r = subprocess.run([cli, "-c", str(tmp_path / "missing.json")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "cannot open configuration file" in r.stderr
    bad = tmp_path / "bad.json"
    bad.write_text('{"N": {"N_E": 10, "N_I": 2, "N_A": 3}, "simulation": {"nope": 1}}')
    r = subprocess.run([cli, "-c", str(bad)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "error:" in r.stderr


def test_split_spikes_reassembles_each_slot_across_drain_rounds():
    """Batch.split_spikes (host side of asg_batch_spikes_all): pairs written
    back to back in slot order, over rounds of a small capacity that cut a
    slot in the middle, come back per occupied slot in order; empty slots
    and slots with no pairs give empty arrays."""
    rng = np.random.default_rng(3)
    occupied = [0, 2, 3, 6]
    want = {s: (np.sort(rng.integers(0, 1000, n)).astype(np.int64),
                rng.integers(0, 500, n).astype(np.uint32))
            for s, n in zip(occupied, (5, 0, 9, 4))}
    flat_s = np.concatenate([want[s][0] for s in occupied])
    flat_n = np.concatenate([want[s][1] for s in occupied])
    owner = np.concatenate([np.full(len(want[s][0]), s) for s in occupied])
    cap, rounds = 7, []
    for at in range(0, len(flat_s) + 1, cap):  # the C side's rounds of `cap` pairs
        piece = slice(at, min(at + cap, len(flat_s)))
        counts = np.bincount(owner[piece], minlength=8).astype(np.int64)
        ends = np.cumsum(counts)
        rounds.append(((ends - counts).tolist(), ends.tolist(), flat_s[piece], flat_n[piece]))
    fake = type("FakeBatch", (), {"models": {s: None for s in occupied}})()
    got = A.Batch.split_spikes(fake, rounds)
    assert sorted(got) == occupied
    for s in occupied:
        assert np.array_equal(got[s][0], want[s][0]) and np.array_equal(got[s][1], want[s][1])


@pytest.mark.parametrize("text,where", [
    ('{"a": }', "byte 6"), ('{"a" 1}', "byte 5"), ('{"a": 1,}', "byte 8"),
    ('[1, 2', "byte 5"), ('{"N": {}} x', "byte 10"), ('"\\q"', "byte 1"),
    ('{"a": tru}', "byte 6"), ('{"a": "\\u12"}', "byte 11"), ("", "byte 0"),
    ('{"a": "open', "byte 6"),
])
def test_malformed_documents_name_the_byte(text, where):
    with pytest.raises(A.AsgError) as e:
        A.Model(text=text)
    assert e.value.status == capi.ASG_ERR_CONFIG
    assert "JSON parse error at " + where in str(e.value), str(e.value)


def test_document_rules_of_the_reference_loader():
    """A later duplicate key wins (json.cpp builds a std::map); a comment may
    end the text without a newline; numbers take strtod's longest prefix of
    the sign/digit/point/exponent run ("12e" is 12); unread keys are listed
    at every level in byte order (parameters.cpp:191-202)."""
    m = A.Model(text='{"N": {"N_A": 4, "N_E": 8, "N_I": 2, "N_A": 6},\n'
                     ' "simulation": {"seed": 12e}} // end')
    assert m.info["n_astro"] == 6 and m.info["seed"] == 12
    with pytest.raises(A.AsgError) as e:
        A.Model(text='{"N": {"N_A": 4, "zz": 1, "N_E": 8, "N_I": 2}, "bogus": {"x": 1},'
                     ' "simulation": {"b": 1, "a": 2}}')
    msg = str(e.value)
    assert msg.index("N.zz") < msg.index("bogus") < msg.index("simulation.a") < msg.index("simulation.b")
    assert "bogus.x" not in msg


def test_assignments_create_and_replace_members():
    """Overrides "a.b=literal" parse the literal as JSON, create missing
    objects on the path and replace what was there."""
    m = A.Model(config_path("quick"), assignments={"N.N_A": 7, "synapse.stp.tau_fac": 250.0,
                                                   "connectivity.pool_type": "random"})
    assert m.info["n_astro"] == 7
    assert m.desc.contents.stp.tau_fac == 250.0
    text = '{"N": {"N_A": 4, "N_E": 8, "N_I": 2}, "input_exc": {"poiss_rate": 5.0}}'
    m = A.Model(text=text, assignments={"input_exc.independent_noise": False})
    assert m.desc.contents.input_exc.independent_noise == 0
