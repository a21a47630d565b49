"""Write the golden fixtures under tests/golden/ from the reference build.

Run in the build container only (it needs /root/reference and the reference
binary built by `make -C oracle ref`):

    python tests/golden/make_golden.py [--full]

Outputs (data only):
  configs/<name>.json   the reference's parameter files with comments removed
                        (their values, as inputs for the tests)
  ref_<case>.json       for each case: the reference binary's network stats
                        (run.txt) and its spike raster as (step, neuron)
                        pairs -- the full list for short cases, a SHA-256 of
                        the little-endian int64 step / uint32 neuron arrays and
                        per-1000-step counts for long ones -- plus the first
                        recorded astrocyte Ca/IP3 CSV rows where recorded.

--full also runs the 10 s Sparse network for seeds 1, 2 and 3 (about a
minute each on 8 cores).

Variant cases exercise the reference's branches that no shipped config
reaches. Each writes configs/<case>.json = a shipped config's values with a
few keys overridden (listed in VARIANTS), and runs the reference on that file:
substeps = 2 (neuron.cpp:183-229), a refractory period (neuron.cpp:193-197),
STP facilitation (network.cpp:237-243), shared noise for neurons and
astrocytes (neuron.cpp:117-124, astrocyte.cpp:247-249), autapses and one
astrocyte output per pair (network.cpp:154-204).
"""

from __future__ import annotations

import argparse
import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_ROOT = "/root/reference"
REF_BIN = os.path.join(REPO, "oracle", "_ref", "astrosimgpu_ref")

CONFIGS = ["quick", "use_case", "bursting", "no_spiking", "random_pool", "paper_sparse",
           "kernel_scaling", "scale_1000", "scale_1000_saturated"]

# name -> (config, seed, sim_time ms, pre_time ms, keep full raster)
SHORT_CASES = {
    "quick_s1": ("quick", 1, 2500.0, 0.0, True),
    "quick_s7": ("quick", 7, 1000.0, 0.0, True),
    "use_case_2s": ("use_case", 1, 2000.0, 0.0, True),
    "bursting_2s": ("bursting", 1, 2000.0, 0.0, True),
    "random_pool_1s": ("random_pool", 1, 1000.0, 0.0, True),
    "sparse_200ms": ("paper_sparse", 1, 200.0, 0.0, True),
    "sparse_1s_s2": ("paper_sparse", 2, 1000.0, 0.0, True),
}
# case -> (base config, {"section.key": value}, seed, sim ms, pre ms, full raster)
VARIANTS = {
    "sparse_sub2": ("paper_sparse", {"simulation.substeps": 2}, 1, 300.0, 0.0, True),
    "quick_tref": ("quick", {"neuron_exc.t_ref": 2.0, "neuron_inh.t_ref": 1.0},
                   1, 1500.0, 0.0, True),
    "quick_fac": ("quick", {"synapse.stp.tau_fac": 600.0, "synapse.stp.U": 0.2},
                  1, 1500.0, 0.0, True),
    "use_case_shared": ("use_case", {"input_exc.independent_noise": False,
                                     "input_inh.independent_noise": False,
                                     "input_astro.independent_noise": False},
                        1, 2000.0, 0.0, True),
    "quick_autapse_unique": ("quick", {"connectivity.allow_autapses": True,
                                       "connectivity.unique_third_out": True,
                                       "connectivity.p_third_if_primary": 0.5},
                             3, 1500.0, 0.0, True),
    "bursting_sub2_tref": ("bursting", {"simulation.substeps": 2, "neuron_exc.t_ref": 1.5,
                                        "synapse.stp.tau_fac": 300.0},
                           2, 1000.0, 0.0, True),
}
# Long noise runs: the use_case network (Gaussian noise on every cell) for
# 10 s simulated with no transient, and as configured (20 s transient).
LONG_CASES = {
    "use_case_10s": ("use_case", 1, 10000.0, 0.0, False),
    "use_case_10s_pre20s": ("use_case", 1, 10000.0, 20000.0, False),
}
FULL_CASES = {
    "sparse_10s_s1": ("paper_sparse", 1, 10000.0, 1000.0, False),
    "sparse_10s_s2": ("paper_sparse", 2, 10000.0, 1000.0, False),
    "sparse_10s_s3": ("paper_sparse", 3, 10000.0, 1000.0, False),
}


def strip_comments(text: str) -> str:
    out = []
    for line in text.splitlines():
        # The configs hold no "//" inside strings, so a plain cut is exact.
        idx = line.find("//")
        out.append(line if idx < 0 else line[:idx])
    return "\n".join(out)


def raster_hash(steps: np.ndarray, neurons: np.ndarray) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(steps, "<i8").tobytes())
    h.update(np.ascontiguousarray(neurons, "<u4").tobytes())
    return h.hexdigest()


def parse_run_txt(path: str) -> dict:
    labels = {
        "astrocytes": "n_astro", "excitatory neurons": "n_exc", "inhibitory neurons": "n_inh",
        "exc -> neuron synapses": "n_primary_exc", "inh -> neuron synapses": "n_primary_inh",
        "neuron -> astrocyte": "n_neuron_to_astro", "astrocyte -> neuron": "n_astro_to_neuron",
        "spikes recorded": "spikes_recorded",
    }
    out = {}
    with open(path) as f:
        for line in f:
            for lab, key in labels.items():
                m = re.match(rf"^{re.escape(lab)}\s+(\d+)\s*$", line)
                if m:
                    out[key] = int(m.group(1))
            m = re.match(r"^mean firing rate\s+([0-9.]+) Hz", line)
            if m:
                out["mean_rate_hz"] = float(m.group(1))
    return out


def write_variant(name, base, overrides):
    with open(os.path.join(HERE, "configs", base + ".json")) as f:
        values = json.load(f)
    for dotted, v in overrides.items():
        node = values
        keys = dotted.split(".")
        for k in keys[:-1]:
            node = node.setdefault(k, {})
        node[keys[-1]] = v
    with open(os.path.join(HERE, "configs", name + ".json"), "w") as f:
        json.dump(values, f, indent=1)


def run_case(name, cfg, seed, sim, pre, full_raster, threads):
    cfg_path = os.path.join(REF_ROOT, "config", cfg + ".json")
    if not os.path.exists(cfg_path):
        cfg_path = os.path.join(HERE, "configs", cfg + ".json")
    with tempfile.TemporaryDirectory() as out:
        cmd = [REF_BIN, "-c", cfg_path, "-s", str(seed), "-t", str(sim), "-p", str(pre),
               "-o", out, "--no-analysis"]
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        subprocess.run(cmd, check=True, env=env, stdout=subprocess.DEVNULL)
        stats = parse_run_txt(os.path.join(out, "run.txt"))
        spikes_path = os.path.join(out, "spikes.csv")
        data = np.loadtxt(spikes_path, delimiter=",", skiprows=1, ndmin=2)
        dt = 0.1
        steps = np.rint(data[:, 0] / dt).astype(np.int64) if len(data) else np.zeros(0, np.int64)
        neurons = data[:, 1].astype(np.uint32) if len(data) else np.zeros(0, np.uint32)
        astro_rows = []
        apath = os.path.join(out, "astrocytes.csv")
        if os.path.exists(apath):
            with open(apath) as f:
                next(f)
                for i, line in enumerate(f):
                    if i >= 400:
                        break
                    astro_rows.append(line.rstrip("\n"))
    pre_steps = int(pre / dt + 0.5)
    sim_steps = int(sim / dt + 0.5)
    rec = {
        "config": cfg, "seed": seed, "sim_time": sim, "pre_time": pre, "dt": dt,
        "first_recorded_step": pre_steps, "total_steps": pre_steps + sim_steps,
        "stats": stats, "n_spikes": int(len(steps)),
        "raster_sha256": raster_hash(steps, neurons),
        "counts_per_1000": np.bincount((steps - pre_steps) // 1000,
                                       minlength=(sim_steps + 999) // 1000).tolist(),
        "astro_csv_head": astro_rows,
        "producer": "reference build (oracle/_ref, g++ -O2 -ffp-contract=off), glibc libm",
    }
    if full_raster:
        rec["steps"] = steps.tolist()
        rec["neurons"] = neurons.tolist()
    with open(os.path.join(HERE, f"ref_{name}.json"), "w") as f:
        json.dump(rec, f, separators=(",", ":"))
    print(f"{name}: {len(steps)} spikes, stats {stats}", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--full", action="store_true")
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--only", nargs="*", help="run only these cases")
    args = ap.parse_args()
    if not os.path.exists(REF_BIN):
        sys.exit(f"{REF_BIN} missing: run `make -C oracle ref` first")
    os.makedirs(os.path.join(HERE, "configs"), exist_ok=True)
    for c in CONFIGS:
        with open(os.path.join(REF_ROOT, "config", c + ".json")) as f:
            values = json.loads(strip_comments(f.read()))
        with open(os.path.join(HERE, "configs", c + ".json"), "w") as f:
            json.dump(values, f, indent=1)
    cases = dict(SHORT_CASES)
    cases.update(LONG_CASES)
    for name, (base, overrides, seed, sim, pre, full) in VARIANTS.items():
        write_variant(name, base, overrides)
        cases[name] = (name, seed, sim, pre, full)
    if args.full:
        cases.update(FULL_CASES)
    for name, (cfg, seed, sim, pre, full) in cases.items():
        if args.only and name not in args.only:
            continue
        run_case(name, cfg, seed, sim, pre, full, args.threads)


if __name__ == "__main__":
    main()
