#!/usr/bin/env python3
"""Writes tests/golden/cli_<case>.json: what the reference simulator's command
line (src/main.cpp, built from its own sources as oracle/_ref/astrosimgpu_ref)
writes for a small run -- the SHA-256, line count and first lines of
spikes.csv, astrocytes.csv and neurons.csv, and the network statistics and
spike count of run.txt. The test runs astrosim-amd with the same arguments
and compares.

Run from the repository root in the build container:
    python tests/golden/make_cli_golden.py
"""

from __future__ import annotations

import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(REPO, "oracle", "_ref", "astrosimgpu_ref")

# name: (base config, edits to its "simulation" block, seed, sim ms, pre ms)
CASES = {
    "quick_neurons": ("quick", {"record_neuron": True, "record_neuron_max": 50}, 2, 300.0, 100.0),
    "use_case_noise": ("use_case", {"record_every": 20, "record_astro_max": 30}, 1, 400.0, 200.0),
}

STAT_LABELS = {
    "astrocytes": "n_astro", "excitatory neurons": "n_exc", "inhibitory neurons": "n_inh",
    "exc -> neuron synapses": "n_primary_exc", "inh -> neuron synapses": "n_primary_inh",
    "neuron -> astrocyte": "n_neuron_to_astro", "astrocyte -> neuron": "n_astro_to_neuron",
    "spikes recorded": "spikes_recorded",
}


def strip_comments(text: str) -> str:
    return "\n".join(line for line in text.splitlines() if not line.lstrip().startswith("//"))


def case_config(case: str) -> dict:
    base, edits, _, _, _ = CASES[case]
    with open(os.path.join(HERE, "configs", base + ".json")) as f:
        cfg = json.loads(strip_comments(f.read()))
    cfg["simulation"].update(edits)
    return cfg


def run_stats(text: str) -> dict:
    out = {}
    for line in text.splitlines():
        for lab, key in STAT_LABELS.items():
            m = re.match(rf"^{re.escape(lab)}\s+(\d+)\s*$", line)
            if m:
                out[key] = int(m.group(1))
        m = re.match(r"^mean firing rate\s+([0-9.]+) Hz", line)
        if m:
            out["mean_rate_hz"] = float(m.group(1))
    return out


def describe_outputs(out_dir: str) -> dict:
    files = {}
    for name in ("spikes.csv", "astrocytes.csv", "neurons.csv"):
        path = os.path.join(out_dir, name)
        if not os.path.exists(path):
            continue
        data = open(path, "rb").read()
        lines = data.decode().splitlines()
        files[name] = {"sha256": hashlib.sha256(data).hexdigest(), "bytes": len(data),
                       "lines": len(lines), "head": lines[:4]}
    with open(os.path.join(out_dir, "run.txt")) as f:
        stats = run_stats(f.read())
    return {"files": files, "stats": stats}


def main():
    if not os.path.exists(REF):
        sys.exit(f"{REF} is missing: make -C oracle ref")
    for case, (_, _, seed, sim, pre) in CASES.items():
        with tempfile.TemporaryDirectory() as tmp:
            cfg_path = os.path.join(tmp, "cfg.json")
            with open(cfg_path, "w") as f:
                json.dump(case_config(case), f)
            out = os.path.join(tmp, "out")
            subprocess.run([REF, "-c", cfg_path, "-s", str(seed), "-t", str(sim), "-p", str(pre),
                            "-o", out, "--no-analysis"], check=True, stdout=subprocess.DEVNULL,
                           env=dict(os.environ, OMP_NUM_THREADS="8"))
            d = describe_outputs(out)
        d.update({"seed": seed, "sim_time": sim, "pre_time": pre,
                  "producer": "reference build (oracle/_ref), --no-analysis"})
        with open(os.path.join(HERE, f"cli_{case}.json"), "w") as f:
            json.dump(d, f, indent=1)
        print(case, {k: (v["lines"], v["sha256"][:12]) for k, v in d["files"].items()}, d["stats"])


if __name__ == "__main__":
    main()
