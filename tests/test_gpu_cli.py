"""astrosim-amd (the command-line simulator on the device step) writes the
same files as the reference build's command line for the same configuration,
seed and durations: spikes.csv, astrocytes.csv and neurons.csv byte for byte
(FP64), and the same statistics in run.txt."""

import hashlib
import json
import os
import subprocess
import sys

import pytest

from conftest import GOLDEN, REPO

sys.path.insert(0, GOLDEN)
import make_cli_golden as M  # noqa: E402

pytestmark = pytest.mark.gpu

CLI = os.path.join(REPO, "astrosimgpu_amd", "astrosim-amd")


@pytest.mark.parametrize("case", sorted(M.CASES))
def test_cli_files_equal_reference_build(case, tmp_path):
    print(f"[gpu case] cli {case}", flush=True)
    with open(os.path.join(GOLDEN, f"cli_{case}.json")) as f:
        want = json.load(f)
    cfg = tmp_path / "cfg.json"
    cfg.write_text(json.dumps(M.case_config(case)))
    out = tmp_path / "out"
    r = subprocess.run([CLI, "-c", str(cfg), "-s", str(want["seed"]), "-t", str(want["sim_time"]),
                        "-p", str(want["pre_time"]), "-o", str(out), "--no-analysis"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    got = M.describe_outputs(str(out))
    assert sorted(got["files"]) == sorted(want["files"])
    for name, w in want["files"].items():
        g = got["files"][name]
        assert g["head"] == w["head"], name
        assert g["lines"] == w["lines"], name
        assert g["sha256"] == w["sha256"], name
    assert got["stats"] == want["stats"]


def test_cli_runs_analysis_and_fp32(tmp_path):
    print("[gpu case] cli analysis fp32", flush=True)
    out = tmp_path / "out"
    r = subprocess.run([CLI, "-c", os.path.join(GOLDEN, "configs", "quick.json"), "-t", "500",
                        "-p", "100", "-o", str(out), "--fp32"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    text = (out / "run.txt").read_text()
    assert "Calcium analysis" in text and "FP32 + FP64 islands" in text
    assert (out / "spikes.csv").read_text().startswith("time_ms,neuron\n")
