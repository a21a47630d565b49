#!/usr/bin/env python3
"""Writes the calcium-analysis fixtures (tests/golden/analysis_*.npz + .json).

For each case the reference simulator, built from its own sources
(oracle/_ref/astrosimgpu_ref, see oracle/Makefile), runs a configuration
with astrocyte recording on. The fixture keeps its recorded Ca traces
(astrocytes.csv as written, 6 significant digits) and the calcium-analysis
block it prints in run.txt (main.cpp:264-313). The test feeds the same traces
to asg_analyze_calcium and compares with those printed numbers.

Run from the repository root in the build container:
    python tests/golden/make_analysis_golden.py
"""

from __future__ import annotations

import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(REPO, "oracle", "_ref", "astrosimgpu_ref")

# name: (config, seed, sim_time ms, pre_time ms, astrocytes recorded)
CASES = {
    "use_case_60s": ("use_case", 1, 60000.0, 0.0, 12),
    "bursting_20s": ("bursting", 2, 20000.0, 0.0, 20),
    "use_case_short": ("use_case", 3, 1500.0, 500.0, 20),  # hist_window >= sim_time
}


def parse_analysis(text: str) -> dict:
    pats = {
        "threshold": r"Calcium analysis \(threshold ([0-9.]+) uM, window ([0-9.]+)/([0-9.]+) ms\)",
        "active": r"astrocytes with transients\s+(\d+) / (\d+)",
        "transients": r"transients detected\s+(\d+)",
        "duration": r"mean duration\s+([0-9.]+) ms \(sd ([0-9.]+)\)",
        "pairs": r"pairs correlated\s+(\d+)",
        "correlation": r"mean pairwise correlation\s+(-?[0-9.]+)",
    }
    out = {}
    m = re.search(pats["threshold"], text)
    out["threshold"], out["width"], out["shift"] = (float(m.group(i)) for i in (1, 2, 3))
    m = re.search(pats["active"], text)
    out["active"], out["cells"] = int(m.group(1)), int(m.group(2))
    out["transients"] = int(re.search(pats["transients"], text).group(1))
    m = re.search(pats["duration"], text)
    out["mean_duration"], out["sd_duration"] = float(m.group(1)), float(m.group(2))
    out["pairs"] = int(re.search(pats["pairs"], text).group(1))
    out["mean_correlation"] = float(re.search(pats["correlation"], text).group(1))
    return out


# libstdc++'s std::ofstream writes its 8192-byte buffer to the file only when
# it fills; the reference analyses astrocytes.csv (main.cpp:268) while the
# Recorder that writes it is still open, so it reads the whole 8192-byte
# chunks written so far -- the last line possibly cut short.
FILEBUF = 8192

_NUMBER = re.compile(r"^[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?")


def _strtod(field: str) -> float:
    m = _NUMBER.match(field)
    return float(m.group(0)) if m else 0.0


def traces_as_analysed(data: bytes):
    """Per-cell (times, Ca) as read_astro_traces (main.cpp:46-74) reads the
    flushed prefix of the file: header skipped, empty lines skipped, a line
    short of three fields skipped, numbers parsed by strtod."""
    text = data[: len(data) // FILEBUF * FILEBUF].decode()
    lines = text.split("\n")[1:]
    times, calcium = {}, {}
    for line in lines:
        if not line:
            continue
        fields = line.split(",")
        if len(fields) < 3 or fields[2] == "":
            continue
        cell = int(_NUMBER.match(fields[1]).group(0)) if _NUMBER.match(fields[1]) else 0
        times.setdefault(cell, []).append(_strtod(fields[0]))
        calcium.setdefault(cell, []).append(_strtod(fields[2]))
    cells = sorted(calcium)
    return [times[c] for c in cells], [calcium[c] for c in cells]


def strip_comments(text: str) -> str:
    return "\n".join(line for line in text.splitlines() if not line.lstrip().startswith("//"))


def main():
    if not os.path.exists(REF):
        sys.exit(f"{REF} is missing: make -C oracle ref")
    for name, (config, seed, sim, pre, n_rec) in CASES.items():
        with open(os.path.join(HERE, "configs", config + ".json")) as f:
            cfg = json.loads(strip_comments(f.read()))
        cfg["simulation"]["record_astro"] = True
        cfg["simulation"]["record_astro_max"] = n_rec
        cfg.setdefault("analysis", {})["enabled"] = True
        with tempfile.TemporaryDirectory() as tmp:
            path = os.path.join(tmp, "cfg.json")
            with open(path, "w") as f:
                json.dump(cfg, f)
            out = os.path.join(tmp, "out")
            subprocess.run([REF, "-c", path, "-s", str(seed), "-t", str(sim), "-p", str(pre),
                            "-o", out], check=True, stdout=subprocess.DEVNULL,
                           env=dict(os.environ, OMP_NUM_THREADS="8"))
            with open(os.path.join(out, "run.txt")) as f:
                expected = parse_analysis(f.read())
            with open(os.path.join(out, "astrocytes.csv"), "rb") as f:
                data = f.read()
        times, ca = traces_as_analysed(data)
        # Stored losslessly: Ca as integer counts of 1e-8 (every value read
        # has at most 8 decimals), differenced along each cell's series;
        # q / 1e8 is the double nearest the decimal, as parsing it gives.
        offsets = np.zeros(len(ca) + 1, np.int64)
        offsets[1:] = np.cumsum([len(x) for x in ca])
        t_cat = np.concatenate(times)
        tq = np.rint(t_cat * 1e3).astype(np.int64)   # times have at most 3 decimals
        assert np.array_equal(tq / 1e3, t_cat)
        dq = []
        for x in ca:
            q = np.rint(np.array(x) * 1e8).astype(np.int64)
            assert np.array_equal(q / 1e8, np.array(x))
            dq.append(np.diff(q, prepend=0))
        np.savez_compressed(os.path.join(HERE, f"analysis_{name}.npz"),
                            times_dq=np.diff(tq, prepend=0), ca_dq=np.concatenate(dq).astype(np.int32),
                            offsets=offsets)
        expected["csv_bytes"] = len(data)
        expected["csv_bytes_read"] = len(data) // FILEBUF * FILEBUF
        params = cfg.get("analysis", {})
        # the threshold the analysis used (main.cpp:271-273); run.txt prints
        # it rounded to 4 decimals
        ca_th = params.get("ca_threshold", -1.0)
        expected["threshold_printed"] = expected.pop("threshold")
        expected["threshold"] = ca_th if ca_th >= 0.0 else cfg["astrocyte"]["SIC_th"]
        expected.update({
            "config": config, "seed": seed, "sim_time": sim, "pre_time": pre,
            "merge_gap": params.get("merge_gap", 2000.0),
            "hist_window": params.get("hist_window", 2000.0),
            "hist_shift": params.get("hist_shift", 400.0),
            "producer": "reference build (oracle/_ref) run.txt calcium analysis",
        })
        with open(os.path.join(HERE, f"analysis_{name}.json"), "w") as f:
            json.dump(expected, f, indent=1)
        print(name, expected)


if __name__ == "__main__":
    main()
