#!/usr/bin/env python3
"""Throughput of the simulation step on one MI355X.

Workload (default): the reference's "Sparse" network -- 8,000 excitatory +
2,000 inhibitory AdEx neurons and 10,000 Li-Rinzel astrocytes built from
paper_sparse.json -- whole step on the device in FP64 (the plain drop-in
path), seeds 1, 2 and 3 resident side by side, each on its own HIP stream so
their launches overlap. One bench step advances each of the three networks by
1,000 time steps (0.1 s simulated) and drains its spike raster to the host. The default
--warmup 10 --steps 100 is the reference's run of those seeds: a 1 s transient
(untimed) and 10 s recorded (timed).

Reported on one JSON line (rank 0):
  value      neuron-steps/s over the timed region = networks x 10,000 neurons
             x time steps / wall seconds (inputs already resident on the device)
  roofline   the dominant kernel (the cell-update kernel): algorithmic bytes
             per launch / mean launch time from HIP events on its stream
  cpu_baseline  the reference simulator built from its own sources
             (oracle/_ref) on this host's cores, first 2 s of seed 1
"""

from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "neuron-steps/s batched FP32 vs plain FP64 on 1 GPU; HBM GB/s; p99 job latency"
CONFIG = os.path.join(HERE, "tests", "golden", "configs", "paper_sparse.json")
HBM_PEAK_GBS = 8000.0
SIM_STEPS_PER_BENCH_STEP = 1000
SEEDS = (1, 2, 3)
# Published mean rates of these seeds (reference docs/validation.md:83-89).
REFERENCE_RATE_HZ = {1: 4.7390, 2: 4.7309, 3: 4.7636}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cells_bytes_per_launch(model, window: int) -> float:
    """Algorithmic HBM bytes of one cell-update launch (one window).

    The kernel keeps each cell's state in registers for the whole window, so
    state moves once per window; ring slots move every step. Per neuron:
    state V, w, g_ex, dg_ex, g_in, dg_in, I_sic (8 B each) read + written,
    refractory counter (4 B) read + written, V_reset and b read (136 B); per
    neuron-step: the exc and inh ring slots read and cleared (32 B) and the
    SIC slot read (8 B). Per astrocyte: Ca, IP3, h read + written and Ca_tot,
    IP3_0, tau_IP3, delta_IP3 read (80 B); per astrocyte-step: its ring slot
    read and cleared (16 B) and its SIC factor written (8 B).
    """
    n = model.n_neurons
    na = model.n_astro
    per_window = n * 136.0 + na * 80.0
    per_step = n * 40.0 + na * 24.0
    return per_window + window * per_step


def reference_cpu_baseline(sample_ms: float = 2000.0):
    """Time the reference build (oracle/_ref) on the Sparse network, seed 1,
    first `sample_ms` of simulated time, on min(16, cpus) OpenMP threads.
    Falls back to the repo's CPU oracle if the reference build is absent."""
    threads = min(16, os.cpu_count() or 1)
    ref = os.path.join(HERE, "oracle", "_ref", "astrosimgpu_ref")
    steps = int(sample_ms / 0.1 + 0.5)
    sample = (f"paper_sparse.json seed 1, steps 0-{steps} ({sample_ms / 1000:g} s simulated, "
              "the start of the run), phase-profile total of the step loop")
    if os.path.exists(ref):
        with tempfile.TemporaryDirectory() as out:
            env = dict(os.environ, OMP_NUM_THREADS=str(threads))
            t0 = time.time()
            subprocess.run([ref, "-c", CONFIG, "-s", "1", "-t", str(sample_ms), "-p", "0",
                            "-o", out, "--no-analysis"], check=True, env=env,
                           stdout=subprocess.DEVNULL)
            wall = time.time() - t0
            text = open(os.path.join(out, "run.txt")).read()
        m = re.search(r"^total\s+([0-9.]+)", text, flags=re.M)
        loop_s = float(m.group(1)) if m else wall
        return {"value": 10000.0 * steps / loop_s, "unit": "neuron-steps/s", "cores": threads,
                "kind": "reference", "sample": sample, "seconds": round(loop_s, 3)}
    import oracle
    import astrosimgpu_amd as A
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    mdl = A.Model(CONFIG, seed=1)
    o = oracle.OracleRun(mdl.desc, oracle.LIBM_GLIBC)
    t0 = time.time()
    o.step(steps)
    loop_s = time.time() - t0
    return {"value": 10000.0 * steps / loop_s, "unit": "neuron-steps/s", "cores": threads,
            "kind": "port", "sample": sample, "seconds": round(loop_s, 3)}


def load_traffic():
    path = os.path.join(HERE, "profiles", "k_cells_sparse_fp64_traffic.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    return None


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed bench steps (1,000 time steps each)")
    ap.add_argument("--warmup", type=int, default=10, help="untimed bench steps before timing")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-ms", type=float, default=2000.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (got --gpus {args.gpus}); the path "
              "is single-device, see DESIGN.md for how it would shard", file=sys.stderr)
        sys.exit(2)

    import astrosimgpu_amd as A

    t_build = time.time()
    models = [A.Model(CONFIG, seed=s) for s in SEEDS]
    nets = [A.DeviceNetwork(m) for m in models]
    log(f"built {len(models)} Sparse networks in {time.time() - t_build:.1f} s "
        f"(synapses {models[0].stats()})")

    spikes = {s: [] for s in SEEDS}

    def bench_step():
        # Each network has its own HIP stream: queue all three, then wait,
        # so their kernels overlap on the device.
        for net in nets:
            net.step(SIM_STEPS_PER_BENCH_STEP, wait=False)
        for s, net in zip(SEEDS, nets):
            net.sync()
            st, _ = net.spikes()
            spikes[s].append(st)

    for i in range(args.warmup):
        bench_step()
    for net in nets:
        net.set_profiling(True)
    # asg_net_step returns after the device has finished (hipStreamSynchronize),
    # so the wall clock below brackets exactly the timed steps.
    t0 = time.perf_counter()
    for i in range(args.steps):
        bench_step()
        if (i + 1) % 20 == 0:
            log(f"  timed bench step {i + 1}/{args.steps}")
    elapsed = time.perf_counter() - t0

    kt = [net.kernel_times() for net in nets]
    n_total = sum(m.n_neurons for m in models)
    sim_steps = args.steps * SIM_STEPS_PER_BENCH_STEP
    value = n_total * sim_steps / elapsed
    windows = sum(k["windows"] for k in kt)
    cells_ms = sum(k["cells"] for k in kt)
    window_len = 10  # min(d_e, d_i, d_a2n) = 10 steps for this configuration
    bytes_launch = cells_bytes_per_launch(models[0], window_len)
    cells_avg_s = cells_ms / 1000.0 / max(windows, 1)
    achieved = bytes_launch / cells_avg_s / 1e9
    first_rec = args.warmup * SIM_STEPS_PER_BENCH_STEP
    rates = {}
    for s in SEEDS:
        st = np.concatenate(spikes[s]) if spikes[s] else np.zeros(0)
        rec = st[st >= first_rec]
        rates[s] = round(len(rec) * 1000.0 / (sim_steps * 0.1 * 10000), 4)

    cpu = None
    if not args.no_cpu_baseline:
        log("timing the reference build on the host cores ...")
        cpu = reference_cpu_baseline(args.cpu_sample_ms)

    counters = [net.counters() for net in nets]
    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "neuron-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: Sparse network built from paper_sparse.json, Poisson drive",
        "config": {
            "workload": "Sparse network (8000 E + 2000 I AdEx neurons, 10000 Li-Rinzel "
                        "astrocytes), plain FP64 whole step, seeds 1-3, 1 s transient untimed "
                        "+ 10 s timed, 1000 time steps (dt 0.1 ms) per bench step",
            "networks": len(models),
            "neurons_per_network": models[0].n_neurons,
            "astrocytes_per_network": models[0].n_astro,
            "time_steps_per_bench_step": SIM_STEPS_PER_BENCH_STEP,
            "precision": "fp64",
        },
        "roofline": {
            "kernel": "k_cells (neuron + astrocyte update, one window of 10 steps)",
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": load_traffic(),
            "algorithmic_bytes_per_launch": round(bytes_launch),
            "mean_launch_us": round(cells_avg_s * 1e6, 3),
            "launches": windows,
        },
        "kernel_ms": {k: round(sum(x[k] for x in kt), 3) for k in ("cells", "compact", "deliver", "sum")},
        "mean_rate_hz": {str(s): rates[s] for s in SEEDS},
        "reference_rate_hz": {str(s): REFERENCE_RATE_HZ[s] for s in SEEDS},
        "synapse_events": sum(c["synapse_events"] for c in counters),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
