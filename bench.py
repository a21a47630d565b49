#!/usr/bin/env python3
"""Throughput of the simulation step on one MI355X: the 256-network sweep,
batched FP32 against the plain FP64 path.

Workload: 256 independent tripartite networks of uneven size -- N drawn
log-uniformly from [200, 5000] neurons (80/20 excitatory/inhibitory),
round(0.2 N) astrocytes, IP3 weight w_n2a uniform in [0.10, 0.40], seed
1000 + index, every other parameter from use_case.json, fixed mean in-degree
(p_primary = 100 / N) -- built on the host once, resident on the device.

  value        batched FP32 (state FP32, FP64 near threshold and for the Ca
               threshold test): all 256 networks in one asg_batch, every
               window one launch per kernel for all of them. Neuron-steps/s =
               sum over networks of N x time steps / wall seconds of the timed
               bench steps. One bench step = 1,000 time steps (0.1 s
               simulated) of every network plus draining their spike rasters;
               --warmup bench steps run first as the untimed transient, then
               --steps are timed (default 5 + 50: 0.5 s transient, 5 s timed).
  plain_fp64   the drop-in path: each network alone on the device in FP64,
               one after the other (every 8th network of the sweep, timed
               over the same time steps).
  batched_fp64 the same batch in FP64; it is bit-identical to the plain runs
               and serves as the FP64 reference of the accuracy figures.
  accuracy     per network, FP32 against FP64: firing rate over the timed
               window, active-astrocyte fraction, mean pairwise Ca
               correlation (the reference's analysis of Ca recorded every
               record_every steps).
  roofline     the kernel with the largest share of device time: algorithmic
               bytes per launch / mean launch time from HIP events recorded on
               the batch's stream.
  cpu_baseline the reference simulator built from its own sources
               (oracle/_ref) on this host's cores, on a sample of the sweep's
               networks.
  sparse_fp64  the reference's 'Sparse' network (10,000 neurons, 10,000
               astrocytes, paper_sparse.json), FP64, seeds 1-3 side by side,
               1 s transient + --sparse-seconds timed.
  astro_scaling the isolated astrocyte update at 1e7 cells, 1,000 steps, one
               launch per step: HBM GB/s of the update launches, FP64 and FP32,
               and the FP32-vs-FP64 state error.
  serving      a bounded sample of the serving loop (--serve-jobs jobs,
               Poisson arrivals at 0.8 x the capacity measured first): job
               latency p50/p99 and throughput.
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
# Kernel durations come from HIP events around one window in PROFILE_EVERY of
# the timed region (200 of the 5,000 windows): events around every window
# would cost ~14% of the wall time they are meant to explain, around one in
# 25 about 1%.
PROFILE_EVERY = 25
SWEEP_CONFIG = os.path.join(HERE, "tests", "golden", "configs", "use_case.json")
SPARSE_CONFIG = os.path.join(HERE, "tests", "golden", "configs", "paper_sparse.json")
HBM_PEAK_GBS = 8000.0
STEPS_PER_BENCH_STEP = 1000
SPARSE_SEEDS = (1, 2, 3)
# Published mean rates of the Sparse seeds over 1 s transient + 10 s
# (reference docs/validation.md:83-89).
SPARSE_REFERENCE_RATE_HZ = {1: 4.7390, 2: 4.7309, 3: 4.7636}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# Algorithmic HBM bytes of the cell-update kernel. It keeps each cell's
# state in registers for a whole window of W steps, so state moves once per
# window and ring slots every step.
#   neuron, per window: V, w, g_ex, dg_ex, g_in, dg_in, I_sic read + written
#     (7 x 2 x s), refractory counter read + written (8), V_reset and b read
#     (16); per step: exc, inh and SIC ring slots read (3 s), the exc and inh
#     slots cleared (2 s), and the external drive k_drive wrote for it: the
#     int32 Poisson count (4) and, with noise on, the noise current (s).
#   astrocyte, per window: Ca, IP3, h read + written (6 s), Ca_tot, IP3_0,
#     1/tau_IP3 and delta_IP3 read (4 s); per step: the IP3 ring slot read
#     and cleared (2 s) and the SIC factor written (s).
# s = 8 bytes in FP64, 4 in FP32.
def cell_bytes(n_neurons: int, n_astro: int, window: int, state_bytes: int,
               ring_bytes: int, noise_values: int = 1) -> float:
    """Bytes one cell-kernel launch must move for a window: per neuron the 7
    state values read and written (14 s), the refractory counter (8) and
    V_reset, b (16), per step the exc and inh ring slots (2 r) and the
    Poisson count (4), and per noise interval of the window its noise value
    (s; noise_values intervals, one when the noise interval is the window);
    per astrocyte 3 state values read and written and 4 parameters (10 s),
    and per step the n2a ring slot (r) and the SIC factor (s). s is the state
    value size, r the ring slot size (8 for the FP32 path's fixed-point
    rings). Reads of SIC slots of live exchanges and clears of nonzero ring
    slots depend on the activity and are not counted, so the bandwidth
    derived from these bytes is a lower bound."""
    s, r = state_bytes, ring_bytes
    neuron = 14 * s + 8 + 16 + window * (2 * r + 4) + noise_values * s
    astro = 10 * s + window * (r + s)
    return n_neurons * neuron + n_astro * astro


def build_sweep(count: int, transient_ms: float, timed_ms: float):
    from astrosimgpu_amd import sweep
    specs = sweep.sweep_specs(count)
    return specs, [sweep.build(SWEEP_CONFIG, s, timed_ms, transient_ms) for s in specs]


def run_batch(models, precision, warmup: int, steps: int, profile: bool):
    """One asg_batch of all models; returns elapsed seconds of the timed
    bench steps, per-network spike steps, Ca traces and kernel times."""
    import astrosimgpu_amd as A
    from astrosimgpu_amd import sweep
    batch = A.Batch(len(models))
    slots = []
    for m in models:
        m.desc.contents.precision = precision
        start, every, n_rec, cap, total = sweep._record_plan(m)
        s = batch.add(m, total)
        batch.set_recording(s, start, every, n_rec, cap)
        slots.append(s)
    spikes = [[] for _ in models]

    def keep(raw):
        drained = batch.split_spikes(raw)
        for k, s in enumerate(slots):
            spikes[k].append(drained[s][0])

    for _ in range(warmup):
        batch.step(STEPS_PER_BENCH_STEP)
        keep(batch.spikes_all_raw())
    if profile:
        batch.set_profiling(True, every=PROFILE_EVERY)
    # Each bench step launches its steps without waiting, splits the previous
    # step's raster per network on the host while they run, then drains its
    # own (the drain waits for the device: hipStreamSynchronize on the
    # batch's stream). The clock brackets exactly the timed work.
    t0 = time.perf_counter()
    raw = None
    for i in range(steps):
        batch.step(STEPS_PER_BENCH_STEP, wait=False)
        if raw is not None:
            keep(raw)
        raw = batch.spikes_all_raw()
        if (i + 1) % 10 == 0:
            log(f"    bench step {i + 1}/{steps}")
    keep(raw)
    elapsed = time.perf_counter() - t0
    kt = batch.kernel_times() if profile else None
    traces = []
    for s in slots:
        st, ca, _ = batch.read_trace(s)
        traces.append((st, ca))
    batch.close()
    spk = [np.concatenate(x) if x else np.zeros(0, np.int64) for x in spikes]
    return elapsed, spk, traces, kt


def per_window_us(kt: dict) -> dict:
    """Mean device microseconds per window of each kernel over the timed
    windows, and how many windows were timed."""
    w = max(kt["windows"], 1)
    out = {k: round(v * 1000.0 / w, 3) for k, v in kt.items() if k != "windows"}
    out["windows_timed"] = kt["windows"]
    return out


def run_plain(models, total_steps: int):
    """Each model alone in FP64 through the plain per-network path."""
    import astrosimgpu_amd as A
    from astrosimgpu_amd import _capi as capi
    elapsed = 0.0
    ns = 0
    for m in models:
        m.desc.contents.precision = capi.ASG_FP64
        net = A.DeviceNetwork(m)
        t0 = time.perf_counter()
        done = 0
        while done < total_steps:
            k = min(10 * STEPS_PER_BENCH_STEP, total_steps - done)
            net.step(k)
            net.spikes(1 << 16)
            done += k
        elapsed += time.perf_counter() - t0
        ns += m.n_neurons * total_steps
        net.close()
    return ns, elapsed


def sparse_fp64(transient_bench_steps: int, timed_bench_steps: int):
    """The Sparse network, FP64, seeds 1-3 resident together, each on its
    own stream."""
    import astrosimgpu_amd as A
    models = [A.Model(SPARSE_CONFIG, seed=s) for s in SPARSE_SEEDS]
    nets = [A.DeviceNetwork(m) for m in models]
    spikes = {s: [] for s in SPARSE_SEEDS}

    def bench_step():
        for net in nets:
            net.step(STEPS_PER_BENCH_STEP, wait=False)
        for s, net in zip(SPARSE_SEEDS, nets):
            net.sync()
            st, _ = net.spikes()
            spikes[s].append(st)

    for _ in range(transient_bench_steps):
        bench_step()
    t0 = time.perf_counter()
    for _ in range(timed_bench_steps):
        bench_step()
    elapsed = time.perf_counter() - t0
    first = transient_bench_steps * STEPS_PER_BENCH_STEP
    steps = timed_bench_steps * STEPS_PER_BENCH_STEP
    rates = {}
    for s in SPARSE_SEEDS:
        st = np.concatenate(spikes[s])
        rates[str(s)] = round(int((st >= first).sum()) * 1000.0 / (steps * 0.1 * 10000), 4)
    for net in nets:
        net.close()
    out = {"value": round(3 * 10000 * steps / elapsed, 1), "unit": "neuron-steps/s",
           "seconds_timed": steps / 10000.0, "mean_rate_hz": rates}
    if transient_bench_steps == 10 and timed_bench_steps == 100:
        out["reference_rate_hz"] = {str(s): r for s, r in SPARSE_REFERENCE_RATE_HZ.items()}
    return out


def reference_cpu_baseline(specs, models, sample_ms: float, every: int, transient_ms: float):
    """Time the reference build (oracle/_ref) on every `every`-th sweep
    network: `transient_ms` untimed, then `sample_ms` timed by the reference's
    own phase profile (it times the recorded window only), on min(16, cpus)
    OpenMP threads. Falls back to the CPU oracle when the reference build is
    absent."""
    threads = min(16, os.cpu_count() or 1)
    ref = os.path.join(HERE, "oracle", "_ref", "astrosimgpu_ref")
    picked = list(range(0, len(specs), every))
    steps = int(sample_ms / 0.1 + 0.5)
    sample = (f"sweep networks {', '.join(str(i) for i in picked)} "
              f"({sum(specs[i].n for i in picked)} neurons), {sample_ms:g} ms each after a "
              f"{transient_ms:g} ms transient, step-loop time of the recorded window")
    ns = 0
    loop_s = 0.0
    if os.path.exists(ref):
        with open(SWEEP_CONFIG) as f:
            base = json.load(f)
        with tempfile.TemporaryDirectory() as tmp:
            for i in picked:
                sp = specs[i]
                cfg = json.loads(json.dumps(base))
                cfg["N"].update({"N_E": sp.n_exc, "N_I": sp.n_inh, "N_A": sp.n_astro})
                cfg["synapse"]["w_n2a"] = sp.w_n2a
                cfg["connectivity"]["p_primary"] = min(1.0, 100.0 / sp.n)
                cfg["simulation"]["record_astro"] = False
                path = os.path.join(tmp, f"net{i}.json")
                with open(path, "w") as f:
                    json.dump(cfg, f)
                out = os.path.join(tmp, f"out{i}")
                env = dict(os.environ, OMP_NUM_THREADS=str(threads))
                subprocess.run([ref, "-c", path, "-s", str(sp.seed), "-t", str(sample_ms), "-p", str(transient_ms),
                                "-o", out, "--no-analysis"], check=True, env=env,
                               stdout=subprocess.DEVNULL)
                with open(os.path.join(out, "run.txt")) as f:
                    m = re.search(r"^total\s+([0-9.]+)", f.read(), flags=re.M)
                loop_s += float(m.group(1))
                ns += sp.n * steps
        kind = "reference"
    else:
        import oracle
        os.environ.setdefault("OMP_NUM_THREADS", str(threads))
        for i in picked:
            o = oracle.OracleRun(models[i].desc, oracle.LIBM_GLIBC)
            o.step(int(transient_ms / 0.1 + 0.5))
            t0 = time.time()
            o.step(steps)
            loop_s += time.time() - t0
            ns += specs[i].n * steps
        kind = "port"
    return {"value": round(ns / loop_s, 1), "unit": "neuron-steps/s", "cores": threads,
            "kind": kind, "sample": sample, "seconds": round(loop_s, 3)}


def load_traffic(kernel: str):
    path = os.path.join(HERE, "profiles", f"{kernel}_sweep_fp32_traffic.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    return None


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50, help="timed bench steps (1,000 time steps each)")
    ap.add_argument("--warmup", type=int, default=5, help="untimed bench steps before timing")
    ap.add_argument("--networks", type=int, default=256)
    ap.add_argument("--plain-every", type=int, default=8,
                    help="time the plain FP64 path on every n-th network")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-ms", type=float, default=1500.0)
    ap.add_argument("--cpu-every", type=int, default=16)
    ap.add_argument("--no-sparse", action="store_true")
    ap.add_argument("--fp32-only", action="store_true",
                    help="run only the batched FP32 sweep (for kernel profiles)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy kernel of MIB MiB five times "
                         "(the reference point for HBM byte counters) and exit")
    ap.add_argument("--sparse-seconds", type=int, default=2,
                    help="timed simulated seconds of the Sparse FP64 runs (10 = the full run)")
    ap.add_argument("--no-astro-scaling", action="store_true")
    ap.add_argument("--serve-jobs", type=int, default=120,
                    help="jobs of the serving sample (0 = skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (got --gpus {args.gpus}); the "
              "networks of a sweep would be bin-packed over GPUs by footprint, one batch per "
              "GPU, which is not built here", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration > 0:
        import torch
        n = args.copy_calibration << 20
        src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_kernel_bytes_read": n, "copy_kernel_bytes_written": n,
                          "launches": 5}), flush=True)
        return

    from astrosimgpu_amd import _capi as capi
    from astrosimgpu_amd import sweep

    t_build = time.time()
    transient_ms = args.warmup * STEPS_PER_BENCH_STEP * 0.1
    timed_ms = args.steps * STEPS_PER_BENCH_STEP * 0.1
    specs, models = build_sweep(args.networks, transient_ms, timed_ms)
    total_steps = (args.warmup + args.steps) * STEPS_PER_BENCH_STEP
    n_total = sum(m.n_neurons for m in models)
    na_total = sum(m.n_astro for m in models)
    log(f"built {len(models)} networks ({n_total} neurons, {na_total} astrocytes) "
        f"in {time.time() - t_build:.1f} s")

    log("batched FP32 ...")
    e32, spk32, tr32, kt32 = run_batch(models, capi.ASG_FP32, args.warmup, args.steps, True)
    timed_ns = n_total * args.steps * STEPS_PER_BENCH_STEP
    value = timed_ns / e32
    log(f"  {value:.3e} neuron-steps/s  kernel us per window {per_window_us(kt32)}")

    if args.fp32_only:
        args.no_sparse = args.no_cpu_baseline = args.no_astro_scaling = True
        args.serve_jobs = 0
        e64 = kt64 = plain64 = cmp = floor = sub = None
    else:
        log("batched FP64 ...")
        e64, spk64, tr64, kt64 = run_batch(models, capi.ASG_FP64, args.warmup, args.steps, True)
        log(f"  {timed_ns / e64:.3e} neuron-steps/s")

        log("plain FP64 (one network at a time) ...")
        sub = models[::args.plain_every]
        p_ns, p_s = run_plain(sub, total_steps)
        plain64 = p_ns / p_s
        log(f"  {plain64:.3e} neuron-steps/s over {len(sub)} networks")

        r64 = sweep.RunResult(e64, timed_ns, spk64, tr64)
        r32 = sweep.RunResult(e32, timed_ns, spk32, tr32)
        cmp = sweep.compare(models, r64, r32)
        cmp.pop("rows")

        # The networks are chaotic: any perturbation, however small, changes
        # individual spike times and then whole trajectories. The floor of
        # per-network agreement is what FP64 itself does when every initial
        # membrane potential moves by 1e-6 mV, a quarter of one FP32 ulp of V
        # (a one-ulp FP64 change is absorbed by the first step's rounding).
        log("FP64 with the initial V 1e-6 mV higher (chaos floor) ...")
        saved = []
        for m in models:
            d = m.desc.contents
            saved.append((d.neuron[0].V_m_init, d.neuron[1].V_m_init))
            for k in (0, 1):
                d.neuron[k].V_m_init = d.neuron[k].V_m_init + 1e-6
        _, spkp, trp, _ = run_batch(models, capi.ASG_FP64, args.warmup, args.steps, False)
        for m, (v0, v1) in zip(models, saved):
            m.desc.contents.neuron[0].V_m_init = v0
            m.desc.contents.neuron[1].V_m_init = v1
        floor = sweep.compare(models, r64, sweep.RunResult(0.0, timed_ns, spkp, trp))
        floor.pop("rows")

    # Roofline of the kernel with the largest share of device time.
    windows = kt32["windows"]
    share = {k: kt32[k] for k in ("cells", "compact", "deliver", "sum")}
    dominant = max(share, key=share.get)
    window_len = 10  # min(d_e, d_i, d_a2n) of use_case.json = 1 ms = 10 steps
    if dominant == "cells":
        # use_case.json drives neurons with Poisson input and noise
        # noise_dt of use_case.json = 1 ms = the window: one noise value each
        bytes_launch = cell_bytes(n_total, na_total, window_len, 4, 8, noise_values=1)
        avg_s = kt32["cells"] / 1000.0 / max(windows, 1)
        achieved = bytes_launch / avg_s / 1e9
        roof = {"kernel": "k_cells (every neuron and astrocyte of the batch, one window of "
                          "10 steps, FP32)",
                "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": load_traffic("k_cells"),
                "algorithmic_bytes_per_launch": round(bytes_launch),
                "mean_launch_us": round(avg_s * 1e6, 3), "launches": windows}
    else:
        roof = {"kernel": dominant, "bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": None, "traffic": None,
                "mean_launch_us": round(share[dominant] * 1000.0 / max(windows, 1), 3),
                "launches": windows}

    sparse = None
    if not args.no_sparse:
        log("Sparse FP64 ...")
        sparse = sparse_fp64(10, 10 * args.sparse_seconds)

    astro = None
    if not args.no_astro_scaling:
        log("astrocyte update at 1e7 cells ...")
        from astrosimgpu_amd import astro_bench
        row = astro_bench.scaling([10_000_000], 1000)[0]
        astro = {"cells": row["n"], "steps": row["steps"],
                 "fp64_GBps": row["fp64"]["GB/s"], "fp32_GBps": row["fp32"]["GB/s"],
                 "peak_GBps": HBM_PEAK_GBS,
                 "fp32_rel_err_max": {k: v["max"] for k, v in row["fp32_vs_fp64_rel_err"].items()}}

    serving = None
    if args.serve_jobs > 0:
        log(f"serving sample, {args.serve_jobs} jobs ...")
        from astrosimgpu_amd import serve
        cap = serve.calibrate(serve.DEFAULT_CONFIG, 48, capi.ASG_FP32, 2000, 4)
        rate = 0.8 * cap / serve._mean_job_neuron_steps()
        jobs = serve.job_stream(args.serve_jobs, rate, seed=7)
        srv = serve.Server(serve.DEFAULT_CONFIG, chunk_steps=2000, builders=4)
        rep_ = srv.run(jobs)
        srv.close()
        serving = {"jobs": rep_.jobs, "rate_per_s": round(rate, 3),
                   "latency_p50_s": round(rep_.latency_p50_s, 3),
                   "latency_p99_s": round(rep_.latency_p99_s, 3),
                   "neuron_steps_per_s": round(rep_.neuron_steps_per_s, 1),
                   "capacity_neuron_steps_per_s": round(cap, 1),
                   "workload": "Poisson arrivals at 0.8 x capacity; N log-uniform [500, 1e5], "
                               "astrocytes U[0.2, 1] N, 1-10 s each, FP32"}

    cpu = None
    if not args.no_cpu_baseline:
        log("timing the reference build on the host cores ...")
        cpu = reference_cpu_baseline(specs, models, args.cpu_sample_ms, args.cpu_every,
                                     transient_ms)

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "neuron-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * e32 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: networks built from use_case.json with drawn sizes, IP3 weights "
                "and seeds; Poisson and Gaussian drive",
        "config": {
            "workload": f"{len(models)}-network sweep (N log-uniform in [200, 5000], 80/20 E/I, "
                        "0.2 N astrocytes, w_n2a U[0.10, 0.40], in-degree 100), batched FP32 "
                        f"path, {transient_ms / 1000:g} s transient untimed + "
                        f"{timed_ms / 1000:g} s timed, 1000 time steps (dt 0.1 ms) per bench step",
            "networks": len(models), "neurons": n_total, "astrocytes": na_total,
            "time_steps_per_bench_step": STEPS_PER_BENCH_STEP,
        },
        "plain_fp64": None if plain64 is None else
        {"value": round(plain64, 1), "unit": "neuron-steps/s", "networks_timed": len(sub)},
        "batched_fp64": None if e64 is None else
        {"value": round(timed_ns / e64, 1), "unit": "neuron-steps/s"},
        "speedup_batched_fp32_vs_plain_fp64": None if plain64 is None else round(value / plain64, 2),
        "accuracy_fp32_vs_fp64": None if cmp is None else
        {k: (round(v, 5) if isinstance(v, float) else v) for k, v in cmp.items()},
        "chaos_floor_fp64_vs_fp64_dV_1e-6_mV": None if floor is None else
        {k: (round(v, 5) if isinstance(v, float) else v) for k, v in floor.items()},
        "roofline": roof,
        "kernel_us_per_window_fp32": per_window_us(kt32),
        "kernel_us_per_window_fp64": None if kt64 is None else per_window_us(kt64),
        "sparse_fp64": sparse,
        "astro_scaling": astro,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
