#!/usr/bin/env python3
"""Throughput of the simulation step on one MI355X: the 256-network sweep,
batched FP32 against the plain FP64 path.

Workload: 256 independent tripartite networks of uneven size -- N drawn
log-uniformly from [200, 5000] neurons (80/20 excitatory/inhibitory),
round(0.2 N) astrocytes, IP3 weight w_n2a uniform in [0.10, 0.40], seed
1000 + index, every other parameter from use_case.json, fixed mean in-degree
(p_primary = 100 / N) -- built on the host once, resident on the device.

  value        batched FP32 (FP32 conductances, rings in fixed point, the slow
               state V, w, Ca, IP3, h carried in FP64): all 256 networks in
               one asg_batch, schedule picked by the autotuner, then each
               call's windows replayed from captured hipGraphs or launched
               eagerly, whichever a short probe of both finds faster (the
               line says which). Neuron-steps/s =
               sum over networks of N x time steps / wall seconds of the
               timed bench steps. One bench step = 1,000 time steps (0.1 s
               simulated) of every network plus draining their spike
               rasters; --warmup bench steps run first, untimed, then
               --steps are timed.
  batched_fp64 the same batch in FP64 (bit-identical to the plain runs), timed
               over the accuracy run below.
  plain_fp64   the drop-in path: each network alone on the device in FP64,
               one after the other (every 8th network of the sweep).
  accuracy     always over 0.5 s transient + 5 s recorded (the sweep's
               duration), per network, FP32 against FP64: firing rate,
               active-astrocyte fraction, mean pairwise Ca correlation (the
               reference's analysis of Ca recorded every record_every steps),
               and the same figures for FP64 against FP64 with every initial
               V moved by 1e-6 mV (the chaos floor).
  roofline     the kernel with the largest share of device time in a
               profiled eager run of the same batch: algorithmic bytes per
               launch / mean launch time from HIP events on the batch's
               stream; traffic from the committed rocprofv3 counter summary.
  cpu_baseline the reference simulator built from its own sources
               (oracle/_ref) on this host: the sweep sample with one network
               per core on every core of the box's share, the same runs' one-
               thread rate, and the C1 / C2 configurations on all cores and on
               one thread.
  sparse_fp64  the reference's 'Sparse' network (10,000 neurons, 10,000
               astrocytes, paper_sparse.json), FP64, seeds 1-3 in one batch
               (graph replay), 1 s transient + --sparse-seconds timed; beside
               it the drop-in path (one asg_net per seed, eager) and seed 1's
               cell-kernel roofline.
  astro_scaling the isolated astrocyte update at 1e7 cells, 1,000 steps.
  serving      a bounded sample of the serving loop (--serve-jobs jobs,
               Poisson arrivals at 0.8 x the capacity measured first).
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
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
# the profiled run: events around every window would cost a noticeable share
# of the time they are meant to explain, around one in 25 about 1%.
PROFILE_EVERY = 25
SWEEP_CONFIG = os.path.join(HERE, "tests", "golden", "configs", "use_case.json")
SPARSE_CONFIG = os.path.join(HERE, "tests", "golden", "configs", "paper_sparse.json")
REF_BIN = os.path.join(HERE, "oracle", "_ref", "astrosimgpu_ref")
HBM_PEAK_GBS = 8000.0
STEPS_PER_BENCH_STEP = 1000
PROBE_STEPS = 3  # bench steps per launch-mode probe
WINDOW = 10  # min(d_e, d_i, d_a2n) of use_case.json and paper_sparse.json = 1 ms = 10 steps
ACC_TRANSIENT_BS, ACC_TIMED_BS = 5, 50  # accuracy run: 0.5 s + 5 s, in bench steps
SPARSE_SEEDS = (1, 2, 3)
# Published mean rates of the Sparse seeds over 1 s transient + 10 s
# (reference docs/validation.md:83-89).
SPARSE_REFERENCE_RATE_HZ = {1: 4.7390, 2: 4.7309, 3: 4.7636}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cell_bytes(n_neurons: int, n_astro: int, window: int, fp32: bool, noise_values: int = 1,
               per_cell_params: bool = True) -> float:
    """Algorithmic HBM bytes of one cell-kernel launch (one window of
    `window` steps; the cell state stays in registers across the window).

    Neuron, per window: V and w read + written (FP64 in both paths, 32),
    g_ex, dg_ex, g_in, dg_in, I_sic read + written (5 x 2 x s), refractory
    counter read + written (8), V_reset and b read (16); per step the exc and
    inh ring slots read (2 x 8: FP64 sums, or 64-bit fixed point); per noise
    interval of the window its noise value (s). Astrocyte, per window: Ca,
    IP3, h read + written (FP64 in both paths, 48) and 4 parameters (4 s);
    per step the n2a ring slot read (8) and the SIC factor written (s). s is
    8 (FP64) or 4 (FP32). Not counted: clears and SIC reads of nonzero slots
    (activity-dependent, so the bandwidth derived is a lower bound), and the
    Poisson counts the drive pass stages for the cell pass (an artefact of
    splitting the drive off, not of the step)."""
    s = 4 if fp32 else 8
    neuron = 32 + 10 * s + 8 + (16 if per_cell_params else 0) + window * 16 + noise_values * s
    astro = 48 + 4 * s + window * (8 + s)
    return n_neurons * neuron + n_astro * astro


def build_sweep(count: int, transient_ms: float, timed_ms: float):
    from astrosimgpu_amd import sweep
    specs = sweep.sweep_specs(count)
    return specs, [sweep.build(SWEEP_CONFIG, s, timed_ms, transient_ms) for s in specs]


def new_batch(models, precision, total_steps, record=True):
    import astrosimgpu_amd as A
    from astrosimgpu_amd import sweep
    batch = A.Batch(len(models))
    slots = []
    for m in models:
        m.desc.contents.precision = precision
        s = batch.add(m, total_steps)
        if record:
            start, every, n_rec, cap, _ = sweep._record_plan(m)
            batch.set_recording(s, start, every, n_rec, cap)
        slots.append(s)
    return batch, slots


def timed_steps(batch, slots, warmup: int, steps: int, keep_spikes: bool):
    """warmup untimed bench steps, then `steps` timed ones. Each launches
    its steps without waiting, splits the previous step's raster per network
    on the host while they run, then drains its own. Returns the timed
    seconds and, with keep_spikes, each network's spike steps."""
    spikes = [[] for _ in slots]

    def keep(raw):
        if keep_spikes:
            drained = batch.split_spikes(raw)
            for k, s in enumerate(slots):
                spikes[k].append(drained[s][0])

    for _ in range(warmup):
        batch.step(STEPS_PER_BENCH_STEP)
        keep(batch.spikes_all_raw())
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
    spk = [np.concatenate(x) if x else np.zeros(0, np.int64) for x in spikes]
    return elapsed, spk


def accuracy_run(models, precision, profile=False, lanes=0):
    """0.5 s + 5 s of the whole batch, eager launches, default schedule
    (or `lanes` lanes): timed seconds of the 5 s, spikes, Ca traces, and
    with `profile` per-window kernel times (a profiled run launches its
    windows in one lane, so its time is not the batch's throughput)."""
    batch, slots = new_batch(models, precision, (ACC_TRANSIENT_BS + ACC_TIMED_BS) * 1000)
    if lanes:
        batch.set_schedule(lanes=lanes)
    if profile:
        batch.set_profiling(True, every=PROFILE_EVERY)
    elapsed, spk = timed_steps(batch, slots, ACC_TRANSIENT_BS, ACC_TIMED_BS, True)
    kt = batch.kernel_times() if profile else None
    traces = [batch.read_trace(s)[:2] for s in slots]
    batch.close()
    return elapsed, spk, traces, kt


def per_window_us(kt: dict) -> dict:
    """Mean device microseconds per window of each pass over the timed
    windows (FP32 networks: the drive rides in the post pass, reported as
    "deliver"), and how many windows were timed."""
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


def sparse_fp64(transient_bench_steps: int, timed_bench_steps: int, lanes: int = 0):
    """The Sparse network, FP64, seeds 1-3 resident together in one batch
    (the FP64 batch is bit-identical to each network run alone), windows
    replayed from captured graphs; the drop-in path (each seed its own
    asg_net on its own stream, eager launches) timed the same way beside it;
    then seed 1 alone with per-kernel event timing for the roofline of its
    cell kernel."""
    import astrosimgpu_amd as A
    models = [A.Model(SPARSE_CONFIG, seed=s) for s in SPARSE_SEEDS]
    total = (transient_bench_steps + timed_bench_steps) * STEPS_PER_BENCH_STEP
    batch = A.Batch(len(models))
    slots = [batch.add(m, total) for m in models]
    if lanes:
        batch.set_schedule(lanes=lanes)
    batch.set_graphs(True)
    spikes = {s: [] for s in SPARSE_SEEDS}

    def batch_step():
        batch.step(STEPS_PER_BENCH_STEP)
        drained = batch.spikes_all()
        for s, slot in zip(SPARSE_SEEDS, slots):
            spikes[s].append(drained[slot][0])

    for _ in range(transient_bench_steps):
        batch_step()
    t0 = time.perf_counter()
    for _ in range(timed_bench_steps):
        batch_step()
    elapsed = time.perf_counter() - t0
    batch.close()
    first = transient_bench_steps * STEPS_PER_BENCH_STEP
    steps = timed_bench_steps * STEPS_PER_BENCH_STEP
    rates = {}
    for s in SPARSE_SEEDS:
        st = np.concatenate(spikes[s])
        rates[str(s)] = round(int((st >= first).sum()) * 1000.0 / (steps * 0.1 * 10000), 4)
    # the drop-in path: one asg_net per seed, each on its own stream, with
    # eager launches and then with its window sequences replayed from graphs
    def drop_in(graphs):
        nets = [A.DeviceNetwork(m) for m in models]
        for net in nets:
            net.set_graphs(graphs)

        def nets_step():
            for net in nets:
                net.step(STEPS_PER_BENCH_STEP, wait=False)
            for net in nets:
                net.sync()
                net.spikes()

        for _ in range(min(transient_bench_steps, 5)):
            nets_step()
        t1 = time.perf_counter()
        for _ in range(timed_bench_steps):
            nets_step()
        elapsed = time.perf_counter() - t1
        for net in nets:
            net.close()
        return elapsed

    elapsed_nets = drop_in(False)
    elapsed_nets_graphs = drop_in(True)
    # seed 1 alone, profiled: its cell kernel's bytes per window / time
    m = models[0]
    net = A.DeviceNetwork(m)
    net.set_profiling(True, every=1)
    net.step(10 * STEPS_PER_BENCH_STEP)
    kt = net.kernel_times()
    net.close()
    d = m.desc.contents
    noise = d.input_exc.gauss_noise_std > 0 or d.input_inh.gauss_noise_std > 0
    by = cell_bytes(m.n_neurons, m.n_astro, WINDOW, False, 1 if noise else 0)
    us = kt["cells"] * 1000.0 / max(kt["windows"], 1)
    out = {"value": round(3 * 10000 * steps / elapsed, 1), "unit": "neuron-steps/s",
           "launch": "the three seeds in one FP64 batch, graph replay",
           "drop_in_asg_net_value": round(3 * 10000 * steps / elapsed_nets, 1),
           "drop_in_asg_net_graphs_value": round(3 * 10000 * steps / elapsed_nets_graphs, 1),
           "seconds_timed": steps / 10000.0, "mean_rate_hz": rates,
           "kernel_us_per_window_seed1": per_window_us(kt),
           "roofline": {"kernel": "k_cells<FP64> (one Sparse network, one window of 10 steps)",
                        "bound": "hbm", "algorithmic_bytes_per_launch": round(by),
                        "mean_launch_us": round(us, 3),
                        "achieved": round(by / us / 1e3, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(by / us / 1e3 / HBM_PEAK_GBS, 5),
                        "note": "one 10,000-neuron network launches ~20k cells per window: "
                                "far too few to fill 256 CUs, so latency-bound by construction"}}
    if transient_bench_steps == 10 and timed_bench_steps == 100:
        out["reference_rate_hz"] = {str(s): r for s, r in SPARSE_REFERENCE_RATE_HZ.items()}
    return out


def _run_reference(cfg_path, seed, sim_ms, pre_ms, threads, out_dir):
    """One reference run; returns (neurons, recorded steps, step-loop seconds
    as its run.txt reports them)."""
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    subprocess.run([REF_BIN, "-c", cfg_path, "-s", str(seed), "-t", str(sim_ms), "-p", str(pre_ms),
                    "-o", out_dir, "--no-analysis"], check=True, env=env,
                   stdout=subprocess.DEVNULL)
    with open(os.path.join(out_dir, "run.txt")) as f:
        text = f.read()
    loop = float(re.search(r"^total\s+([0-9.]+)", text, flags=re.M).group(1))
    ne = int(re.search(r"^excitatory neurons\s+(\d+)", text, flags=re.M).group(1))
    ni = int(re.search(r"^inhibitory neurons\s+(\d+)", text, flags=re.M).group(1))
    return ne + ni, int(sim_ms / 0.1 + 0.5), loop


def reference_cpu_baseline(specs, sample_ms: float, every: int, transient_ms: float):
    """The reference build (oracle/_ref) on this host.

    Sweep sample: every `every`-th sweep network, `transient_ms` untimed then
    `sample_ms` timed by the reference's own step-loop timer, one network
    per core (single-threaded processes, as many at once as the cores of the
    box's share). value = the sum of the per-core rates (what the cores
    deliver while each holds a network); one_thread = the same runs' total
    neuron-steps / total loop seconds.

    C1 (use_case, 10 s, no transient) and C2 (paper_sparse seed 1, a 200 ms
    sample of its 10 s run): on one thread and on all cores."""
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    cores = max(1, min(cores, os.cpu_count() or 1))
    if not os.path.exists(REF_BIN):
        return None
    picked = list(range(0, len(specs), every))
    with open(SWEEP_CONFIG) as f:
        base = json.load(f)
    out = {"unit": "neuron-steps/s", "cores": cores, "kind": "reference"}
    with tempfile.TemporaryDirectory() as tmp:
        jobs = []
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
            jobs.append((path, sp.seed, os.path.join(tmp, f"out{i}")))
        with cf.ThreadPoolExecutor(max_workers=cores) as ex:
            runs = list(ex.map(lambda j: _run_reference(j[0], j[1], sample_ms, transient_ms, 1,
                                                        j[2]), jobs))
        out["value"] = round(sum(n * st / loop for n, st, loop in runs), 1)
        out["one_thread"] = round(sum(n * st for n, st, _ in runs) / sum(l for _, _, l in runs), 1)
        out["sample"] = (f"sweep networks {picked[0]}, {picked[1]}, ... every {every}th "
                         f"({len(picked)} networks, {sum(specs[i].n for i in picked)} neurons), "
                         f"{sample_ms:g} ms each after a {transient_ms:g} ms transient, one "
                         f"single-threaded reference process per core on {cores} cores; "
                         f"step-loop time of the recorded window (run.txt 'total')")
        use_case = os.path.join(HERE, "tests", "golden", "configs", "use_case.json")
        c1 = {}
        for threads in (1, cores):
            n, st, loop = _run_reference(use_case, 1, 10000.0, 0.0, threads,
                                         os.path.join(tmp, f"c1_{threads}"))
            c1[f"threads_{threads}"] = round(n * st / loop, 1)
        c1["workload"] = "use_case.json seed 1, 10 s simulated, no transient"
        out["C1_use_case"] = c1
        c2 = {}
        for threads in (1, cores):
            n, st, loop = _run_reference(SPARSE_CONFIG, 1, 200.0, 0.0, threads,
                                         os.path.join(tmp, f"c2_{threads}"))
            c2[f"threads_{threads}"] = round(n * st / loop, 1)
        c2["workload"] = "paper_sparse.json seed 1, the first 200 ms (2,000 steps) of its run"
        out["C2_sparse"] = c2
    return out


def load_traffic(kernel: str):
    path = os.path.join(HERE, "profiles", f"{kernel}_sweep_fp32_traffic.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    return None


def rounded(d):
    return None if d is None else {k: (round(v, 5) if isinstance(v, float) else v)
                                   for k, v in d.items()}


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
    ap.add_argument("--no-tune", action="store_true", help="default schedule, eager launches")
    ap.add_argument("--lanes", type=int, default=0, choices=(0, 1, 2, 3),
                    help="with --no-tune: lanes of the schedule (0 = the default, 2; the "
                         "FP32 lanes, and the FP64 groups of --fp64-only's timed run and of "
                         "--sparse-only's batch); "
                         "1 puts every kernel over the whole batch, as the profiled "
                         "run does, for kernel traces")
    ap.add_argument("--teams", type=int, default=0,
                    help="with --no-tune: run each call's FP32 windows in one persistent "
                         "launch of this many teams of networks (0 = per-kernel launches)")
    ap.add_argument("--fp32-only", action="store_true",
                    help="run only the batched FP32 sweep (for kernel profiles)")
    ap.add_argument("--fp64-only", action="store_true",
                    help="run only the batched FP64 sweep's 5 s run, profiled (for kernel "
                         "profiles)")
    ap.add_argument("--sparse-only", action="store_true",
                    help="run only the Sparse FP64 leg (for kernel profiles)")
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

    if args.sparse_only:
        print(json.dumps({"sparse_fp64": sparse_fp64(10, 10 * args.sparse_seconds, args.lanes)}),
              flush=True)
        return

    t_build = time.time()
    specs, models = build_sweep(args.networks, ACC_TRANSIENT_BS * 100.0, ACC_TIMED_BS * 100.0)
    n_total = sum(m.n_neurons for m in models)
    na_total = sum(m.n_astro for m in models)
    log(f"built {len(models)} networks ({n_total} neurons, {na_total} astrocytes) "
        f"in {time.time() - t_build:.1f} s")

    if args.fp64_only:
        e64, _, _, _ = accuracy_run(models, capi.ASG_FP64, lanes=args.lanes)
        _, _, _, kt64 = accuracy_run(models, capi.ASG_FP64, profile=True)
        print(json.dumps({"batched_fp64": n_total * ACC_TIMED_BS * STEPS_PER_BENCH_STEP / e64,
                          "kernel_us_per_window_fp64": per_window_us(kt64)}), flush=True)
        return

    # Headline: batched FP32, autotuned schedule, graph replay.
    log("batched FP32 ...")
    # <= 40 candidates x 3 rounds x 20 windows (x 2 for the lanes' untimed
    # calls), then the launch probe
    tune_steps = 0 if args.no_tune else (44 * 3 * 20 * WINDOW + 4 * (PROBE_STEPS + 1) * 1000)
    batch, slots = new_batch(models, capi.ASG_FP32,
                             tune_steps + (args.warmup + args.steps + 10) * 1000, record=False)
    schedule = None
    if args.no_tune and (args.lanes or args.teams):
        batch.set_schedule(lanes=args.lanes, teams=args.teams)
    # The transient first: the activity the schedule is tuned on (how many
    # networks have a live SIC exchange, how many blocks spike) should be
    # the timed run's.
    for _ in range(args.warmup):
        batch.step(STEPS_PER_BENCH_STEP)
        batch.spikes_all_raw()
    launch = "eager"
    probe = None
    if not args.no_tune:
        batch.set_graphs(True)
        schedule = batch.autotune(windows_per_try=20, rounds=3)
        batch.spikes_all_raw()
        log(f"  autotuned schedule {schedule}")
        # Graph replay or eager launches of the tuned schedule, whichever is
        # faster on a short probe of each (alternating, twice).
        probe = {"graphs": [], "eager": []}
        for _ in range(2):
            for mode in ("graphs", "eager"):
                batch.set_graphs(mode == "graphs")
                t, _ = timed_steps(batch, slots, 1, PROBE_STEPS, False)
                probe[mode].append(1000.0 * t / PROBE_STEPS)
        probe = {k: round(min(v), 3) for k, v in probe.items()}
        launch = "graphs" if probe["graphs"] <= probe["eager"] else "eager"
        batch.set_graphs(launch == "graphs")
        log(f"  launch probe ms per bench step {probe}: {launch}")
    e32, _ = timed_steps(batch, slots, 0, args.steps, False)
    timed_ns = n_total * args.steps * STEPS_PER_BENCH_STEP
    value = timed_ns / e32
    log(f"  {value:.3e} neuron-steps/s")
    # the same batch, eager with per-kernel events, for the roofline
    batch.set_graphs(False)
    batch.set_profiling(True, every=4)
    timed_steps(batch, slots, 0, 10, False)
    kt32 = batch.kernel_times()
    batch.close()
    log(f"  kernel us per window {per_window_us(kt32)}")

    cmp = floor = None
    e64 = kt64 = plain64 = sub = None
    if not args.fp32_only:
        log("accuracy, 0.5 s + 5 s: batched FP64 ...")
        e64, spk64, tr64, _ = accuracy_run(models, capi.ASG_FP64)
        log("  batched FP32 ...")
        a32, spk32, tr32, _ = accuracy_run(models, capi.ASG_FP32)
        acc_ns = n_total * ACC_TIMED_BS * STEPS_PER_BENCH_STEP
        r64 = sweep.RunResult(e64, acc_ns, spk64, tr64)
        cmp = sweep.compare(models, r64, sweep.RunResult(a32, acc_ns, spk32, tr32))
        cmp.pop("rows")
        # The networks are chaotic: any perturbation, however small, changes
        # individual spike times and then whole trajectories. The floor of
        # per-network agreement is what FP64 itself does when every initial
        # membrane potential moves by 1e-6 mV, a quarter of one FP32 ulp of V.
        log("  FP64 with the initial V 1e-6 mV higher (chaos floor) ...")
        saved = []
        for m in models:
            d = m.desc.contents
            saved.append((d.neuron[0].V_m_init, d.neuron[1].V_m_init))
            for k in (0, 1):
                d.neuron[k].V_m_init = d.neuron[k].V_m_init + 1e-6
        # (its kernel times, in one lane, are the FP64 per-window figures:
        # the same work as the run above)
        _, spkp, trp, kt64 = accuracy_run(models, capi.ASG_FP64, profile=True)
        for m, (v0, v1) in zip(models, saved):
            m.desc.contents.neuron[0].V_m_init = v0
            m.desc.contents.neuron[1].V_m_init = v1
        floor = sweep.compare(models, r64, sweep.RunResult(0.0, acc_ns, spkp, trp))
        floor.pop("rows")
        e64 = acc_ns / e64  # neuron-steps/s of the batched FP64 run

        log("plain FP64 (one network at a time) ...")
        sub = models[::args.plain_every]
        p_ns, p_s = run_plain(sub, (args.warmup + args.steps) * STEPS_PER_BENCH_STEP)
        plain64 = p_ns / p_s
        log(f"  {plain64:.3e} neuron-steps/s over {len(sub)} networks")

    # Roofline of the kernel with the largest share of device time.
    windows = kt32["windows"]
    share = {k: kt32[k] for k in ("cells", "compact", "deliver", "sum")}
    dominant = max(share, key=share.get)
    if dominant == "cells":
        # use_case.json: noise_dt = 1 ms = the window, one noise value per neuron
        bytes_launch = cell_bytes(n_total, na_total, WINDOW, True, noise_values=1)
        avg_s = kt32["cells"] / 1000.0 / max(windows, 1)
        achieved = bytes_launch / avg_s / 1e9
        roof = {"kernel": "k_cells<FP32> (every neuron and astrocyte of the batch, one window "
                          "of 10 steps)",
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
    fp64_roof = None
    if kt64 is not None:
        b64 = cell_bytes(n_total, na_total, WINDOW, False, noise_values=1)
        us64 = kt64["cells"] * 1000.0 / max(kt64["windows"], 1)
        fp64_roof = {"kernel": "k_cells<FP64>", "algorithmic_bytes_per_launch": round(b64),
                     "mean_launch_us": round(us64, 3), "achieved": round(b64 / us64 / 1e3, 2),
                     "frac": round(b64 / us64 / 1e3 / HBM_PEAK_GBS, 5)}

    sparse = None
    if not args.no_sparse and not args.fp32_only:
        log("Sparse FP64 ...")
        sparse = sparse_fp64(10, 10 * args.sparse_seconds)

    astro = None
    if not args.no_astro_scaling and not args.fp32_only:
        log("astrocyte update at 1e7 cells ...")
        from astrosimgpu_amd import astro_bench
        row = astro_bench.scaling([10_000_000], 1000)[0]
        astro = {"cells": row["n"], "steps": row["steps"],
                 "fp64_GBps": row["fp64"]["GB/s"], "fp32_GBps": row["fp32"]["GB/s"],
                 "peak_GBps": HBM_PEAK_GBS,
                 "fp32_rel_err_max": {k: v["max"] for k, v in row["fp32_vs_fp64_rel_err"].items()}}

    serving = None
    if args.serve_jobs > 0 and not args.fp32_only:
        log(f"serving sample, {args.serve_jobs} jobs ...")
        from astrosimgpu_amd import serve
        cap = serve.calibrate(serve.DEFAULT_CONFIG, 48, capi.ASG_FP32, 2000, 8)
        rate = 0.8 * cap / serve._mean_job_neuron_steps()
        jobs = serve.job_stream(args.serve_jobs, rate, seed=7)
        srv = serve.Server(serve.DEFAULT_CONFIG, chunk_steps=2000, builders=8)
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
    if not args.no_cpu_baseline and not args.fp32_only:
        log("timing the reference build on the host cores ...")
        cpu = reference_cpu_baseline(specs, args.cpu_sample_ms, args.cpu_every,
                                     ACC_TRANSIENT_BS * 100.0)

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
                        "path (FP32 conductances, fixed-point rings, V/w/Ca/IP3/h carried in "
                        f"FP64), {'autotuned schedule, ' if schedule else ''}{launch} launches, "
                        f"{args.warmup} bench steps untimed + {args.steps} timed, 1000 time "
                        "steps (dt 0.1 ms) per bench step; accuracy over 0.5 s + 5 s",
            "networks": len(models), "neurons": n_total, "astrocytes": na_total,
            "time_steps_per_bench_step": STEPS_PER_BENCH_STEP,
        },
        "schedule": schedule,
        "launch": launch,
        "launch_probe_ms_per_step": probe,
        "plain_fp64": None if plain64 is None else
        {"value": round(plain64, 1), "unit": "neuron-steps/s", "networks_timed": len(sub)},
        "batched_fp64": None if e64 is None else {"value": round(e64, 1), "unit": "neuron-steps/s"},
        "speedup_batched_fp32_vs_plain_fp64": None if plain64 is None else round(value / plain64, 2),
        "speedup_batched_fp32_vs_batched_fp64": None if e64 is None else round(value / e64, 2),
        "accuracy_fp32_vs_fp64_5s": rounded(cmp),
        "chaos_floor_fp64_vs_fp64_dV_1e-6_mV_5s": rounded(floor),
        "roofline": roof,
        "roofline_fp64_cells": fp64_roof,
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
