"""Serving loop: simulation jobs arrive over time and share the device.

Continuous batching: all running jobs are networks of one asg_batch and
advance together, one chunk of time steps per device call. Between chunks the
loop
  * hands newly arrived jobs to host builder threads (network construction
    with the per-row generator, which scales to 1e5-neuron jobs),
  * admits built jobs in arrival order while a slot is free and their device
    footprint fits the memory budget (asg_batch_add refuses otherwise),
  * retires finished jobs: their raster and summary are collected and their
    device memory is freed at once, which is what lets waiting jobs in.
A job's result does not depend on which other jobs shared its launches: each
network's arithmetic is its own, so a job run alone and in the mix gives the
same bits.

Reported: neuron-steps/s over the whole run (every job's neurons x steps /
makespan) and p50 / p99 of job latency (completion - arrival).
"""

from __future__ import annotations

import concurrent.futures as cf
import math
import os
import sys
import time
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _capi as capi
from .network import Batch, Model


DEFAULT_CONFIG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                              "tests", "golden", "configs", "paper_sparse.json")


@dataclass
class Job:
    id: int
    arrival_s: float
    n: int
    n_astro: int
    duration_ms: float
    w_n2a: float
    seed: int
    precision: int = capi.ASG_FP32
    # filled by the server
    model: Optional[Model] = None
    slot: Optional[int] = None
    total_steps: int = 0
    admitted_s: float = math.nan
    done_s: float = math.nan
    spike_steps: list = field(default_factory=list)
    spike_src: list = field(default_factory=list)

    @property
    def latency_s(self) -> float:
        return self.done_s - self.arrival_s

    def raster(self):
        if not self.spike_steps:
            return np.zeros(0, np.int64), np.zeros(0, np.uint32)
        return np.concatenate(self.spike_steps), np.concatenate(self.spike_src)


def job_stream(count: int, rate_per_s: float, seed: int = 7, n_min: int = 500,
               n_max: int = 100000, astro_ratio=(0.2, 1.0), duration_ms=(1000.0, 10000.0),
               w_range=(0.10, 0.40), precision: int = capi.ASG_FP32):
    """Poisson arrivals at `rate_per_s`; sizes log-uniform, astrocyte ratio,
    duration and IP3 weight uniform, independent seeds."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for i in range(count):
        t += rng.exponential(1.0 / rate_per_s)
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        na = max(1, int(round(rng.uniform(*astro_ratio) * n)))
        jobs.append(Job(i, t, n, na, float(rng.uniform(*duration_ms)), float(rng.uniform(*w_range)),
                        int(rng.integers(1, 2**31)), precision))
    return jobs


def build_job(base_config: str, job: Job, in_degree: float = 100.0) -> Model:
    n_exc = int(round(0.8 * job.n))
    m = Model(base_config, seed=job.seed, sim_time=job.duration_ms, pre_time=0.0,
              assignments={"N.N_E": n_exc, "N.N_I": job.n - n_exc, "N.N_A": job.n_astro,
                           "synapse.w_n2a": job.w_n2a,
                           "connectivity.p_primary": min(1.0, in_degree / job.n),
                           "connectivity.generator": "per_row",
                           "simulation.record_astro": False})
    m.desc.contents.precision = job.precision
    # the host side of admission (derived connectivity, plan) here, on the
    # builder's thread, rather than on the loop's thread between chunks
    m.prepare()
    return m


@dataclass
class ServeReport:
    jobs: int
    makespan_s: float
    neuron_steps: int
    neuron_steps_per_s: float
    latency_p50_s: float
    latency_p99_s: float
    refused_admissions: int    # jobs that waited for a slot or memory
    peak_running: int
    peak_bytes: int


class Server:
    """`batch` is the device batch (an asg_batch with `max_slots` slots and a
    memory budget of `budget_bytes`, 0 = free device memory, by default);
    `builder(job) -> Model` constructs a job's network on the host, on
    `executor` (a thread pool of `builders` threads by default)."""

    def __init__(self, base_config: str, max_slots: int = 512, budget_bytes: int = 0,
                 chunk_steps: int = 2000, builders: int = 2, batch=None, builder=None,
                 executor=None):
        self.base_config = base_config
        self.batch = batch if batch is not None else Batch(max_slots, budget_bytes)
        self.builder = builder if builder is not None else (lambda j: build_job(base_config, j))
        self.chunk_steps = chunk_steps
        self.pool = executor if executor is not None else cf.ThreadPoolExecutor(max_workers=builders)
        self.refused: set = set()   # ids of jobs that had to wait for room

    def run(self, jobs, clock=None, sleep=None, progress=None) -> ServeReport:
        """Serve `jobs` (arrival times in seconds from the start). `clock()`
        returns seconds since the start and `sleep(s)` waits; both default to
        the wall clock."""
        if clock is None:
            t0 = time.perf_counter()
            clock = lambda: time.perf_counter() - t0  # noqa: E731
        if sleep is None:
            sleep = time.sleep
        pending = sorted(jobs, key=lambda j: j.arrival_s)
        building: list = []      # (job, future)
        ready: list = []         # built, waiting for room
        running: dict = {}       # slot -> job
        done: list = []
        peak = [0, 0]            # running jobs, device bytes
        i = 0
        in_flight = False        # a chunk launched and not yet drained

        def intake():
            """Host work that needs no finished chunk: arrivals to the
            builders, built networks to the queue, admission (a network's
            upload does not wait for the chunk in flight)."""
            nonlocal i, building
            now = clock()
            while i < len(pending) and pending[i].arrival_s <= now:
                j = pending[i]
                building.append((j, self.pool.submit(self.builder, j)))
                i += 1
            still = []
            for j, fut in building:
                if fut.done():
                    j.model = fut.result()
                    j.total_steps = j.model.info["sim_steps"]
                    ready.append(j)
                else:
                    still.append((j, fut))
            building = still
            ready.sort(key=lambda j: j.arrival_s)
            # admission by slot and footprint, first come first served: the
            # oldest waiting job blocks later ones, so none is starved
            while ready and not (building and building[0][0].arrival_s < ready[0].arrival_s):
                j = ready[0]
                slot = self.batch.try_add(j.model, j.total_steps)
                if slot is None:
                    self.refused.add(j.id)
                    break
                ready.pop(0)
                j.slot = slot
                j.admitted_s = clock()
                running[slot] = j
            peak[0] = max(peak[0], len(running))
            peak[1] = max(peak[1], self.batch.bytes_in_use())

        while len(done) < len(jobs):
            intake()
            if in_flight:
                # the chunk's rasters (waits for the device), then the jobs
                # it finished leave and free their memory
                drained = self.batch.spikes_all()
                in_flight = False
                finished = []
                for slot, j in running.items():
                    if slot not in drained:
                        continue  # admitted while the chunk ran: not in it
                    st, src = drained[slot]
                    j.spike_steps.append(st)
                    j.spike_src.append(src)
                    if self.batch.net_step(slot) >= j.total_steps:
                        finished.append(slot)
                for slot in finished:
                    j = running.pop(slot)
                    self.batch.remove(slot)
                    j.done_s = clock()
                    j.model = None
                    done.append(j)
                    if progress is not None:
                        progress(len(done), len(jobs))
                if finished:
                    intake()  # waiting jobs into the room just freed
            if running:
                # the next chunk for every running job; the host takes in
                # arrivals and uploads new jobs while it runs
                self.batch.step(self.chunk_steps, wait=False)
                in_flight = True
                continue
            if building:
                cf.wait([f for _, f in building], return_when=cf.FIRST_COMPLETED)
            elif i < len(pending):
                sleep(max(0.0, pending[i].arrival_s - clock()))
            elif ready:
                raise RuntimeError(f"job {ready[0].id} does not fit an empty batch")
        peak_running, peak_bytes = peak
        makespan = clock()
        ns = sum(j.n * j.total_steps for j in done)
        lat = np.array([j.latency_s for j in done])
        return ServeReport(len(done), makespan, ns, ns / makespan if makespan > 0 else 0.0,
                           float(np.percentile(lat, 50)) if len(lat) else 0.0,
                           float(np.percentile(lat, 99)) if len(lat) else 0.0,
                           len(self.refused), peak_running, peak_bytes)

    def close(self):
        self.pool.shutdown(wait=True)
        self.batch.close()


def _progress_printer(t0):
    last = [0.0]

    def cb(done, total):
        now = time.perf_counter() - t0
        if now - last[0] >= 10.0 or done == total:
            last[0] = now
            print(f"  {done}/{total} jobs done at {now:.1f} s", file=sys.stderr, flush=True)
    return cb


def calibrate(base_config: str, count: int, precision: int, chunk_steps: int, builders: int,
              seed: int = 99, chunks: int = 5) -> float:
    """Capacity in neuron-steps/s: `count` jobs of the serving distribution
    resident in one batch, advanced `chunks` chunks of `chunk_steps` steps
    (after one untimed chunk) with every chunk's rasters drained, as the
    serving loop advances its running jobs. A steady-state figure: no job
    finishes during the timed chunks, so no tail of a few long jobs left
    alone on the device counts against it."""
    jobs = job_stream(count, rate_per_s=1e9, seed=seed, precision=precision)
    with cf.ThreadPoolExecutor(max_workers=max(1, builders)) as pool:
        models = list(pool.map(lambda j: build_job(base_config, j), jobs))
    batch = Batch(max(count, 1))
    try:
        for m in models:
            batch.add(m, (chunks + 1) * chunk_steps)
        batch.step(chunk_steps)
        batch.spikes_all_raw()
        t0 = time.perf_counter()
        for _ in range(chunks):
            batch.step(chunk_steps)
            batch.spikes_all_raw()
        seconds = time.perf_counter() - t0
    finally:
        batch.close()
    return sum(m.n_neurons for m in models) * chunks * chunk_steps / seconds


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description="Serve a Poisson stream of simulation jobs on one "
                                             "GPU and report throughput and job latency.")
    ap.add_argument("--jobs", type=int, default=2000)
    ap.add_argument("--rate", type=float, default=0.0,
                    help="arrivals per second; 0 = 0.8 x the capacity measured first")
    ap.add_argument("--calibration-jobs", type=int, default=64)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--precision", choices=("fp32", "fp64"), default="fp32")
    ap.add_argument("--chunk-steps", type=int, default=2000)
    ap.add_argument("--builders", type=int, default=8,
                    help="host threads that build and prepare jobs' networks")
    ap.add_argument("--max-slots", type=int, default=512)
    ap.add_argument("--budget-gb", type=float, default=0.0, help="0 = free device memory")
    ap.add_argument("--config", default=DEFAULT_CONFIG)
    ap.add_argument("--calibrate-only", action="store_true",
                    help="print the capacity measured with --calibration-jobs resident jobs "
                         "and exit")
    args = ap.parse_args(argv)
    prec = capi.ASG_FP32 if args.precision == "fp32" else capi.ASG_FP64
    if args.calibrate_only:
        cap = calibrate(args.config, args.calibration_jobs, prec, args.chunk_steps, args.builders)
        print(json.dumps({"capacity_neuron_steps_per_s": round(cap, 1),
                          "resident_jobs": args.calibration_jobs,
                          "chunk_steps": args.chunk_steps}), flush=True)
        return
    mean_ns = _mean_job_neuron_steps()
    capacity = None
    rate = args.rate
    if rate <= 0.0:
        print(f"calibrating with {args.calibration_jobs} jobs resident at once ...",
              file=sys.stderr, flush=True)
        capacity = calibrate(args.config, args.calibration_jobs, prec, args.chunk_steps,
                             args.builders)
        rate = 0.8 * capacity / mean_ns
        print(f"  capacity {capacity:.3e} neuron-steps/s -> rate {rate:.3f} jobs/s",
              file=sys.stderr, flush=True)
    jobs = job_stream(args.jobs, rate, seed=args.seed, precision=prec)
    srv = Server(args.config, max_slots=args.max_slots,
                 budget_bytes=int(args.budget_gb * (1 << 30)), chunk_steps=args.chunk_steps,
                 builders=args.builders)
    t0 = time.perf_counter()
    rep = srv.run(jobs, progress=_progress_printer(t0))
    srv.close()
    lat = np.array([j.latency_s for j in jobs])
    wait = np.array([j.admitted_s - j.arrival_s for j in jobs])
    out = {
        "jobs": rep.jobs, "rate_per_s": round(rate, 4), "precision": args.precision,
        "capacity_neuron_steps_per_s": capacity,
        "makespan_s": round(rep.makespan_s, 3),
        "neuron_steps": rep.neuron_steps,
        "neuron_steps_per_s": round(rep.neuron_steps_per_s, 1),
        "latency_p50_s": round(rep.latency_p50_s, 3),
        "latency_p99_s": round(rep.latency_p99_s, 3),
        "latency_mean_s": round(float(lat.mean()), 3),
        "admission_wait_p99_s": round(float(np.percentile(wait, 99)), 3),
        "jobs_that_waited_for_room": rep.refused_admissions,
        "peak_running": rep.peak_running,
        "peak_device_bytes": rep.peak_bytes,
        "chunk_steps": args.chunk_steps,
        "workload": "Poisson arrivals; N log-uniform [500, 1e5] (80/20 E/I, in-degree 100, "
                    "per-row generator), astrocytes U[0.2, 1.0] N, duration U[1, 10] s, "
                    "w_n2a U[0.10, 0.40], own seed; paper_sparse.json otherwise",
    }
    print(json.dumps(out), flush=True)


def _mean_job_neuron_steps(n_min=500, n_max=100000, d_min=1000.0, d_max=10000.0, dt=0.1):
    """Expected neuron-steps of one job: E[N] of the log-uniform size times
    E[steps] of the uniform duration."""
    mean_n = (n_max - n_min) / math.log(n_max / n_min)
    return mean_n * 0.5 * (d_min + d_max) / dt


if __name__ == "__main__":
    sys.exit(main())
