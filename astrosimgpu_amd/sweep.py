"""Parameter sweeps: many independent networks of uneven size stepped in
shared launches, each compared against its FP64 run.

The reference runs one network per process; its scripts sweep the IP3
weight and seeds one run at a time (scripts/sic_interval.sh,
scripts/convergence.sh). Here a sweep is one asg_batch: every network's
cells, spikes and deliveries go through the same four kernels per window.

Sweep definition (the 256-network sweep): network sizes N drawn
log-uniformly from [200, 5000] (numpy default_rng(4)), 80/20 excitatory /
inhibitory, astrocytes round(0.2 N) in block pools; the IP3 weight w_n2a
uniform in [0.10, 0.40]; seed 1000 + index; every other parameter from
use_case.json. All networks use the same connectivity rule at fixed mean
in-degree: p_primary = 100 / N (use_case's 0.2 at N = 500), which keeps every
size in the same operating regime (the reference's scale_1000.json does the
same; fixed p would saturate the large networks).
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np

from . import _capi as capi
from .analysis import analyze_calcium
from .network import Batch, DeviceNetwork, Model


@dataclass
class NetSpec:
    index: int
    n: int
    n_exc: int
    n_inh: int
    n_astro: int
    w_n2a: float
    seed: int


def sweep_specs(count: int = 256, n_min: int = 200, n_max: int = 5000, gen_seed: int = 4,
                astro_ratio=(0.2, 0.2), w_range=(0.10, 0.40), seed0: int = 1000):
    rng = np.random.default_rng(gen_seed)
    specs = []
    for i in range(count):
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        ratio = rng.uniform(*astro_ratio) if astro_ratio[0] != astro_ratio[1] else astro_ratio[0]
        w = float(rng.uniform(*w_range))
        n_exc = int(round(0.8 * n))
        specs.append(NetSpec(i, n, n_exc, n - n_exc, max(1, int(round(ratio * n))), w, seed0 + i))
    return specs


def build(base_config: str, spec: NetSpec, sim_time_ms: float, pre_time_ms: float,
          in_degree: float = 100.0) -> Model:
    assignments = {
        "N.N_E": spec.n_exc, "N.N_I": spec.n_inh, "N.N_A": spec.n_astro,
        "synapse.w_n2a": spec.w_n2a,
        "connectivity.p_primary": min(1.0, in_degree / spec.n),
    }
    return Model(base_config, seed=spec.seed, sim_time=sim_time_ms, pre_time=pre_time_ms,
                 assignments=assignments)


@dataclass
class RunResult:
    elapsed_s: float
    neuron_steps: int
    spikes: list      # per network: int64 steps of its spikes
    traces: list      # per network: (steps, Ca[r, n])


def _record_plan(model: Model, record_astro_max: int = 100):
    info = model.info
    n_rec = min(record_astro_max, info["n_astro"])
    every = info["record_every"]
    total = info["pre_steps"] + info["sim_steps"]
    cap = (total - info["pre_steps"]) // every + 2
    return info["pre_steps"], every, n_rec, cap, total


def run_batched(models, precision: int, chunk_steps: int = 10000, budget_bytes: int = 0,
                profile: bool = False, schedule: dict | None = None) -> RunResult:
    """All networks in one batch (launch schedule `schedule`, keyword
    arguments of Batch.set_schedule, if given); returns after every network
    finished."""
    batch = Batch(len(models), budget_bytes)
    if schedule:
        batch.set_schedule(**schedule)
    slots = []
    totals = []
    for m in models:
        m.desc.contents.precision = precision
        start, every, n_rec, cap, total = _record_plan(m)
        s = batch.add(m, total)
        batch.set_recording(s, start, every, n_rec, cap)
        slots.append(s)
        totals.append(total)
    if profile:
        batch.set_profiling(True)
    spikes = [[] for _ in models]
    longest = max(totals)
    t0 = time.perf_counter()
    done = 0
    while done < longest:
        batch.step(chunk_steps)
        done += chunk_steps
        drained = batch.spikes_all()
        for k, s in enumerate(slots):
            spikes[k].append(drained[s][0])
    elapsed = time.perf_counter() - t0
    traces = []
    for s in slots:
        st, ca, _ = batch.read_trace(s)
        traces.append((st, ca))
    res = RunResult(elapsed, sum(m.n_neurons * t for m, t in zip(models, totals)),
                    [np.concatenate(x) for x in spikes], traces)
    res.kernel_times = batch.kernel_times() if profile else None
    batch.close()
    return res


def run_plain(models, chunk_steps: int = 10000) -> RunResult:
    """One network at a time on the device (the plain drop-in path)."""
    spikes, traces = [], []
    elapsed = 0.0
    neuron_steps = 0
    for m in models:
        m.desc.contents.precision = capi.ASG_FP64
        start, every, n_rec, cap, total = _record_plan(m)
        net = DeviceNetwork(m)
        net.set_recording(start, every, n_rec, cap)
        parts = []
        t0 = time.perf_counter()
        done = 0
        while done < total:
            k = min(chunk_steps, total - done)
            net.step(k)
            done += k
            st, _ = net.spikes()
            parts.append(st)
        elapsed += time.perf_counter() - t0
        st, ca, _ = net.read_trace()
        spikes.append(np.concatenate(parts))
        traces.append((st, ca))
        neuron_steps += m.n_neurons * total
        net.close()
    return RunResult(elapsed, neuron_steps, spikes, traces)


def network_metrics(model: Model, spike_steps: np.ndarray, trace) -> dict:
    """Rate over the recorded window, active-astrocyte fraction and mean
    pairwise Ca correlation, as the reference reports them."""
    info = model.info
    pre = info["pre_steps"]
    rec = spike_steps[spike_steps >= pre]
    rate = len(rec) * 1000.0 / (info["sim_time"] * model.n_neurons) if info["sim_time"] > 0 else 0.0
    steps, ca = trace
    times = steps.astype(np.float64) * info["dt"]
    thr = info["ca_threshold"] if info["ca_threshold"] >= 0 else info["SIC_th"]
    a = analyze_calcium(times, ca, thr, info["merge_gap"], info["pre_sim_time"], info["sim_time"],
                        info["hist_window"], info["hist_shift"])
    n_rec = ca.shape[1] if ca.ndim == 2 else 0
    return {"rate_hz": rate, "active": a["active"], "cells": n_rec,
            "active_frac": a["active"] / n_rec if n_rec else 0.0,
            "corr": a["mean_correlation"], "transients": a["transients"]}


def compare(models, ref: RunResult, test: RunResult, rate_tol=0.01, corr_tol=0.02,
            regime_corr=0.1) -> dict:
    rows = []
    for k, m in enumerate(models):
        a = network_metrics(m, ref.spikes[k], ref.traces[k])
        b = network_metrics(m, test.spikes[k], test.traces[k])
        rel = abs(b["rate_hz"] - a["rate_hz"]) / a["rate_hz"] if a["rate_hz"] > 0 else abs(b["rate_hz"])
        rows.append({"n": m.n_neurons, "rate_ref": a["rate_hz"], "rate": b["rate_hz"],
                     "rate_rel_err": rel, "active_ref": a["active_frac"], "active": b["active_frac"],
                     "corr_ref": a["corr"], "corr": b["corr"],
                     "corr_abs_err": abs(b["corr"] - a["corr"]),
                     "regime_ref": a["corr"] > regime_corr, "regime": b["corr"] > regime_corr})
    rate_err = np.array([r["rate_rel_err"] for r in rows])
    corr_err = np.array([r["corr_abs_err"] for r in rows])
    act_err = np.array([abs(r["active"] - r["active_ref"]) for r in rows])
    tot_ref = sum(r["rate_ref"] * r["n"] for r in rows)
    tot = sum(r["rate"] * r["n"] for r in rows)
    return {
        "networks": len(rows),
        "rate_rel_err_median": float(np.median(rate_err)),
        "rate_rel_err_p90": float(np.percentile(rate_err, 90)),
        "rate_rel_err_max": float(rate_err.max()),
        "frac_rate_within_tol": float(np.mean(rate_err < rate_tol)),
        "pooled_rate_rel_err": abs(tot - tot_ref) / tot_ref if tot_ref > 0 else 0.0,
        "corr_abs_err_median": float(np.median(corr_err)),
        "corr_abs_err_max": float(corr_err.max()),
        "frac_corr_within_tol": float(np.mean(corr_err < corr_tol)),
        "active_frac_abs_err_mean": float(act_err.mean()),
        "regimes_ref": int(sum(r["regime_ref"] for r in rows)),
        "regimes": int(sum(r["regime"] for r in rows)),
        "regime_agreement": float(np.mean([r["regime"] == r["regime_ref"] for r in rows])),
        "rows": rows,
    }
