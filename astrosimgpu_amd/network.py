"""Python face of the C ABI: build a network from a reference JSON config and
step it on the device.

Mirrors the reference's host interface (src/network.cpp, network.hpp:58-67):
``Model`` is ``load_config`` + ``Network::build`` (host side), ``DeviceNetwork``
is the device-resident step loop of ``Network::run`` (network.cpp:385-468)
with the spike raster and state read back on demand, ``Batch`` holds many
independent networks stepped in shared launches.
"""

from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _capi as capi
from ._capi import check, lib


def _overrides(seed=-1, sim_time=-1.0, pre_time=-1.0, n_astro=-1, assignments=None):
    ov = capi.ModelOverrides()
    ov.seed = int(seed)
    ov.sim_time = float(sim_time)
    ov.pre_time = float(pre_time)
    ov.n_astro = int(n_astro)
    ov.assignments = assignments.encode() if assignments else None
    return ov


class Model:
    """A configured and built network on the host.

    ``config`` is a path to a reference-format JSON file, or pass ``text``.
    ``assignments`` is a dict or "a.b=v;c.d=w" string of JSON overrides applied
    before the file is read (e.g. {"synapse.w_n2a": 0.3, "N.N_E": 800}).
    """

    def __init__(self, config: Optional[str] = None, *, text: Optional[str] = None,
                 seed: int = -1, sim_time: float = -1.0, pre_time: float = -1.0,
                 n_astro: int = -1, assignments=None):
        if isinstance(assignments, dict):
            assignments = ";".join(f"{k}={_json_literal(v)}" for k, v in assignments.items())
        ov = _overrides(seed, sim_time, pre_time, n_astro, assignments)
        handle = C.c_void_p()
        if text is not None:
            st = lib().asg_model_load_text(text.encode(), C.byref(ov), C.byref(handle))
        else:
            if config is None:
                raise ValueError("give a config path or text")
            st = lib().asg_model_load(str(config).encode(), C.byref(ov), C.byref(handle))
        check(st)
        self._h = handle
        info = capi.ModelInfo()
        check(lib().asg_model_get_info(self._h, C.byref(info)))
        self.info = {name: getattr(info, name) for name, _ in capi.ModelInfo._fields_}
        self.info["output_dir"] = info.output_dir.decode()

    @property
    def desc(self):
        """Pointer to the asg_net_desc (valid while this object lives)."""
        return lib().asg_model_desc(self._h)

    @property
    def n_neurons(self) -> int:
        return self.info["n_exc"] + self.info["n_inh"]

    @property
    def n_astro(self) -> int:
        return self.info["n_astro"]

    def stats(self) -> dict:
        keys = ("n_astro", "n_exc", "n_inh", "n_primary_exc", "n_primary_inh",
                "n_neuron_to_astro", "n_astro_to_neuron")
        return {k: self.info[k] for k in keys}

    def csr(self, name: str):
        """(row_start, target, weight, delay_steps) of one set as numpy copies."""
        d = self.desc.contents
        c = getattr(d, name)
        rows = np.ctypeslib.as_array(c.row_start, shape=(c.n_rows + 1,)).copy()
        nnz = int(rows[-1])
        tgt = (np.ctypeslib.as_array(c.target, shape=(nnz,)).copy() if nnz
               else np.zeros(0, np.uint32))
        return rows, tgt, c.weight, c.delay_steps

    def footprint(self) -> int:
        return int(lib().asg_query_footprint(self.desc))

    def prepare(self) -> None:
        """The host-side part of admitting this network to a batch
        (asg_net_prepare: validation, derived connectivity, device plan),
        done now, on the calling thread; the next Batch.add / try_add then
        only allocates and uploads. Change nothing in the desc afterwards."""
        self.release_prepared()
        handle = C.c_void_p()
        check(lib().asg_net_prepare(self.desc, C.byref(handle)))
        self._prepared = handle

    def release_prepared(self) -> None:
        if getattr(self, "_prepared", None):
            lib().asg_prepared_free(self._prepared)
            self._prepared = None

    def close(self):
        self.release_prepared()
        if getattr(self, "_h", None):
            lib().asg_model_free(self._h)
            self._h = None

    def __del__(self):
        self.close()


def _json_literal(v) -> str:
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, str):
        return '"' + v + '"'
    return repr(float(v)) if isinstance(v, float) else str(v)


class DeviceNetwork:
    """One network whose whole step runs on the device (FP64)."""

    def __init__(self, model: Model):
        self.model = model  # keeps the host arrays alive while the desc is read
        self._h = C.c_void_p()
        check(lib().asg_net_create(model.desc, C.byref(self._h)))

    def step(self, n_steps: int, wait: bool = True) -> None:
        """Advance n_steps; with wait=False only queue the work (see sync)."""
        if wait:
            check(lib().asg_net_step(self._h, int(n_steps)))
        else:
            check(lib().asg_net_step_async(self._h, int(n_steps)))

    def sync(self) -> None:
        check(lib().asg_net_sync(self._h))

    @property
    def current_step(self) -> int:
        return int(lib().asg_net_current_step(self._h))

    def spikes(self, capacity: int = 1 << 20):
        """Drain the recorded raster: (steps int64, sources uint32)."""
        out_s, out_n = [], []
        step = np.empty(capacity, np.int64)
        src = np.empty(capacity, np.uint32)
        n = C.c_int64()
        while True:
            check(lib().asg_net_spikes(self._h, step.ctypes.data_as(capi.I64P),
                                       src.ctypes.data_as(capi.U32P), capacity, C.byref(n)))
            out_s.append(step[:n.value].copy())
            out_n.append(src[:n.value].copy())
            if n.value < capacity:
                break
        return np.concatenate(out_s), np.concatenate(out_n)

    def state(self, field: str, offset: int = 0, count: Optional[int] = None) -> np.ndarray:
        size = self.model.n_astro if field in capi.ASTRO_FIELDS else self.model.n_neurons
        if count is None:
            count = size - offset
        out = np.empty(count, np.float64)
        check(lib().asg_net_read_state(self._h, capi.FIELDS[field], offset, count,
                                       out.ctypes.data_as(capi.DBLP)))
        return out

    def counters(self) -> dict:
        c = capi.Counters()
        check(lib().asg_net_counters(self._h, C.byref(c)))
        return {name: getattr(c, name) for name, _ in capi.Counters._fields_}

    def set_profiling(self, on: bool, every: int = 1) -> None:
        """Time the window kernels with HIP events, one window in `every`."""
        check(lib().asg_net_set_profiling(self._h, int(every) if on else 0))

    def set_graphs(self, on: bool) -> None:
        check(lib().asg_net_set_graphs(self._h, 1 if on else 0))

    def set_recording(self, start_step: int, every: int, n_astro: int, capacity: int) -> None:
        """Record Ca and IP3 of astrocytes [0, n_astro) every `every` steps from
        start_step on, into a device ring of `capacity` records."""
        check(lib().asg_net_set_recording(self._h, int(start_step), int(every), int(n_astro),
                                          int(capacity)))
        self._rec_n = int(n_astro)

    def read_trace(self, max_records: int = 1 << 16):
        """Drain recorded samples: (steps int64[r], Ca[r, n], IP3[r, n])."""
        return _read_trace(lambda out, st, cap, n: lib().asg_net_read_trace(
            self._h, out, st, cap, n), self._rec_n, max_records)

    def kernel_times(self) -> dict:
        """Device milliseconds per kernel, summed over the timed windows since
        profiling was switched on."""
        t = capi.KernelTimes()
        check(lib().asg_net_kernel_times(self._h, C.byref(t)))
        return {"cells": t.ms[0], "compact": t.ms[1], "deliver": t.ms[2], "sum": t.ms[3],
                "drive": t.ms[4], "windows": t.windows}

    def close(self):
        if getattr(self, "_h", None):
            lib().asg_net_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class Batch:
    """Independent networks stepped together; slots can be added and removed."""

    def __init__(self, max_nets: int, budget_bytes: int = 0):
        self._h = C.c_void_p()
        check(lib().asg_batch_create(int(max_nets), int(budget_bytes), C.byref(self._h)))
        self.max_nets = int(max_nets)
        self.models = {}
        self._rec_n = {}
        self._drain = None  # staging arrays of spikes_all, kept between calls

    def _add(self, model: Model, total_steps: int, slot) -> int:
        prep = getattr(model, "_prepared", None)
        if prep:
            return lib().asg_batch_add_prepared(self._h, model.desc, prep, int(total_steps),
                                                C.byref(slot))
        return lib().asg_batch_add(self._h, model.desc, int(total_steps), C.byref(slot))

    def add(self, model: Model, total_steps: int) -> int:
        """Admit a network (with the work of Model.prepare done beforehand
        if it was called)."""
        slot = C.c_int32()
        check(self._add(model, total_steps, slot))
        model.release_prepared()
        self.models[slot.value] = model
        return slot.value

    def try_add(self, model: Model, total_steps: int) -> Optional[int]:
        """Like add, but returns None when the batch has no room now."""
        slot = C.c_int32()
        st = self._add(model, total_steps, slot)
        if st == capi.ASG_ERR_CAPACITY:
            return None
        check(st)
        model.release_prepared()
        self.models[slot.value] = model
        return slot.value

    def remove(self, slot: int) -> None:
        check(lib().asg_batch_remove(self._h, int(slot)))
        self.models.pop(slot, None)

    def step(self, n_steps: int, wait: bool = True) -> None:
        fn = lib().asg_batch_step if wait else lib().asg_batch_step_async
        check(fn(self._h, int(n_steps)))

    def net_step(self, slot: int) -> int:
        return int(lib().asg_batch_net_step(self._h, int(slot)))

    def spikes(self, slot: int, capacity: int = 1 << 20):
        out_s, out_n = [], []
        step = np.empty(capacity, np.int64)
        src = np.empty(capacity, np.uint32)
        n = C.c_int64()
        while True:
            check(lib().asg_batch_spikes(self._h, int(slot), step.ctypes.data_as(capi.I64P),
                                         src.ctypes.data_as(capi.U32P), capacity, C.byref(n)))
            out_s.append(step[:n.value].copy())
            out_n.append(src[:n.value].copy())
            if n.value < capacity:
                break
        return np.concatenate(out_s), np.concatenate(out_n)

    def spikes_all(self, capacity: int = 1 << 20) -> dict:
        """Drain every network's raster in one call (asg_batch_spikes_all):
        {slot: (steps int64, sources uint32)} for every occupied slot, each
        slot's pairs in the order asg_batch_spikes would return them."""
        return self.split_spikes(self.spikes_all_raw(capacity))

    def spikes_all_raw(self, capacity: int = 1 << 20) -> list:
        """The drain of spikes_all without the split per slot: a list of
        (starts, ends, steps, sources) rounds, each slot's pairs at
        [starts[slot], ends[slot]) of its round. split_spikes turns it into
        spikes_all's dict, e.g. while the next steps run."""
        if self._drain is None or len(self._drain[1]) != capacity:
            self._drain = (np.empty(self.max_nets, np.int64), np.empty(capacity, np.int64),
                           np.empty(capacity, np.uint32))
        counts, step, src = self._drain
        n = C.c_int64()
        rounds = []
        while True:
            check(lib().asg_batch_spikes_all(self._h, counts.ctypes.data_as(capi.I64P),
                                             step.ctypes.data_as(capi.I64P),
                                             src.ctypes.data_as(capi.U32P), capacity,
                                             C.byref(n)))
            # one right-sized copy per round; the slots' arrays are views of it
            ends = np.cumsum(counts)
            rounds.append(((ends - counts).tolist(), ends.tolist(), step[:n.value].copy(),
                           src[:n.value].copy()))
            if n.value < capacity:
                break
        return rounds

    def split_spikes(self, rounds: list) -> dict:
        out = {}
        for slot in self.models:
            parts_s, parts_n = [], []
            for starts, ends, st, sr in rounds:
                parts_s.append(st[starts[slot]:ends[slot]])
                parts_n.append(sr[starts[slot]:ends[slot]])
            if len(rounds) == 1:
                out[slot] = (parts_s[0], parts_n[0])
            else:
                out[slot] = (np.concatenate(parts_s), np.concatenate(parts_n))
        return out

    def state(self, slot: int, field: str, offset: int = 0, count: Optional[int] = None):
        m = self.models[slot]
        size = m.n_astro if field in capi.ASTRO_FIELDS else m.n_neurons
        if count is None:
            count = size - offset
        out = np.empty(count, np.float64)
        check(lib().asg_batch_read_state(self._h, int(slot), capi.FIELDS[field], offset, count,
                                         out.ctypes.data_as(capi.DBLP)))
        return out

    def counters(self, slot: int) -> dict:
        c = capi.Counters()
        check(lib().asg_batch_counters(self._h, int(slot), C.byref(c)))
        return {name: getattr(c, name) for name, _ in capi.Counters._fields_}

    def set_profiling(self, on: bool, every: int = 1) -> None:
        """Time the window kernels with HIP events, one window in `every`."""
        check(lib().asg_batch_set_profiling(self._h, int(every) if on else 0))

    def set_recording(self, slot: int, start_step: int, every: int, n_astro: int,
                      capacity: int) -> None:
        check(lib().asg_batch_set_recording(self._h, int(slot), int(start_step), int(every),
                                            int(n_astro), int(capacity)))
        self._rec_n[slot] = int(n_astro)

    def read_trace(self, slot: int, max_records: int = 1 << 16):
        return _read_trace(lambda out, st, cap, n: lib().asg_batch_read_trace(
            self._h, int(slot), out, st, cap, n), self._rec_n.get(slot, 0), max_records)

    def kernel_times(self) -> dict:
        t = capi.KernelTimes()
        check(lib().asg_batch_kernel_times(self._h, C.byref(t)))
        return {"cells": t.ms[0], "compact": t.ms[1], "deliver": t.ms[2], "sum": t.ms[3],
                "drive": t.ms[4], "windows": t.windows}

    def bytes_in_use(self) -> int:
        return int(lib().asg_batch_bytes_in_use(self._h))

    @staticmethod
    def _schedule_dict(sch) -> dict:
        return {"cell_waves": sch.cell_waves, "sic_group": sch.sic_group,
                "step_group": sch.step_group, "deliver_waves": sch.deliver_waves,
                "sic_workers": sch.sic_workers, "lanes": sch.lanes, "teams": sch.teams,
                "ms_per_window": sch.ms_per_window}

    def autotune(self, windows_per_try: int = 20, rounds: int = 3) -> dict:
        """Pick the fastest launch schedule on the live batch (advances every
        network; results do not depend on the schedule)."""
        sch = capi.Schedule()
        check(lib().asg_batch_autotune(self._h, int(windows_per_try), int(rounds), C.byref(sch)))
        return self._schedule_dict(sch)

    def schedule(self) -> dict:
        sch = capi.Schedule()
        check(lib().asg_batch_get_schedule(self._h, C.byref(sch)))
        return self._schedule_dict(sch)

    def set_schedule(self, cell_waves: int = 1, sic_group: int = 0, step_group: int = 0,
                     deliver_waves: int = 0, sic_workers: int = 0, lanes: int = 0,
                     teams: int = 0) -> None:
        """Launch geometry of the windows (results do not depend on it);
        teams > 0 runs each call's FP32 windows in one persistent launch of
        that many teams of networks."""
        sch = capi.Schedule(int(cell_waves), int(sic_group), int(step_group), int(deliver_waves),
                            int(sic_workers), int(lanes), int(teams), 0.0)
        check(lib().asg_batch_set_schedule(self._h, C.byref(sch)))

    def set_graphs(self, on: bool) -> None:
        check(lib().asg_batch_set_graphs(self._h, 1 if on else 0))

    def close(self):
        if getattr(self, "_h", None):
            lib().asg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def _read_trace(call, n_rec: int, max_records: int):
    steps_out, ca_out, ip3_out = [], [], []
    if n_rec == 0:
        return np.zeros(0, np.int64), np.zeros((0, 0)), np.zeros((0, 0))
    buf = np.empty((max_records, 2, n_rec), np.float64)
    st = np.empty(max_records, np.int64)
    n = C.c_int64()
    while True:
        check(call(buf.ctypes.data_as(capi.DBLP), st.ctypes.data_as(capi.I64P), max_records,
                   C.byref(n)))
        k = n.value
        steps_out.append(st[:k].copy())
        ca_out.append(buf[:k, 0, :].copy())
        ip3_out.append(buf[:k, 1, :].copy())
        if k < max_records:
            break
    return np.concatenate(steps_out), np.concatenate(ca_out), np.concatenate(ip3_out)


def device_math(fn: str, x: np.ndarray) -> np.ndarray:
    """Evaluate exp/log/sin/cos/sqrt of the device step, or the FP32 path's
    2^y of the spike current ("exp2_f32"), on the device; "div": x holds
    (a, b) pairs and the result's even entries are a / b by the FP64 step's
    division through the reciprocal of b."""
    codes = {"exp": 0, "log": 1, "sin": 2, "cos": 3, "sqrt": 4, "exp2_f32": 5, "div": 6}
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty_like(x)
    check(lib().asg_device_math(codes[fn], x.ctypes.data_as(capi.DBLP),
                                y.ctypes.data_as(capi.DBLP), x.size))
    return y
