"""Launch schedules, the autotuner and graph replay change launch geometry
and host overhead only: every variant gives the bits of the default eager
run."""

import numpy as np
import pytest

import astrosimgpu_amd as A
from astrosimgpu_amd import _capi as capi

from conftest import config_path

pytestmark = pytest.mark.gpu

SPECS = [("quick", 3, capi.ASG_FP64), ("use_case", 4, capi.ASG_FP32), ("random_pool", 5, capi.ASG_FP64),
         ("quick", 6, capi.ASG_FP32)]


def _models():
    ms = []
    for cfg, seed, prec in SPECS:
        m = A.Model(config_path(cfg), seed=seed)
        m.desc.contents.precision = prec
        ms.append(m)
    return ms


def _run(configure, chunks=(700, 1000, 1000, 300), total=3000, drain_all=None):
    """Every chunk's spikes are drained (slot by slot, or all slots in one
    call with that capacity) and a state array read between the calls
    (eager work between graph replays)."""
    models = _models()
    b = A.Batch(8)
    slots = [b.add(m, total) for m in models]
    configure(b)
    parts = {s: ([], []) for s in slots}
    for c in chunks:
        b.step(c)
        drained = b.spikes_all(drain_all) if drain_all else None
        for s in slots:
            st, src = drained[s] if drain_all else b.spikes(s)
            parts[s][0].append(st)
            parts[s][1].append(src)
        b.state(slots[1], "V")
    out = []
    for s in slots:
        out.append((np.concatenate(parts[s][0]), np.concatenate(parts[s][1]), b.state(s, "V"),
                    b.state(s, "Ca")))
    b.close()
    return out


def _same(a, b):
    for (s1, n1, v1, c1), (s2, n2, v2, c2) in zip(a, b):
        assert np.array_equal(s1, s2) and np.array_equal(n1, n2)
        assert np.array_equal(v1.view(np.uint64), v2.view(np.uint64))
        assert np.array_equal(c1.view(np.uint64), c2.view(np.uint64))


def test_every_schedule_gives_the_same_bits():
    print("[gpu case] schedules", flush=True)
    ref = _run(lambda b: None)
    for waves in (1, 2, 4):
        for group in (1, 2, 4, 8, 16):
            for steps in (1, 3, 16):   # 3 does not divide the window; 16 clamps to it
                _same(ref, _run(lambda b, w=waves, g=group, sg=steps: b.set_schedule(w, g, sg)))
    # the delivery and SIC wavefront counts that share out the lists
    for dw in (64, 512, 4096):
        for sw in (64, 2048, 8192):
            _same(ref, _run(lambda b, d=dw, e=sw: b.set_schedule(1, 0, 0, d, e)))
    # the FP32 networks in one lane, and split over two streams
    for lanes in (1, 2):
        _same(ref, _run(lambda b, n=lanes: b.set_schedule(1, 0, 0, 0, 0, n)))


def test_a_lane_that_stops_and_runs_again_starts_from_empty_lists():
    """Calls alternate between one lane and two, each an odd number of
    10-step windows, so the second lane runs again on the window parity of
    its last window before it stopped (whose lists that lane's post pass had
    not emptied): the bits of a run that never changed lanes."""
    print("[gpu case] lanes stop and resume", flush=True)
    chunks = (30, 10, 30, 10, 30, 10, 30, 50)
    total = sum(chunks)
    ref = _run(lambda b: None, chunks=chunks, total=total)
    models = _models()
    b = A.Batch(8)
    slots = [b.add(m, total) for m in models]
    parts = {s: ([], []) for s in slots}
    for k, c in enumerate(chunks):
        b.set_schedule(1, 0, 0, 0, 0, 2 if k % 2 == 0 else 1)
        b.step(c)
        for s in slots:
            st, src = b.spikes(s)
            parts[s][0].append(st)
            parts[s][1].append(src)
        b.state(slots[1], "V")
    out = [(np.concatenate(parts[s][0]), np.concatenate(parts[s][1]), b.state(s, "V"),
            b.state(s, "Ca")) for s in slots]
    b.close()
    _same(ref, out)


def test_graph_replays_equal_eager_launches():
    """The same 1000-step call replayed from one graph several times, with
    eager work (raster drains, state reads) and syncs between the replays;
    mixed FP64 / FP32 networks (both cell kernels in the graph)."""
    print("[gpu case] graphs", flush=True)
    ref = _run(lambda b: None, chunks=(1000, 1000, 1000))
    _same(ref, _run(lambda b: b.set_graphs(True), chunks=(1000, 1000, 1000)))

    def one_lane_graphs(b):
        b.set_schedule(1, 0, 0, 0, 0, 1)
        b.set_graphs(True)
    _same(ref, _run(one_lane_graphs, chunks=(1000, 1000, 1000)))
    # uneven chunks: several graphs (one per window count)
    ref = _run(lambda b: None)
    _same(ref, _run(lambda b: b.set_graphs(True)))


def test_drain_of_all_slots_equals_slot_by_slot():
    """asg_batch_spikes_all: the same pairs per slot as asg_batch_spikes,
    also when a small capacity splits the drain over several calls with
    slots cut in the middle (4 occupied slots of 8)."""
    print("[gpu case] drain all", flush=True)
    ref = _run(lambda b: None)
    assert sum(len(x[0]) for x in ref) > 200  # spikes enough to split
    for cap in (1 << 20, 97):
        _same(ref, _run(lambda b: None, drain_all=cap))


def test_autotune_then_step_equals_plain_run():
    print("[gpu case] autotune", flush=True)
    models = _models()
    b = A.Batch(8)
    slots = [b.add(m, 4000) for m in models]
    chosen = b.autotune(windows_per_try=1, rounds=2)
    assert chosen["cell_waves"] in (1, 2, 4) and chosen["sic_group"] in (1, 2, 4, 8, 16)
    assert chosen["step_group"] >= 1
    assert chosen["deliver_waves"] in (512, 1024, 2048)
    assert chosen["sic_workers"] == 0
    assert chosen["lanes"] in (1, 2)
    assert chosen["ms_per_window"] > 0.0
    done = b.net_step(slots[0])
    assert 0 < done < 4000 and all(b.net_step(s) == done for s in slots)
    b.set_graphs(True)
    b.step(4000 - done)
    tuned = [(b.spikes(s)[0], b.state(s, "V")) for s in slots]
    b.close()
    plain = A.Batch(8)
    ps = [plain.add(m, 4000) for m in _models()]
    plain.step(4000)
    for (st, v), s in zip(tuned, ps):
        assert np.array_equal(st, plain.spikes(s)[0])
        assert np.array_equal(v.view(np.uint64), plain.state(s, "V").view(np.uint64))
    plain.close()
