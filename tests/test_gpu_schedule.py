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


def _models(all_fp64=False):
    ms = []
    for cfg, seed, prec in SPECS:
        m = A.Model(config_path(cfg), seed=seed)
        m.desc.contents.precision = capi.ASG_FP64 if all_fp64 else prec
        ms.append(m)
    return ms


def _run(configure, chunks=(700, 1000, 1000, 300), total=3000, drain_all=None, all_fp64=False):
    """Every chunk's spikes are drained (slot by slot, or all slots in one
    call with that capacity) and a state array read between the calls
    (eager work between graph replays)."""
    models = _models(all_fp64)
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


def test_fp64_groups_give_the_same_bits():
    """An FP64-only batch split into one, two or three groups of networks,
    each on a stream of its own, eager and from graphs, and with the group
    count changed between calls: the bits of the one-group run."""
    print("[gpu case] FP64 groups", flush=True)
    ref = _run(lambda b: b.set_schedule(lanes=1), all_fp64=True)
    for lanes in (2, 3):
        _same(ref, _run(lambda b, n=lanes: b.set_schedule(lanes=n), all_fp64=True))

        def with_graphs(b, n=lanes):
            b.set_schedule(lanes=n)
            b.set_graphs(True)
        _same(ref, _run(with_graphs, all_fp64=True))
    models = _models(True)
    b = A.Batch(8)
    slots = [b.add(m, 3000) for m in models]
    parts = {s: ([], []) for s in slots}
    for k, c in enumerate((700, 1000, 1000, 300)):
        b.set_schedule(lanes=(1, 3, 2, 1)[k])
        b.step(c)
        for s in slots:
            st, src = b.spikes(s)
            parts[s][0].append(st)
            parts[s][1].append(src)
    out = [(np.concatenate(parts[s][0]), np.concatenate(parts[s][1]), b.state(s, "V"),
            b.state(s, "Ca")) for s in slots]
    b.close()
    _same(ref, out)


def test_persistent_teams_give_the_same_bits():
    """Each call's FP32 windows in one persistent cooperative launch: the two
    FP32 networks in one team or one team each (a count above the network
    count is capped), behind the mixed batch's FP64 networks' own kernels,
    with the delivery and SIC wavefront counts varied and graph replay asked
    for (the FP64 networks' graphs; the persistent launch is not captured)."""
    print("[gpu case] persistent teams", flush=True)
    ref = _run(lambda b: None)
    for teams in (1, 2, 5):
        _same(ref, _run(lambda b, t=teams: b.set_schedule(teams=t)))
    for dw, sw in ((64, 64), (4096, 8192)):
        _same(ref, _run(lambda b, d=dw, e=sw: b.set_schedule(1, 0, 0, d, e, 0, 2)))

    def teams_and_graphs(b):
        b.set_schedule(teams=2)
        b.set_graphs(True)
    _same(ref, _run(teams_and_graphs))


def test_switching_between_teams_lanes_and_graphs_keeps_the_bits():
    """Calls alternate among a persistent launch of two teams, two stream
    lanes with graphs, one team and one lane, each call an odd number of
    windows (list parities carry over between launch modes)."""
    print("[gpu case] teams / lanes switches", flush=True)
    chunks = (30, 10, 30, 10, 30, 10, 30, 50)
    total = sum(chunks)
    ref = _run(lambda b: None, chunks=chunks, total=total)
    modes = [dict(teams=2), dict(lanes=2), dict(teams=1), dict(lanes=1)]
    models = _models()
    b = A.Batch(8)
    slots = [b.add(m, total) for m in models]
    parts = {s: ([], []) for s in slots}
    for k, c in enumerate(chunks):
        b.set_schedule(**modes[k % len(modes)])
        b.set_graphs(k % 4 == 1)
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


def test_persistent_sweep_equals_per_kernel_launches():
    """A 40-network FP32 sweep (200-5000 neurons), 0.3 s: the persistent
    launch with 8 teams and with one team per network gives every network's
    raster and recorded Ca of the default per-kernel launches, bit for bit."""
    print("[gpu case] persistent sweep", flush=True)
    from astrosimgpu_amd import sweep
    specs = sweep.sweep_specs(count=40, gen_seed=5)
    models = [sweep.build(config_path("use_case"), s, 200.0, 100.0) for s in specs]
    ref = sweep.run_batched(models, capi.ASG_FP32, chunk_steps=1000)
    for teams in (8, 40):
        got = sweep.run_batched(models, capi.ASG_FP32, chunk_steps=1000,
                                schedule=dict(teams=teams))
        for k in range(len(models)):
            assert np.array_equal(ref.spikes[k], got.spikes[k]), (teams, k)
            assert np.array_equal(ref.traces[k][1].view(np.uint64),
                                  got.traces[k][1].view(np.uint64)), (teams, k)


def test_a_team_barrier_that_times_out_fails_the_call(monkeypatch):
    """The persistent launch's barrier wait is bounded: with the bound set to
    zero from the host the first barrier gives up, every workgroup leaves
    the kernel, and the call returns ASG_ERR_DEVICE naming the barrier."""
    print("[gpu case] team barrier time-out", flush=True)
    monkeypatch.setenv("ASG_TEAM_WAIT_US", "0")
    b = A.Batch(8)
    for m in _models():
        b.add(m, 3000)
    b.set_schedule(teams=2)
    with pytest.raises(A.AsgError) as e:
        b.step(500)
    assert e.value.status == capi.ASG_ERR_DEVICE
    assert "team barrier" in str(e.value)
    b.close()


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
    assert chosen["lanes"] in (1, 2, 3)
    assert chosen["teams"] in (0, 2)  # two FP32 networks: at most two teams
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
