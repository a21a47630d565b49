"""Serving loop scheduling on the CPU: a stand-in for the device batch keeps
the admission rule of asg_batch_add (a free slot, and the job's footprint
within the budget left) so arrival order, refusals and retirement can be
checked without a GPU. Footprints are the library's own asg_query_footprint
figures of the built networks."""

import concurrent.futures as cf

import numpy as np

import astrosimgpu_amd.serve as S

from conftest import config_path


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, s):
        self.t += s


class InlineExecutor:
    """Builds at submission, so the schedule depends on the fake clock only."""

    def submit(self, fn, *args):
        f = cf.Future()
        f.set_result(fn(*args))
        return f

    def shutdown(self, wait=True):
        pass


class FakeBatch:
    """Device batch stand-in; a chunk of n steps costs n * 20 us of the clock."""

    def __init__(self, max_slots, budget, clock):
        self.max_slots, self.budget, self.clock = max_slots, budget, clock
        self.nets = {}          # slot -> [step, target, bytes]
        self.log = []           # (event, slot, bytes_in_use after)

    def try_add(self, model, total_steps):
        fp = model.footprint()
        if len(self.nets) >= self.max_slots or self.bytes_in_use() + fp > self.budget:
            return None
        slot = min(set(range(self.max_slots)) - set(self.nets))
        self.nets[slot] = [0, total_steps, fp]
        self.log.append(("add", slot, self.bytes_in_use()))
        return slot

    def remove(self, slot):
        del self.nets[slot]
        self.log.append(("remove", slot, self.bytes_in_use()))

    def step(self, n, wait=True):
        # the chunk covers the networks present at the call; ones added
        # while it "runs" (before the drain) start with the next chunk
        self.clock.t += n * 2e-5
        self.log.append(("step", len(self.nets), self.bytes_in_use()))
        for v in self.nets.values():
            v[0] = min(v[1], v[0] + n)

    def net_step(self, slot):
        return self.nets[slot][0]

    def spikes(self, slot, capacity=0):
        return np.zeros(0, np.int64), np.zeros(0, np.uint32)

    def spikes_all(self, capacity=0):
        return {slot: self.spikes(slot) for slot in self.nets}

    def bytes_in_use(self):
        return sum(v[2] for v in self.nets.values())

    def close(self):
        pass


def _jobs():
    return S.job_stream(10, rate_per_s=400.0, seed=3, n_min=200, n_max=3000,
                        duration_ms=(50.0, 300.0))


def test_job_stream_distribution():
    jobs = S.job_stream(2000, rate_per_s=4.0, seed=11)
    arr = np.array([j.arrival_s for j in jobs])
    gaps = np.diff(np.concatenate([[0.0], arr]))
    assert abs(gaps.mean() - 0.25) < 0.02          # Poisson arrivals at 4/s
    n = np.array([j.n for j in jobs])
    assert n.min() >= 500 and n.max() <= 100000
    assert abs(np.log(n).mean() - 0.5 * (np.log(500) + np.log(1e5))) < 0.15  # log-uniform
    ratio = np.array([j.n_astro / j.n for j in jobs])
    assert ratio.min() >= 0.19 and ratio.max() <= 1.01
    dur = np.array([j.duration_ms for j in jobs])
    assert dur.min() >= 1000 and dur.max() <= 10000
    assert len({j.seed for j in jobs}) == len(jobs)


def test_build_job_applies_the_job_shape():
    j = S.Job(0, 0.0, n=1234, n_astro=500, duration_ms=250.0, w_n2a=0.3, seed=5)
    m = S.build_job(config_path("paper_sparse"), j)
    assert m.n_neurons == 1234 and m.n_astro == 500
    assert m.info["sim_steps"] == 2500 and m.info["pre_steps"] == 0
    assert m.desc.contents.precision == S.capi.ASG_FP32
    st = m.stats()
    # in-degree about 100 from p_primary = 100 / N
    assert 80 < (st["n_primary_exc"] + st["n_primary_inh"]) / 1234 < 120


def test_admission_respects_budget_and_arrival_order():
    base = config_path("paper_sparse")
    jobs = _jobs()
    fps = [S.build_job(base, j).footprint() for j in jobs]
    budget = int(1.5 * max(fps))
    clk = FakeClock()
    fb = FakeBatch(4, budget, clk)
    srv = S.Server(base, chunk_steps=500, batch=fb, executor=InlineExecutor())
    rep = srv.run(jobs, clock=clk, sleep=clk.sleep)
    srv.close()
    assert rep.jobs == len(jobs)
    assert rep.refused_admissions > 0                     # the budget made jobs wait
    assert max(b for _, _, b in fb.log) <= budget
    assert rep.peak_running <= 4
    order = sorted(jobs, key=lambda j: j.admitted_s)
    assert [j.id for j in order] == sorted(j.id for j in jobs)   # first come, first admitted
    for j in jobs:
        assert j.arrival_s <= j.admitted_s <= j.done_s
    assert rep.neuron_steps == sum(j.n * j.total_steps for j in jobs)
    assert rep.latency_p50_s <= rep.latency_p99_s


def test_a_job_larger_than_the_budget_is_not_admitted_ahead_of_order():
    """Admission is first come first served: a waiting job blocks later ones
    (no overtaking), so a big job is never starved by small ones."""
    base = config_path("paper_sparse")
    jobs = _jobs()[:4]
    fps = [S.build_job(base, j).footprint() for j in jobs]
    clk = FakeClock()
    fb = FakeBatch(8, max(fps), clk)            # one job at a time at worst
    srv = S.Server(base, chunk_steps=400, batch=fb, executor=InlineExecutor())
    rep = srv.run(jobs, clock=clk, sleep=clk.sleep)
    srv.close()
    adds = [slot for ev, slot, _ in fb.log if ev == "add"]
    assert all(j.total_steps > 0 and fb.log for j in jobs)
    assert len(adds) == 4 and rep.jobs == 4
    ids = [j.id for j in sorted(jobs, key=lambda j: j.admitted_s)]
    assert ids == [0, 1, 2, 3]


def test_a_job_that_cannot_fit_an_empty_batch_is_an_error():
    base = config_path("paper_sparse")
    clk = FakeClock()
    srv = S.Server(base, batch=FakeBatch(4, 1000, clk), executor=InlineExecutor())
    try:
        srv.run(_jobs()[:1], clock=clk, sleep=clk.sleep)
    except RuntimeError as e:
        assert "does not fit" in str(e)
    else:
        raise AssertionError("expected an error")
    finally:
        srv.close()


def test_jobs_admitted_while_a_chunk_runs_join_the_next_one():
    """The loop launches a chunk, then takes in arrivals and admits jobs
    while it runs: every job still runs to its total in whole chunks (its
    last one cut at the total), and no chunk is launched with no job."""
    base = config_path("paper_sparse")
    jobs = _jobs()
    clk = FakeClock()
    fb = FakeBatch(16, 1 << 40, clk)
    srv = S.Server(base, chunk_steps=300, batch=fb, executor=InlineExecutor())
    rep = srv.run(jobs, clock=clk, sleep=clk.sleep)
    srv.close()
    assert rep.jobs == len(jobs)
    assert all(running > 0 for ev, running, _ in fb.log if ev == "step")
    for j in jobs:
        assert j.arrival_s <= j.admitted_s < j.done_s
    # chunks overlapped admission: some job was admitted between a chunk's
    # launch and its drain (an add logged right after a step)
    evs = [ev for ev, _, _ in fb.log]
    assert any(a == "step" and b == "add" for a, b in zip(evs, evs[1:]))
