"""Serving loop on the device: every job's raster is the same whether it ran
alone or in the mix, with jobs joining and leaving the batch between chunks
and the memory budget making some wait."""

import sys

import numpy as np
import pytest

import astrosimgpu_amd as A
import astrosimgpu_amd.serve as S
from astrosimgpu_amd import _capi as capi

from conftest import config_path

pytestmark = pytest.mark.gpu


def _announce(name):
    print(f"[gpu] {name}", flush=True)
    sys.stdout.flush()


def test_job_alone_equals_job_in_the_mix():
    _announce("serve alone vs mix")
    base = config_path("paper_sparse")
    jobs = S.job_stream(8, rate_per_s=200.0, seed=5, n_min=300, n_max=4000,
                        duration_ms=(100.0, 400.0))
    for k, j in enumerate(jobs):
        j.precision = capi.ASG_FP32 if k % 2 else capi.ASG_FP64
    fps = [S.build_job(base, j).footprint() for j in jobs]
    budget = int(2.5 * max(fps))
    srv = S.Server(base, max_slots=6, budget_bytes=budget, chunk_steps=730)
    rep = srv.run(jobs)
    srv.close()
    assert rep.jobs == len(jobs)
    assert rep.peak_bytes <= budget
    for j in jobs:
        m = S.build_job(base, j)
        net = A.DeviceNetwork(m)
        net.step(j.total_steps)
        s_alone, n_alone = net.spikes()
        net.close()
        s_mix, n_mix = j.raster()
        assert len(s_alone) > 0, j
        assert np.array_equal(s_alone, s_mix) and np.array_equal(n_alone, n_mix), \
            f"job {j.id} n={j.n} precision={j.precision}"
