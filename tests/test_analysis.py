"""Calcium analysis (C++ host, reference src/analysis.cpp and main.cpp:264-313)
against the reference build's own printed analysis of the same recorded
traces (fixtures written by tests/golden/make_analysis_golden.py), plus
closed-form checks of its building blocks."""

import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

from astrosimgpu_amd.analysis import (analyze_calcium, analyze_calcium_cells, detect_events,
                                     pearson, sliding_window_rate)


def _fixture(case):
    """Per-cell (times, Ca) exactly as the reference's analysis read them: the
    part of astrocytes.csv its still-open output stream had written (whole
    8192-byte chunks, the last line possibly cut), and its printed summary."""
    z = np.load(os.path.join(GOLDEN, f"analysis_{case}.npz"), allow_pickle=False)
    off = z["offsets"]
    t = np.cumsum(z["times_dq"]) / 1e3
    cq = z["ca_dq"].astype(np.int64)
    times, ca = [], []
    for c in range(len(off) - 1):
        times.append(t[off[c]:off[c + 1]])
        ca.append(np.cumsum(cq[off[c]:off[c + 1]]) / 1e8)
    with open(os.path.join(GOLDEN, f"analysis_{case}.json")) as f:
        return times, ca, json.load(f)


@pytest.mark.parametrize("case", ["use_case_60s", "bursting_20s", "use_case_short"])
def test_calcium_analysis_equals_reference_printout(case):
    times, ca, e = _fixture(case)
    r = analyze_calcium_cells(times, ca, e["threshold"], e["merge_gap"], e["pre_time"],
                              e["sim_time"], e["hist_window"], e["hist_shift"])
    assert r["cells"] == e["cells"]
    assert r["active"] == e["active"]
    assert r["transients"] == e["transients"]
    assert r["pairs"] == e["pairs"]
    assert r["width"] == e["width"] and r["shift"] == e["shift"]
    # the reference prints 4 decimals
    assert round(r["mean_duration"], 4) == pytest.approx(e["mean_duration"], abs=1e-9)
    assert round(r["sd_duration"], 4) == pytest.approx(e["sd_duration"], abs=1e-9)
    assert round(r["mean_correlation"], 4) == pytest.approx(e["mean_correlation"], abs=1e-9)


def test_detect_events_square_pulses():
    t = np.arange(0.0, 100.0, 1.0)
    v = np.zeros_like(t)
    v[10:20] = 1.0   # above at t = 10 .. 19: onset 10, offset (last sample above) 19
    v[25:30] = 1.0   # 6 ms after the first: merged when the gap is below max_gap
    v[70:75] = 1.0
    on, off = detect_events(t, v, 0.5)
    assert on.tolist() == [10.0, 25.0, 70.0] and off.tolist() == [19.0, 29.0, 74.0]
    on, off = detect_events(t, v, 0.5, merge=True, max_gap=6.0)   # gap 6 is not below 6
    assert on.tolist() == [10.0, 25.0, 70.0] and off.tolist() == [19.0, 29.0, 74.0]
    on, off = detect_events(t, v, 0.5, merge=True, max_gap=7.0)
    assert on.tolist() == [10.0, 70.0] and off.tolist() == [29.0, 74.0]
    # an event running at either edge is kept
    v2 = np.ones(10)
    on, off = detect_events(np.arange(10.0), v2, 0.5)
    assert on.tolist() == [0.0] and off.tolist() == [9.0]


def test_pearson_closed_forms():
    a = np.arange(50.0)
    assert pearson(a, 3.0 * a + 2.0) == pytest.approx(1.0, abs=1e-14)
    assert pearson(a, -a) == pytest.approx(-1.0, abs=1e-14)
    assert pearson(a, np.ones(50)) == 0.0          # no variance
    rng = np.random.default_rng(0)
    x, y = rng.normal(size=1000), rng.normal(size=1000)
    assert pearson(x, y) == pytest.approx(np.corrcoef(x, y)[0, 1], abs=1e-12)


def test_sliding_window_rate_counts_events():
    """floor((end - start - width) / shift) windows [lo, lo + width], both
    ends included (analysis.cpp:66-96); events per ms."""
    ev = np.array([1.0, 2.0, 3.0, 11.0])
    assert sliding_window_rate(ev, 0.0, 20.0, 10.0, 10.0).tolist() == [pytest.approx(0.3)]
    r = sliding_window_rate(ev, 0.0, 30.0, 10.0, 10.0)
    assert r.tolist() == [pytest.approx(0.3), pytest.approx(0.1)]
    r = sliding_window_rate(np.array([10.0]), 0.0, 30.0, 10.0, 10.0)
    assert r.tolist() == [pytest.approx(0.1), pytest.approx(0.1)]   # the shared edge counts twice


def test_matrix_and_per_cell_forms_agree():
    """Equal-length traces through the [samples][cells] entry point give the
    per-cell entry point's result, with and without 6-digit rounding."""
    times, ca, e = _fixture("bursting_20s")
    n = min(len(x) for x in ca)
    t = times[0][:n]
    mat = np.stack([x[:n] for x in ca], axis=1) * (1.0 + 1e-9)
    for round6 in (False, True):
        a = analyze_calcium(t, mat, e["threshold"], e["merge_gap"], e["pre_time"], e["sim_time"],
                            e["hist_window"], e["hist_shift"], round6=round6)
        b = analyze_calcium_cells([t] * mat.shape[1], [mat[:, c] for c in range(mat.shape[1])],
                                  e["threshold"], e["merge_gap"], e["pre_time"], e["sim_time"],
                                  e["hist_window"], e["hist_shift"], round6=round6)
        assert a == b


REF_DRIVER = os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref", "analysis_driver")


def _reference_analysis(times, ca, params, tmp_path):
    """The reference's analysis functions (its analysis.cpp behind
    oracle/analysis_driver.cpp) on the same per-cell traces, at full
    precision."""
    import subprocess
    off = np.concatenate([[0], np.cumsum([len(x) for x in ca])]).astype(np.int64)
    path = tmp_path / "traces.bin"
    with open(path, "wb") as f:
        f.write(np.int32(len(ca)).tobytes())
        f.write(np.array(params, np.float64).tobytes())
        f.write(off.tobytes())
        f.write(np.concatenate([np.zeros(0)] + list(times)).astype(np.float64).tobytes())
        f.write(np.concatenate([np.zeros(0)] + list(ca)).astype(np.float64).tobytes())
    out = subprocess.run([REF_DRIVER, str(path)], check=True, capture_output=True,
                         text=True).stdout.split()
    return {"active": int(out[0]), "transients": int(out[1]), "pairs": int(out[2]),
            "mean_duration": float.fromhex(out[3]), "sd_duration": float.fromhex(out[4]),
            "mean_correlation": float.fromhex(out[5]), "width": float.fromhex(out[6]),
            "shift": float.fromhex(out[7])}


def _random_traces(rng, n_cells, threshold):
    times, ca = [], []
    for c in range(n_cells):
        kind = rng.integers(0, 6)
        n = {0: 0, 1: 1}.get(kind, int(rng.integers(2, 400)))
        t = np.cumsum(rng.uniform(0.5, 1.5, n)) + rng.uniform(0, 5)
        if kind == 2:       # values exactly at the threshold in places
            v = np.where(rng.random(n) < 0.4, threshold, threshold - 0.01)
        elif kind == 3:     # a plain pulse train with gaps of exactly 5 samples
            v = np.where((np.arange(n) // 5) % 2 == 0, threshold + 0.1, threshold - 0.1)
            t = np.arange(n, dtype=np.float64)
        else:               # a random walk around the threshold
            v = threshold + np.cumsum(rng.normal(0, 0.02, n))
        times.append(t)
        ca.append(v)
    return times, ca


@pytest.mark.skipif(not os.path.exists(REF_DRIVER), reason="reference build absent")
@pytest.mark.parametrize("seed", range(12))
def test_calcium_analysis_bit_equals_reference_functions(seed, tmp_path):
    """Random traces with empty and one-sample cells, values exactly at the
    threshold, gaps exactly max_gap and the short-window fallback: every
    field of the product's analysis equals the reference's, bit for bit."""
    rng = np.random.default_rng(seed)
    threshold = 0.2
    times, ca = _random_traces(rng, int(rng.integers(0, 30)), threshold)
    sim = float(rng.choice([50.0, 200.0, 400.0]))
    params = [threshold, float(rng.choice([0.0, 2.0, 5.0, 6.0, 20.0])), float(rng.uniform(0, 20)),
              sim, float(rng.choice([20.0, 100.0, 1000.0])), float(rng.choice([2.0, 5.0, 10.0]))]
    ref = _reference_analysis(times, ca, params, tmp_path)
    got = analyze_calcium_cells(times, ca, *params)
    for k, v in ref.items():
        if isinstance(v, float):
            assert float(got[k]).hex() == v.hex(), (k, got[k], v)
        else:
            assert got[k] == v, (k, got[k], v)
