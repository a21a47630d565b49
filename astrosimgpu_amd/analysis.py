"""Calcium analysis of recorded astrocyte traces, through the C++ host
(reference src/analysis.cpp and its use in src/main.cpp:266-317)."""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as capi
from ._capi import check, lib


def analyze_calcium(times_ms, ca, threshold, merge_gap=2000.0, pre_sim_time=0.0,
                    sim_time=None, hist_window=2000.0, hist_shift=400.0, round6=False) -> dict:
    """ca: [samples, cells]. Returns the fields the reference prints under
    "Calcium analysis": active cells, transients, mean/sd duration, pairs and
    mean pairwise correlation (plus the window actually used)."""
    t = np.ascontiguousarray(times_ms, np.float64)
    c = np.ascontiguousarray(ca, np.float64)
    if sim_time is None:
        sim_time = float(t[-1] - pre_sim_time) if len(t) else 0.0
    out = capi.CalciumSummary()
    check(lib().asg_analyze_calcium(t.ctypes.data_as(capi.DBLP), c.ctypes.data_as(capi.DBLP),
                                    c.shape[1] if c.ndim == 2 else 0, len(t), float(threshold),
                                    float(merge_gap), float(pre_sim_time), float(sim_time),
                                    float(hist_window), float(hist_shift), 1 if round6 else 0,
                                    C.byref(out)))
    return {name: getattr(out, name) for name, _ in capi.CalciumSummary._fields_}


def analyze_calcium_cells(times, ca, threshold, merge_gap=2000.0, pre_sim_time=0.0,
                          sim_time=0.0, hist_window=2000.0, hist_shift=400.0,
                          round6=False) -> dict:
    """The same over per-cell traces of their own lengths: times[c], ca[c]
    are cell c's sample times and values."""
    offsets = np.zeros(len(ca) + 1, np.int64)
    offsets[1:] = np.cumsum([len(x) for x in ca])
    t = np.ascontiguousarray(np.concatenate(times) if len(times) else np.zeros(0), np.float64)
    c = np.ascontiguousarray(np.concatenate(ca) if len(ca) else np.zeros(0), np.float64)
    out = capi.CalciumSummary()
    check(lib().asg_analyze_calcium_cells(t.ctypes.data_as(capi.DBLP), c.ctypes.data_as(capi.DBLP),
                                          offsets.ctypes.data_as(capi.I64P), len(ca),
                                          float(threshold), float(merge_gap), float(pre_sim_time),
                                          float(sim_time), float(hist_window), float(hist_shift),
                                          1 if round6 else 0, C.byref(out)))
    return {name: getattr(out, name) for name, _ in capi.CalciumSummary._fields_}


def detect_events(times, values, threshold, merge=False, max_gap=0.0):
    t = np.ascontiguousarray(times, np.float64)
    v = np.ascontiguousarray(values, np.float64)
    cap = max(len(t), 1)
    on = np.empty(cap)
    off = np.empty(cap)
    n = C.c_int64()
    check(lib().asg_detect_events(t.ctypes.data_as(capi.DBLP), v.ctypes.data_as(capi.DBLP),
                                  len(t), float(threshold), 1 if merge else 0, float(max_gap),
                                  on.ctypes.data_as(capi.DBLP), off.ctypes.data_as(capi.DBLP),
                                  cap, C.byref(n)))
    return on[:n.value].copy(), off[:n.value].copy()


def pearson(a, b) -> float:
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    n = min(len(a), len(b))
    return float(lib().asg_pearson(a.ctypes.data_as(capi.DBLP), b.ctypes.data_as(capi.DBLP), n))


def sliding_window_rate(events, start, end, width, shift):
    e = np.ascontiguousarray(events, np.float64)
    cap = int(max((end - start) / max(shift, 1e-12), 0)) + 2
    out = np.empty(cap)
    n = C.c_int64()
    check(lib().asg_sliding_window_rate(e.ctypes.data_as(capi.DBLP), len(e), float(start),
                                        float(end), float(width), float(shift),
                                        out.ctypes.data_as(capi.DBLP), cap, C.byref(n)))
    return out[:n.value].copy()
