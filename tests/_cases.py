"""Shared helpers for the parity tests: build a golden case's network, run the
oracle or the device on it, and fingerprint rasters the way the fixtures do."""

from __future__ import annotations

import hashlib

import numpy as np

from conftest import config_path, golden

import astrosimgpu_amd as A


def raster_hash(steps, neurons) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(steps, "<i8").tobytes())
    h.update(np.ascontiguousarray(neurons, "<u4").tobytes())
    return h.hexdigest()


def model_for(case: str) -> tuple:
    g = golden(case)
    m = A.Model(config_path(g["config"]), seed=g["seed"], sim_time=g["sim_time"],
                pre_time=g["pre_time"])
    return m, g


def recorded(steps, neurons, first_step):
    keep = steps >= first_step
    return steps[keep], neurons[keep]


def stats_match(m, g) -> bool:
    s = m.stats()
    return all(s[k] == g["stats"][k] for k in s)
