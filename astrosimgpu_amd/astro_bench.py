"""Astrocyte-update kernel scaling: 1e4 to 1e7 Li-Rinzel cells in SoA arrays,
one update launch per step reading state and an int32 presynaptic spike count
(Poisson, mean 0.05) from HBM, 1,000 steps, FP64 and FP32 from the same
initial state.

Reported per size and precision: achieved HBM GB/s of the update launches
(algorithmic bytes per launch x launches / their summed device time from HIP
events) against MI355X's 8 TB/s, and after the run the maximum and mean
relative difference of FP32 Ca, IP3 and h from FP64.

Constants and nominal per-cell parameters are use_case.json's astrocyte block
(its randomize_astro block for the randomised variant).

    python -m astrosimgpu_amd.astro_bench [--sizes 1e4,1e5,1e6,1e7] [--steps 1000]
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys

import numpy as np

from . import _capi as capi
from ._capi import check, lib
from .network import Model

HBM_PEAK_GBS = 8000.0
SEED = 20261203
USE_CASE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests",
                        "golden", "configs", "use_case.json")


def setup(config: str = USE_CASE, randomized: bool = False, lam: float = 0.05):
    """asg_astro_constants and asg_astro_scaling_params from a configuration's
    astrocyte block (nominal values: the built model's first cell when
    randomisation is off in the file, else the file's means)."""
    import json as _json

    def strip(t):
        return "\n".join(ln for ln in t.splitlines() if not ln.lstrip().startswith("//"))

    with open(config) as f:
        cfg = _json.loads(strip(f.read()))
    a = cfg["astrocyte"]
    m = Model(config, assignments={"N.N_A": 1, "N.N_E": 4, "N.N_I": 1,
                                   "randomize_astro.enabled": False})
    d = m.desc.contents
    constants = capi.AstroConstants()
    C.memmove(C.byref(constants), C.byref(d.astro.c), C.sizeof(capi.AstroConstants))
    p = capi.AstroScalingParams()
    p.Ca_tot, p.IP3_0, p.tau_IP3, p.delta_IP3 = a["Ca_tot"], a["IP3_0"], a["tau_IP3"], a["delta_IP3"]
    p.dt = m.info["dt"]
    p.lambda_ = lam
    r = cfg.get("randomize_astro", {})
    p.randomized = 1 if randomized else 0
    p.lower, p.upper, p.var = r.get("lower", 1.0), r.get("upper", 1.0), r.get("var", 0.0)
    return constants, p


def run(n: int, steps: int, precision: int, constants, params, seed: int = SEED,
        want_state: bool = True):
    ca = np.empty(n) if want_state else None
    ip3 = np.empty(n) if want_state else None
    h = np.empty(n) if want_state else None
    res = capi.AstroScalingResult()
    ptr = (lambda x: x.ctypes.data_as(capi.DBLP) if x is not None else None)
    check(lib().asg_astro_scaling(n, steps, precision, C.byref(constants), C.byref(params), seed,
                                  ptr(ca), ptr(ip3), ptr(h), C.byref(res)))
    gbs = res.bytes_per_step * steps / (res.update_ms * 1e-3) / 1e9 if res.update_ms > 0 else 0.0
    return {"update_ms": res.update_ms, "total_ms": res.total_ms,
            "bytes_per_launch": res.bytes_per_step, "gbs": gbs,
            "us_per_launch": 1000.0 * res.update_ms / max(steps, 1)}, (ca, ip3, h)


def rel_err(a, b):
    d = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    return float(d.max()), float(d.mean())


def scaling(sizes, steps: int, randomized: bool = False, echo: bool = False):
    constants, params = setup(randomized=randomized)
    rows = []
    for n in sizes:
        r64, s64 = run(n, steps, capi.ASG_FP64, constants, params)
        r32, s32 = run(n, steps, capi.ASG_FP32, constants, params)
        err = {name: rel_err(b, a) for name, a, b in zip(("Ca", "IP3", "h"), s64, s32)}
        rows.append({
            "n": n, "steps": steps, "randomized": randomized,
            "fp64": {"GB/s": round(r64["gbs"], 1), "frac": round(r64["gbs"] / HBM_PEAK_GBS, 4),
                     "us_per_launch": round(r64["us_per_launch"], 3),
                     "bytes_per_launch": r64["bytes_per_launch"]},
            "fp32": {"GB/s": round(r32["gbs"], 1), "frac": round(r32["gbs"] / HBM_PEAK_GBS, 4),
                     "us_per_launch": round(r32["us_per_launch"], 3),
                     "bytes_per_launch": r32["bytes_per_launch"]},
            "fp32_vs_fp64_rel_err": {k: {"max": v[0], "mean": v[1]} for k, v in err.items()},
        })
        if echo:
            print(json.dumps(rows[-1]), flush=True)
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    ap.add_argument("--sizes", default="1e4,1e5,1e6,1e7")
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--randomized", action="store_true")
    args = ap.parse_args(argv)
    sizes = [int(float(x)) for x in args.sizes.split(",")]
    scaling(sizes, args.steps, args.randomized, echo=True)


if __name__ == "__main__":
    sys.exit(main())
