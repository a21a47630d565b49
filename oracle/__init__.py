"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU oracle
(oracle/step_oracle.cpp) and of the reference build (oracle/_ref).

Imported only by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg, as the checker or the timed baseline; the product never loads it.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "astrosimgpu_ref")

_lib = None

LIBM_GLIBC = 0
LIBM_AMATH = 1


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `make -C oracle oracle`")
        h = C.CDLL(LIB_PATH)
        h.oracle_create.restype = C.c_void_p
        h.oracle_create.argtypes = [C.c_void_p, C.c_int]
        h.oracle_destroy.argtypes = [C.c_void_p]
        h.oracle_step.argtypes = [C.c_void_p, C.c_int64]
        h.oracle_current_step.restype = C.c_int64
        h.oracle_current_step.argtypes = [C.c_void_p]
        h.oracle_spike_count.restype = C.c_int64
        h.oracle_spike_count.argtypes = [C.c_void_p]
        h.oracle_spikes.restype = C.c_int64
        h.oracle_spikes.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int64]
        h.oracle_read_state.restype = C.c_int
        h.oracle_read_state.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        h.oracle_phase_seconds.argtypes = [C.c_void_p, C.c_void_p]
        h.oracle_math.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int64]
        h.oracle_astro_scaling.argtypes = [C.c_uint32, C.c_int32, C.c_uint64, C.c_void_p,
                                           C.c_void_p, C.c_double, C.c_double, C.c_int,
                                           C.c_double, C.c_double, C.c_double, C.c_void_p,
                                           C.c_void_p, C.c_void_p]
        _lib = h
    return _lib


FIELDS = {
    "V": 0, "w": 1, "g_ex": 2, "dg_ex": 3, "g_in": 4, "dg_in": 5, "I_sic": 6,
    "Ca": 7, "IP3": 8, "h": 9, "stp_x": 10, "stp_u": 11, "stp_t_last": 12,
}


class OracleRun:
    """The reference's step loop on the CPU for one built network.

    ``desc`` is the ctypes pointer from astrosimgpu_amd.Model.desc (the model
    must stay alive while this object lives).
    """

    def __init__(self, desc, libm: int = LIBM_AMATH):
        self._desc = desc
        d = desc.contents
        self.n_neurons = d.n_exc + d.n_inh
        self.n_astro = d.n_astro
        self._h = lib().oracle_create(C.cast(desc, C.c_void_p), libm)

    def step(self, n: int) -> None:
        lib().oracle_step(self._h, int(n))

    @property
    def current_step(self) -> int:
        return int(lib().oracle_current_step(self._h))

    def spikes(self):
        n = int(lib().oracle_spike_count(self._h))
        st = np.empty(n, np.int64)
        src = np.empty(n, np.uint32)
        if n:
            lib().oracle_spikes(self._h, 0, st.ctypes.data, src.ctypes.data, n)
        return st, src

    def state(self, field: str) -> np.ndarray:
        size = self.n_astro if field in ("Ca", "IP3", "h") else self.n_neurons
        out = np.empty(size, np.float64)
        if lib().oracle_read_state(self._h, FIELDS[field], out.ctypes.data) != 0:
            raise KeyError(field)
        return out

    def phase_seconds(self) -> dict:
        out = np.zeros(6)
        lib().oracle_phase_seconds(self._h, out.ctypes.data)
        names = ("arrival application", "background input", "update astrocytes",
                 "update neurons", "spike delivery", "SIC gather+deliver")
        return dict(zip(names, out.tolist()))

    def close(self):
        if getattr(self, "_h", None):
            lib().oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def math(fn: str, x: np.ndarray, libm: int) -> np.ndarray:
    codes = {"exp": 0, "log": 1, "sin": 2, "cos": 3, "sqrt": 4}
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty_like(x)
    lib().oracle_math(codes[fn], libm, x.ctypes.data, y.ctypes.data, x.size)
    return y


def astro_scaling(n, steps, seed, constants, nominal, dt, lam, randomized=False, lower=1.0,
                  upper=1.0, var=0.0):
    """The isolated astrocyte update at scale on the CPU (FP64); `constants`
    is a ctypes asg_astro_constants, `nominal` (Ca_tot, IP3_0, tau_IP3,
    delta_IP3). Returns the final (Ca, IP3, h)."""
    nom = np.ascontiguousarray(nominal, np.float64)
    ca, ip3, h = (np.empty(n) for _ in range(3))
    lib().oracle_astro_scaling(n, steps, seed, C.byref(constants), nom.ctypes.data, dt, lam,
                               1 if randomized else 0, lower, upper, var, ca.ctypes.data,
                               ip3.ctypes.data, h.ctypes.data)
    return ca, ip3, h
