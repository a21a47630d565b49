"""ctypes view of include/asg.h and the loader of libastrosim_amd.so.

The shared library is built in-tree by ``make -C astrosimgpu_amd`` (or
``__graft_entry__.build()``). Importing this module fails loudly when it is
missing: there is no fallback implementation of the step.
"""

from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libastrosim_amd.so")

# asg_status
ASG_OK = 0
ASG_ERR_INVALID = 1
ASG_ERR_NOMEM = 2
ASG_ERR_DEVICE = 3
ASG_ERR_OVERFLOW = 4
ASG_ERR_CONFIG = 5
ASG_ERR_CAPACITY = 6

ASG_FP64 = 0
ASG_FP32 = 1

FIELDS = {
    "V": 0, "w": 1, "g_ex": 2, "dg_ex": 3, "g_in": 4, "dg_in": 5, "I_sic": 6,
    "Ca": 7, "IP3": 8, "h": 9, "stp_x": 10, "stp_u": 11, "stp_t_last": 12,
}
ASTRO_FIELDS = ("Ca", "IP3", "h")


class NeuronParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "C_m", "g_L", "E_L", "V_th", "Delta_T", "a", "b", "tau_w", "V_reset", "V_peak",
        "t_ref", "E_ex", "E_in", "tau_syn_ex", "tau_syn_in", "I_e", "V_m_init")]


class InputParams(C.Structure):
    _fields_ = [("poiss_rate", C.c_double), ("poiss_weight", C.c_double),
                ("gauss_noise_std", C.c_double), ("noise_dt", C.c_double),
                ("independent_noise", C.c_int32)]


class AstroConstants(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "Kd_IP3_1", "Kd_IP3_2", "Kd_act", "Kd_inh", "Km_SERCA", "k_IP3R", "rate_IP3R",
        "rate_L", "rate_SERCA", "ratio_ER_cyt")]


class AstroParams(C.Structure):
    _fields_ = [("c", AstroConstants), ("SIC_th", C.c_double), ("SIC_scale", C.c_double),
                ("Ca_init", C.c_double), ("IP3_init", C.c_double), ("h_init", C.c_double)]


class Csr(C.Structure):
    _fields_ = [("n_rows", C.c_uint32), ("row_start", C.POINTER(C.c_uint32)),
                ("target", C.POINTER(C.c_uint32)), ("weight", C.c_double),
                ("delay_steps", C.c_int32)]


class Stp(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("U", C.c_double), ("tau_rec", C.c_double),
                ("tau_fac", C.c_double)]


class NetDesc(C.Structure):
    _fields_ = [
        ("n_exc", C.c_uint32), ("n_inh", C.c_uint32), ("n_astro", C.c_uint32),
        ("dt", C.c_double), ("substeps", C.c_int32), ("seed", C.c_uint64),
        ("neuron", NeuronParams * 2),
        ("V_reset", C.POINTER(C.c_double)), ("b", C.POINTER(C.c_double)),
        ("input_exc", InputParams), ("input_inh", InputParams), ("input_astro", InputParams),
        ("astro", AstroParams),
        ("Ca_tot", C.POINTER(C.c_double)), ("IP3_0", C.POINTER(C.c_double)),
        ("tau_IP3", C.POINTER(C.c_double)), ("delta_IP3", C.POINTER(C.c_double)),
        ("exc", Csr), ("inh", Csr), ("n2a", Csr), ("a2n", Csr),
        ("stp", Stp), ("sic_interval", C.c_int32), ("precision", C.c_int32),
    ]


class Counters(C.Structure):
    _fields_ = [("steps", C.c_int64), ("spikes", C.c_int64), ("synapse_events", C.c_int64),
                ("sic_entries", C.c_int64)]


class KernelTimes(C.Structure):
    _fields_ = [("ms", C.c_double * 5), ("windows", C.c_int64)]


class AstroScalingParams(C.Structure):
    _fields_ = [("Ca_tot", C.c_double), ("IP3_0", C.c_double), ("tau_IP3", C.c_double),
                ("delta_IP3", C.c_double), ("dt", C.c_double), ("lambda_", C.c_double),
                ("randomized", C.c_int32), ("lower", C.c_double), ("upper", C.c_double),
                ("var", C.c_double)]


class AstroScalingResult(C.Structure):
    _fields_ = [("update_ms", C.c_double), ("total_ms", C.c_double),
                ("bytes_per_step", C.c_double), ("steps", C.c_int64)]


class Schedule(C.Structure):
    _fields_ = [("cell_waves", C.c_int32), ("sic_group", C.c_int32), ("step_group", C.c_int32),
                ("deliver_waves", C.c_int32), ("sic_workers", C.c_int32),
                ("lanes", C.c_int32), ("teams", C.c_int32), ("ms_per_window", C.c_float)]


class CalciumSummary(C.Structure):
    _fields_ = [("cells", C.c_int32), ("active", C.c_int32), ("transients", C.c_int64),
                ("pairs", C.c_int64), ("mean_duration", C.c_double), ("sd_duration", C.c_double),
                ("mean_correlation", C.c_double), ("threshold", C.c_double),
                ("width", C.c_double), ("shift", C.c_double)]


class ModelOverrides(C.Structure):
    _fields_ = [("seed", C.c_int64), ("sim_time", C.c_double), ("pre_time", C.c_double),
                ("n_astro", C.c_int64), ("assignments", C.c_char_p)]


class ModelInfo(C.Structure):
    _fields_ = [
        ("n_astro", C.c_uint32), ("n_exc", C.c_uint32), ("n_inh", C.c_uint32),
        ("n_primary_exc", C.c_uint32), ("n_primary_inh", C.c_uint32),
        ("n_neuron_to_astro", C.c_uint32), ("n_astro_to_neuron", C.c_uint32),
        ("dt", C.c_double), ("pre_sim_time", C.c_double), ("sim_time", C.c_double),
        ("substeps", C.c_int32), ("pre_steps", C.c_int64), ("sim_steps", C.c_int64),
        ("seed", C.c_uint64),
        ("record_spikes", C.c_int32), ("record_astro", C.c_int32), ("record_neuron", C.c_int32),
        ("record_every", C.c_int32), ("record_astro_max", C.c_uint32),
        ("record_neuron_max", C.c_uint32),
        ("analysis_enabled", C.c_int32), ("ca_threshold", C.c_double), ("merge_gap", C.c_double),
        ("hist_window", C.c_double), ("hist_shift", C.c_double), ("SIC_th", C.c_double),
        ("ring_slots", C.c_int32), ("output_dir", C.c_char * 256),
    ]


P = C.c_void_p
PP = C.POINTER(C.c_void_p)
I64P = C.POINTER(C.c_int64)
U32P = C.POINTER(C.c_uint32)
DBLP = C.POINTER(C.c_double)

# name: (restype, argtypes)
PROTOTYPES = {
    "asg_last_error": (C.c_char_p, []),
    "asg_version": (C.c_char_p, []),
    "asg_model_load": (C.c_int, [C.c_char_p, C.POINTER(ModelOverrides), PP]),
    "asg_model_load_text": (C.c_int, [C.c_char_p, C.POINTER(ModelOverrides), PP]),
    "asg_model_free": (None, [P]),
    "asg_model_desc": (C.POINTER(NetDesc), [P]),
    "asg_model_get_info": (C.c_int, [P, C.POINTER(ModelInfo)]),
    "asg_astro_create": (P, [C.c_uint32, DBLP, DBLP, DBLP, DBLP, DBLP, DBLP, DBLP]),
    "asg_astro_destroy": (None, [P, DBLP, DBLP, DBLP]),
    "asg_astro_push_input": (None, [P, DBLP]),
    "asg_astro_pull_calcium": (None, [P, DBLP]),
    "asg_astro_update": (None, [P, C.POINTER(AstroConstants), C.c_double, C.c_int32,
                                C.c_double, C.c_int32, C.c_double, C.c_uint64, C.c_uint64]),
    "asg_net_create": (C.c_int, [C.POINTER(NetDesc), PP]),
    "asg_net_destroy": (C.c_int, [P]),
    "asg_net_step": (C.c_int, [P, C.c_int64]),
    "asg_net_step_async": (C.c_int, [P, C.c_int64]),
    "asg_net_sync": (C.c_int, [P]),
    "asg_net_current_step": (C.c_int64, [P]),
    "asg_net_spikes": (C.c_int, [P, I64P, U32P, C.c_int64, I64P]),
    "asg_net_read_state": (C.c_int, [P, C.c_int32, C.c_int64, C.c_int64, DBLP]),
    "asg_net_counters": (C.c_int, [P, C.POINTER(Counters)]),
    "asg_net_set_recording": (C.c_int, [P, C.c_int64, C.c_int32, C.c_uint32, C.c_int64]),
    "asg_net_read_trace": (C.c_int, [P, DBLP, I64P, C.c_int64, I64P]),
    "asg_batch_set_recording": (C.c_int, [P, C.c_int32, C.c_int64, C.c_int32, C.c_uint32,
                                           C.c_int64]),
    "asg_batch_read_trace": (C.c_int, [P, C.c_int32, DBLP, I64P, C.c_int64, I64P]),
    "asg_query_footprint": (C.c_int64, [C.POINTER(NetDesc)]),
    "asg_batch_create": (C.c_int, [C.c_int32, C.c_int64, PP]),
    "asg_batch_destroy": (C.c_int, [P]),
    "asg_batch_add": (C.c_int, [P, C.POINTER(NetDesc), C.c_int64, C.POINTER(C.c_int32)]),
    "asg_net_prepare": (C.c_int, [C.POINTER(NetDesc), C.POINTER(P)]),
    "asg_prepared_free": (None, [P]),
    "asg_batch_add_prepared": (C.c_int, [P, C.POINTER(NetDesc), P, C.c_int64,
                                         C.POINTER(C.c_int32)]),
    "asg_batch_remove": (C.c_int, [P, C.c_int32]),
    "asg_batch_step": (C.c_int, [P, C.c_int64]),
    "asg_batch_step_async": (C.c_int, [P, C.c_int64]),
    "asg_batch_net_step": (C.c_int64, [P, C.c_int32]),
    "asg_batch_spikes": (C.c_int, [P, C.c_int32, I64P, U32P, C.c_int64, I64P]),
    "asg_batch_spikes_all": (C.c_int, [P, I64P, I64P, U32P, C.c_int64, I64P]),
    "asg_batch_read_state": (C.c_int, [P, C.c_int32, C.c_int32, C.c_int64, C.c_int64, DBLP]),
    "asg_batch_counters": (C.c_int, [P, C.c_int32, C.POINTER(Counters)]),
    "asg_batch_bytes_in_use": (C.c_int64, [P]),
    "asg_device_math": (C.c_int, [C.c_int32, DBLP, DBLP, C.c_int64]),
    "asg_device_math_noalloc": (C.c_int, [C.c_int32, DBLP, DBLP, C.c_int64]),
    "asg_analyze_calcium": (C.c_int, [DBLP, DBLP, C.c_int32, C.c_int64, C.c_double, C.c_double,
                                      C.c_double, C.c_double, C.c_double, C.c_double, C.c_int32,
                                      C.POINTER(CalciumSummary)]),
    "asg_analyze_calcium_cells": (C.c_int, [DBLP, DBLP, I64P, C.c_int32, C.c_double,
                                            C.c_double, C.c_double, C.c_double, C.c_double,
                                            C.c_double, C.c_int32, C.POINTER(CalciumSummary)]),
    "asg_astro_scaling": (C.c_int, [C.c_uint32, C.c_int32, C.c_int32, C.POINTER(AstroConstants),
                                    C.POINTER(AstroScalingParams),
                                    C.c_uint64, DBLP, DBLP, DBLP, C.POINTER(AstroScalingResult)]),
    "asg_astro_scaling_bytes": (C.c_int64, [C.c_uint32, C.c_int32, C.c_int32]),
    "asg_batch_set_schedule": (C.c_int, [P, C.POINTER(Schedule)]),
    "asg_batch_get_schedule": (C.c_int, [P, C.POINTER(Schedule)]),
    "asg_batch_autotune": (C.c_int, [P, C.c_int32, C.c_int32, C.POINTER(Schedule)]),
    "asg_batch_set_graphs": (C.c_int, [P, C.c_int32]),
    "asg_net_set_graphs": (C.c_int, [P, C.c_int32]),
    "asg_detect_events": (C.c_int, [DBLP, DBLP, C.c_int64, C.c_double, C.c_int32, C.c_double,
                                    DBLP, DBLP, C.c_int64, I64P]),
    "asg_pearson": (C.c_double, [DBLP, DBLP, C.c_int64]),
    "asg_sliding_window_rate": (C.c_int, [DBLP, C.c_int64, C.c_double, C.c_double, C.c_double,
                                          C.c_double, DBLP, C.c_int64, I64P]),
    "asg_net_set_profiling": (C.c_int, [P, C.c_int32]),
    "asg_net_kernel_times": (C.c_int, [P, C.POINTER(KernelTimes)]),
    "asg_batch_set_profiling": (C.c_int, [P, C.c_int32]),
    "asg_batch_kernel_times": (C.c_int, [P, C.POINTER(KernelTimes)]),
}


class AsgError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"asg status {status}: {message}")
        self.status = status


_lib = None


def lib() -> C.CDLL:
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `make -C astrosimgpu_amd` "
                "(or __graft_entry__.build()); there is no fallback implementation")
        handle = C.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status: int) -> None:
    if status != ASG_OK:
        raise AsgError(status, lib().asg_last_error().decode(errors="replace"))
