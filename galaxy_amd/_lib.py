"""Loader and ctypes signatures of libgalaxy_bh.so (include/galaxy_bh.h).

The library is built in-tree by galaxy_amd/build.py (or __graft_entry__.build()).
There is no fallback: if the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
# GALAXY_BH_LIB selects another build of the library (the host-sanitized one of
# galaxy_amd/build.py --sanitize, for CPU test runs).
LIB_PATH = os.environ.get("GALAXY_BH_LIB") or os.path.join(HERE, "libgalaxy_bh.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "galaxy_bh.h")

_lib = None


class GbhConfig(C.Structure):
    _fields_ = [("theta", C.c_double), ("softening", C.c_double), ("g", C.c_double),
                ("bucket", C.c_uint32), ("max_depth", C.c_uint32)]


class StageTiming(C.Structure):
    _fields_ = [("transfer_seconds", C.c_double), ("dispatch_seconds", C.c_double),
                ("readback_seconds", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class BuildTiming(C.Structure):
    _fields_ = [("bounds_seconds", C.c_double), ("morton_seconds", C.c_double),
                ("radix_seconds", C.c_double), ("positions_seconds", C.c_double),
                ("topology_seconds", C.c_double), ("aggregate_seconds", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class SimParams(C.Structure):
    _fields_ = [("config", GbhConfig), ("dt", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("sort_tile", C.c_uint32), ("traverse_block", C.c_uint32),
                ("traverse_waves_per_cu", C.c_uint32), ("use_graph", C.c_uint32), ("streams", C.c_uint32),
                ("topo_tile", C.c_uint32)]


class ServerOptions(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("steps_per_round", C.c_uint32),
                ("max_inflight", C.c_uint32), ("tuning", Tuning)]


class Completion(C.Structure):
    _fields_ = [("id", C.c_uint64), ("submit_time", C.c_double), ("admit_time", C.c_double),
                ("finish_time", C.c_double), ("n", C.c_uint32), ("steps", C.c_uint32),
                ("status", C.c_int32)]


def header_functions() -> list[str]:
    """Names of every function declared in include/galaxy_bh.h."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gbh_[a-z0-9_]+)\s*\(", text)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"{LIB_PATH} is not built; run galaxy_amd/build.py (the HIP path has no fallback)")
    _lib = C.CDLL(LIB_PATH)
    vp, u32, u64, dbl, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_double, C.c_int
    sz = C.c_size_t
    sig = {
        "gbh_last_error": ([], C.c_char_p),
        "gbh_device_open": ([C.c_char_p, i32, vp], i32),
        "gbh_device_close": ([vp], None),
        "gbh_device_info_json": ([vp, vp, sz, vp], i32),
        "gbh_state_create": ([vp, vp, u32, vp, vp], i32),
        "gbh_state_destroy": ([vp], None),
        "gbh_state_count": ([vp], u32),
        "gbh_state_bytes": ([vp], u64),
        "gbh_state_read": ([vp, vp, vp], i32),
        "gbh_accel_read": ([vp, vp, vp], i32),
        "gbh_kick_drift": ([vp, dbl, vp], i32),
        "gbh_final_kick": ([vp, dbl, vp], i32),
        "gbh_runtime_create": ([vp, vp], i32),
        "gbh_runtime_destroy": ([vp], None),
        "gbh_tree_create": ([vp, vp, vp], i32),
        "gbh_tree_create_large": ([vp, vp, vp], i32),
        "gbh_tree_destroy": ([vp], None),
        "gbh_tree_rebuild": ([vp, vp, vp, vp, vp], i32),
        "gbh_force": ([vp, vp, vp, vp, vp], i32),
        "gbh_force_evidence": ([vp, vp, vp, vp, vp, vp], i32),
        "gbh_force_evidence_noalloc": ([vp, vp, vp, vp, vp, vp], i32),
        "gbh_tree_counts": ([vp, vp, vp, vp], i32),
        "gbh_tree_read_evidence": ([vp, vp, vp, vp, u32, vp, vp, vp], i32),
        "gbh_tree_read_flat": ([vp, vp, vp, u32, vp, vp], i32),
        "gbh_tree_buffer_bytes": ([vp], u64),
        "gbh_tree_last_force_ms": ([vp], dbl),
        "gbh_check_rounding": ([vp, vp, vp], i32),
        "gbh_check_thresholds": ([vp, vp, vp, u32, vp], i32),
        "gbh_diag_direct": ([vp, vp, u32, vp, vp, vp], i32),
        "gbh_diag_energy": ([vp, dbl, dbl, vp], i32),
        "gbh_copy_calibration": ([vp, u64, vp], i32),
        "gbh_sparse_layout_bytes": ([u32], u64),
        "gbh_fnv_mix64": ([u64, vp, u64, vp], i32),
        "gbh_traverse_packed": ([vp, vp, vp, vp, vp, vp, vp], i32),
        "gbh_force_packed": ([vp, vp, vp, vp, vp, vp], i32),
        "gbh_host_flat_tree": ([u32, vp, vp, vp, vp, vp, vp, vp, vp, u32, vp], i32),
        "gbh_gpu_tree_create": ([vp, vp, vp], i32),
        "gbh_gpu_tree_destroy": ([vp], None),
        "gbh_gpu_tree_rebuild": ([vp, vp, vp, vp], i32),
        "gbh_gpu_tree_force": ([vp, vp, vp, vp], i32),
        "gbh_gpu_tree_read_evidence": ([vp, vp, vp, vp, u32, vp], i32),
        "gbh_gpu_tree_buffer_bytes": ([vp], u64),
        "gbh_batch_create": ([vp, u32, vp, vp, vp, vp, vp], i32),
        "gbh_batch_destroy": ([vp], None),
        "gbh_batch_step": ([vp, u32], i32),
        "gbh_batch_read": ([vp, u32, vp, vp], i32),
        "gbh_batch_force_evidence": ([vp, u32, vp, vp], i32),
        "gbh_batch_device_bytes": ([vp], u64),
        "gbh_batch_tree_counts": ([vp, u32, vp, vp], i32),
        "gbh_batch_profile_step": ([vp, vp, u32, vp], i32),
        "gbh_sim_device_bytes": ([u32], u64),
        "gbh_server_create": ([vp, vp, vp], i32),
        "gbh_server_destroy": ([vp], None),
        "gbh_server_now": ([vp], dbl),
        "gbh_server_submit": ([vp, vp, u32, vp, u32, vp], i32),
        "gbh_server_round": ([vp, vp, u32, vp], i32),
        "gbh_server_start": ([vp], i32),
        "gbh_server_stop": ([vp], i32),
        "gbh_server_poll": ([vp, vp, u32, vp, dbl], i32),
        "gbh_server_take_result": ([vp, u64, vp, vp], i32),
        "gbh_server_stats": ([vp, vp, vp, vp], i32),
        "gbh_ic_rotating_disc": ([u32, u64, dbl, dbl, dbl, dbl, dbl, dbl, vp], i32),
        "gbh_ic_uniform_collapse": ([u32, u64, dbl, dbl, vp], i32),
        "gbh_ic_binary_disc": ([u32, u32, u64, u64, dbl, dbl, dbl, dbl, dbl, dbl, dbl, vp], i32),
        "gbh_ic_reference_disc": ([u32, u64, dbl, dbl, vp], i32),
        "gbh_ic_reference_collision": ([u32, u64, vp], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(_lib, name, None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = res
    return _lib
