"""ctypes declarations of the C ABI in include/smp_hip.h.

Loads pactae_amd/_lib/libsmp_hip.so (built in-tree by pactae_amd.build.build()).
There is no fallback: if the library is missing or fails to load, importing the
package raises, so nothing can silently run without the HIP code.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libsmp_hip.so")

SMP_OK = 0
SMP_ERR_INVALID = 1
SMP_ERR_HIP = 2
SMP_ERR_NOMEM = 3
SMP_ERR_STATE = 4
SMP_ERR_TIMEOUT = 5
SMP_PENDING = 6
SMP_MAX_N = 65535

SMP_TAIL_AUTO = 0
SMP_TAIL_GPU = 1
SMP_TAIL_HOST = 2

i32 = ctypes.c_int32
i64 = ctypes.c_int64
u64 = ctypes.c_uint64
i32p = ctypes.POINTER(ctypes.c_int32)
i64p = ctypes.POINTER(ctypes.c_int64)
vp = ctypes.c_void_p


class SmpOpts(ctypes.Structure):
    _fields_ = [
        ("prefix_cols", i32),
        ("tail_mode", i32),
        ("tail_threshold", i32),
        ("device", i32),
        ("precheck", i32),
        ("waves_per_cu", i32),
        ("validate", i32),
        ("coherent_hint", i32),
        ("poll_interval", i32),
        ("host_spin", i32),
        ("prefix_copy", i32),
        ("schedule", i32),
        ("init_mode", i32),
        ("tail_window_entries", i32),
        ("start_order", i32),
        ("reserved", i32 * 1),
    ]


class SmpTimes(ctypes.Structure):
    _fields_ = [
        ("precheck_ms", ctypes.c_double),
        ("init_ms", ctypes.c_double),
        ("core_ms", ctypes.c_double),
        ("post_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double),
        ("rank_kernel_ms", ctypes.c_double),
        ("gather_kernel_ms", ctypes.c_double),
        ("propose_kernel_ms", ctypes.c_double),
        ("handoff_ms", ctypes.c_double),
        ("host_tail_ms", ctypes.c_double),
        ("copy_wait_ms", ctypes.c_double),
        ("warm_ms", ctypes.c_double),
        ("resume_ms", ctypes.c_double),
        ("bucket_kernel_ms", ctypes.c_double),
        ("proposals", i64),
        ("atomics", i64),
        ("host_proposals", i64),
        ("steps", i64),
        ("is_perfect", i32),
        ("handoff_free", i32),
        ("prefix_cols", i32),
        ("tail_threshold", i32),
        ("bailed", i32),
        ("master_list", i32),
        ("schedule", i32),
        ("sparse_init", i32),
        ("rank_matrix", i32),
        ("copy_requeues", i32),
        ("copy_serialized", i32),
        ("tail_windows", i32),
        ("tail_full_matrix", i32),
        ("tail_rounds", i32),
        ("tail_bands", i32),
        ("deep_restart", i32),
        ("master_both", i32),
        ("reserved", i32 * 1),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_ if not f.startswith("reserved")}


class SmpInitTimes(ctypes.Structure):
    _fields_ = [
        ("rank_ms", ctypes.c_double),
        ("pr_ms", ctypes.c_double),
        ("transfer_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double),
        ("rank_kernel_ms", ctypes.c_double),
        ("pr_kernel_ms", ctypes.c_double),
        ("pr_cols", i32),
        ("reserved", i32 * 7),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_ if not f.startswith("reserved")}


class SmpTuneResult(ctypes.Structure):
    _fields_ = [
        ("prefix_cols", i32),
        ("tail_threshold", i32),
        ("best_ms", ctypes.c_double),
        ("gpu_only_ms", ctypes.c_double),
        ("candidates", i32),
        ("schedule", i32),
        ("waves_per_cu", i32),
        ("reserved", i32 * 5),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_ if not f.startswith("reserved")}


class SmpBatchOpts(ctypes.Structure):
    _fields_ = [
        ("device", i32),
        ("use_graph", i32),
        ("block_threads", i32),
        ("wgs_per_cu", i32),
        ("mode", i32),
        ("group_lanes", i32),
        ("pool_slots", i32),
        ("rank_layout", i32),
        ("reserved", i32 * 8),
    ]


class SmpBatchTimes(ctypes.Structure):
    _fields_ = [
        ("exec_ms", ctypes.c_double),
        ("kernel_ms", ctypes.c_double),
        ("solves_per_s", ctypes.c_double),
        ("proposals", i64),
        ("count", i32),
        ("block_threads", i32),
        ("wgs_per_cu", i32),
        ("mode", i32),
        ("group_lanes", i32),
        ("pool_slots", i32),
        ("rank_layout", i32),
        ("rank_ms", ctypes.c_double),
        ("reserved", i32 * 4),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_ if not f.startswith("reserved")}


class SmpServerOpts(ctypes.Structure):
    _fields_ = [
        ("device", i32),
        ("mode", i32),
        ("lanes", i32),
        ("max_batch", i32),
        ("max_n", i32),
        ("block_threads", i32),
        ("hbm_budget_bytes", i64),
        ("large_n", i32),
        ("large_workers", i32),
        ("large_spill_us", i32),
        ("small_to_idle_worker", i32),
        ("fused_lanes", i32),
        ("reserved", i32 * 3),
    ]


class SmpRequestInfo(ctypes.Structure):
    _fields_ = [
        ("queue_ms", ctypes.c_double),
        ("service_ms", ctypes.c_double),
        ("latency_ms", ctypes.c_double),
        ("proposals", i64),
        ("n", i32),
        ("batch_size", i32),
        ("lane", i32),
        ("reserved", i32 * 5),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_ if not f.startswith("reserved")}


class SmpServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", i64),
        ("completed", i64),
        ("launches", i64),
        ("arena_bytes", i64),
        ("peak_arena_bytes", i64),
        ("peak_inflight", i32),
        ("peak_lanes_busy", i32),
        ("large_solved", i32),
        ("reserved", i32 * 7),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_ if not f.startswith("reserved")}


SMP_SCHEDULE_BAMBOO = 0
SMP_SCHEDULE_MW_CAS = 1
SMP_SCHEDULE_LA_CAS = 2
SMP_SCHEDULE_MW_MIN = 3
SMP_SCHEDULE_LA_MIN = 4
SMP_SCHEDULE_FRONTIER = 5
SMP_INIT_AUTO = 0
SMP_INIT_FULL = 1
SMP_INIT_SPARSE = 2
SMP_BATCH_AUTO = 0
SMP_BATCH_FUSED = 1
SMP_BATCH_POOL = 2
SMP_RANK_WOMEN_ROWS = 1
SMP_RANK_MEN_ROWS = 2
SMP_SERVE_BATCHED = 0
SMP_SERVE_PLAIN = 1

# (name, restype, argtypes) for every entry point declared in include/smp_hip.h.
SIGNATURES = [
    ("smp_generate", ctypes.c_int, [i32, i32, u64, i32, i32p, i32p, i32]),
    ("smp_workload_save", ctypes.c_int, [ctypes.c_char_p, i32, i32p, i32p]),
    ("smp_workload_load", ctypes.c_int, [ctypes.c_char_p, i32p, i32p, i32p, ctypes.c_size_t]),
    ("smp_workload_filename", ctypes.c_int, [i32, i32, i32, ctypes.c_char_p, ctypes.c_size_t]),
    ("smp_default_opts", None, [ctypes.POINTER(SmpOpts)]),
    ("smp_create", ctypes.c_int, [i32p, i32p, i32, ctypes.POINTER(SmpOpts), ctypes.POINTER(vp)]),
    ("smp_create_device", ctypes.c_int, [vp, vp, i32, ctypes.POINTER(SmpOpts), ctypes.POINTER(vp)]),
    ("smp_create_ranked", ctypes.c_int, [i32p, i32p, i32, ctypes.POINTER(SmpOpts), ctypes.POINTER(vp)]),
    ("smp_create_device_ranked", ctypes.c_int, [vp, vp, i32, ctypes.POINTER(SmpOpts), ctypes.POINTER(vp)]),
    ("smp_exec", ctypes.c_int, [vp]),
    ("smp_result", ctypes.c_int, [vp, i32p]),
    ("smp_result_device", ctypes.c_int, [vp, ctypes.POINTER(vp)]),
    ("smp_profile", ctypes.c_int, [vp, ctypes.POINTER(SmpTimes)]),
    ("smp_init_exec", ctypes.c_int, [vp, i32, ctypes.POINTER(SmpInitTimes)]),
    ("smp_tune", ctypes.c_int, [vp, i32, ctypes.POINTER(SmpTuneResult)]),
    ("smp_rank_matrix", ctypes.c_int, [vp, i32p]),
    ("smp_pr_matrix", ctypes.c_int, [vp, i32p, i32, i32p]),
    ("smp_destroy", None, [vp]),
    ("smp_exec_device_noalloc", ctypes.c_int, [vp, vp, i32, i32, vp, i64p]),
    ("smp_batch_default_opts", None, [ctypes.POINTER(SmpBatchOpts)]),
    ("smp_batch_create", ctypes.c_int,
     [i32, i32p, i64p, i32p, i32p, ctypes.POINTER(SmpBatchOpts), ctypes.POINTER(vp)]),
    ("smp_batch_exec", ctypes.c_int, [vp]),
    ("smp_batch_exec_looped", ctypes.c_int, [vp]),
    ("smp_batch_result", ctypes.c_int, [vp, i32p]),
    ("smp_batch_profile", ctypes.c_int, [vp, ctypes.POINTER(SmpBatchTimes)]),
    ("smp_batch_destroy", None, [vp]),
    ("smp_batch_plan", ctypes.c_int,
     [i32, i32, i32, u64, ctypes.POINTER(ctypes.c_double), i32p, i32p]),
    ("smp_server_default_opts", None, [ctypes.POINTER(SmpServerOpts)]),
    ("smp_server_create", ctypes.c_int, [ctypes.POINTER(SmpServerOpts), ctypes.POINTER(vp)]),
    ("smp_server_submit", ctypes.c_int, [vp, vp, vp, i32, i32, ctypes.POINTER(u64)]),
    ("smp_server_poll", ctypes.c_int, [vp, u64, i32p, ctypes.POINTER(SmpRequestInfo)]),
    ("smp_server_wait", ctypes.c_int, [vp, u64, i32p, ctypes.POINTER(SmpRequestInfo), i32]),
    ("smp_server_stats_get", ctypes.c_int, [vp, ctypes.POINTER(SmpServerStats)]),
    ("smp_server_destroy", None, [vp]),
    ("smp_device_count", ctypes.c_int, [i32p]),
    ("smp_device_alloc", ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(vp)]),
    ("smp_device_free", ctypes.c_int, [vp]),
    ("smp_copy_h2d", ctypes.c_int, [vp, vp, ctypes.c_size_t]),
    ("smp_copy_d2h", ctypes.c_int, [vp, vp, ctypes.c_size_t]),
    ("smp_device_copy", ctypes.c_int, [vp, vp, ctypes.c_size_t]),
    ("smp_device_sync", ctypes.c_int, []),
    ("smp_last_error", ctypes.c_char_p, []),
    ("smp_version", ctypes.c_char_p, []),
]


def load() -> ctypes.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"pactae_amd: HIP library {LIB_PATH} is not built; run "
            "`python -c 'import __graft_entry__ as g; g.build()'` (or pactae_amd.build.build())")
    lib = ctypes.CDLL(LIB_PATH)
    for name, restype, argtypes in SIGNATURES:
        fn = getattr(lib, name)  # AttributeError here means the library lacks a declared symbol
        fn.restype = restype
        fn.argtypes = argtypes
    return lib
