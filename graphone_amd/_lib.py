"""ctypes binding of libgraphone_hip.so (declared in include/graphone_hip.h).

Loading fails loudly: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgraphone_hip.so")

G1_OK = 0
G1_ERR_INVALID_ARG = 1
G1_ERR_HIP = 2
G1_ERR_OUT_OF_MEMORY = 3
G1_ERR_BUDGET = 4
G1_ERR_NOT_FOUND = 5
G1_ERR_BUSY = 6
G1_ERR_CANCELLED = 7
G1_ERR_INTERNAL = 8
G1_ERR_NOT_CONVERGED = 9
G1_INF = 0xFFFFFFFF

G1_GRAPH_TRANSPOSE = 1
G1_GRAPH_SYMMETRIC = 2

G1_ENGINE_ADVANCE = 0
G1_ENGINE_BLAS_MERGE = 1
G1_ENGINE_BLAS_CSRVEC = 2
G1_ENGINE_DIROPT = 3
G1_ENGINE_DELTA = 4

G1_PR_L1_F64 = 0
G1_PR_MAX_ABS = 1

G1_WORK_PUSH = 1
G1_WORK_PULL = 2
G1_WORK_RANK = 4
G1_WORK_RELABEL = 8
G1_WORK_DELTA = 16
G1_WORK_ROAD = 32
G1_WORK_HUB = 64
G1_WORK_ALL = 127

G1_BATCH_MINPLUS = 0
G1_BATCH_PPR = 1

G1_QUERY_BFS = 0
G1_QUERY_SSSP = 1
G1_QUERY_PPR = 2

G1_Q_QUEUED = 0
G1_Q_RUNNING = 1
G1_Q_DONE = 2
G1_Q_FAILED = 3
G1_Q_CANCELLED = 4
G1_Q_FETCHED = 5


class G1Error(RuntimeError):
    def __init__(self, status: int, where: str, message: str):
        super().__init__(f"{where}: status {status} ({message})")
        self.status = status


class HostCSR(C.Structure):
    _fields_ = [
        ("n", C.c_uint32),
        ("m", C.c_uint32),
        ("row_offsets", C.POINTER(C.c_uint32)),
        ("col_indices", C.POINTER(C.c_uint32)),
        ("weights", C.POINTER(C.c_uint32)),
    ]


class RunStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_uint32),
        ("push_iterations", C.c_uint32),
        ("pull_iterations", C.c_uint32),
        ("dense_iterations", C.c_uint32),
        ("frontier_sum", C.c_uint64),
        ("edges_sum", C.c_uint64),
        ("reached", C.c_uint64),
        ("reached_edges", C.c_uint64),
        ("elapsed_ms", C.c_double),
        ("last_delta", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class BatchStats(C.Structure):
    _fields_ = [
        ("k", C.c_uint32),
        ("tiles", C.c_uint32),
        ("iterations_sum", C.c_uint64),
        ("elapsed_ms", C.c_double),
        ("merge_tile", C.c_uint32),
        ("reserved", C.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class ServerConfig(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_uint64),
        ("batch_k", C.c_uint32),
        ("batch_min_sssp", C.c_uint32),
        ("batch_min_ppr", C.c_uint32),
        ("ppr_restart", C.c_float),
        ("single_workers", C.c_uint32),
        ("reserved", C.c_uint32 * 2),
    ]


class Query(C.Structure):
    _fields_ = [
        ("graph_id", C.c_uint32),
        ("kind", C.c_int32),
        ("source", C.c_uint32),
        ("max_iter", C.c_uint32),
        ("tol", C.c_double),
    ]


class QueryStatus(C.Structure):
    _fields_ = [
        ("state", C.c_int32),
        ("status", C.c_int32),
        ("path", C.c_int32),
        ("iterations", C.c_uint32),
        ("t_submit_ms", C.c_double),
        ("t_admit_ms", C.c_double),
        ("t_done_ms", C.c_double),
        ("result_bytes", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class ServerStats(C.Structure):
    _fields_ = [(name, C.c_uint64) for name in (
        "submitted", "completed", "failed", "cancelled", "queued", "running", "admission_waits", "batched",
        "single", "session_steps", "hbm_budget", "hbm_used", "hbm_peak")]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class KernelProfile(C.Structure):
    _fields_ = [
        ("name", C.c_char * 40),
        ("launches", C.c_uint64),
        ("total_ms", C.c_double),
        ("algorithmic_bytes", C.c_double),
    ]


_u32 = C.c_uint32
_u64 = C.c_uint64
_i32 = C.c_int32
_vp = C.c_void_p
_st = C.c_int32
_pu32 = C.POINTER(C.c_uint32)

_SIGNATURES = {
    "g1_status_string": (C.c_char_p, [_st]),
    "g1_last_error_message": (C.c_char_p, []),
    "g1_abi_version": (_i32, []),
    "g1_device_info": (_st, [C.POINTER(_i32), C.POINTER(_i32), C.POINTER(_u64), C.c_char_p, C.c_size_t]),
    "g1_malloc": (_st, [C.POINTER(_vp), C.c_size_t]),
    "g1_free": (_st, [_vp]),
    "g1_memcpy_h2d": (_st, [_vp, _vp, C.c_size_t, _vp]),
    "g1_memcpy_d2h": (_st, [_vp, _vp, C.c_size_t, _vp]),
    "g1_memcpy_d2d": (_st, [_vp, _vp, C.c_size_t, _vp]),
    "g1_fill_u32": (_st, [_vp, _u32, C.c_size_t, _vp]),
    "g1_fill_f32": (_st, [_vp, C.c_float, C.c_size_t, _vp]),
    "g1_stream_create": (_st, [C.POINTER(_vp)]),
    "g1_stream_destroy": (_st, [_vp]),
    "g1_stream_synchronize": (_st, [_vp]),
    "g1_event_create": (_st, [C.POINTER(_vp)]),
    "g1_event_destroy": (_st, [_vp]),
    "g1_event_record": (_st, [_vp, _vp]),
    "g1_event_synchronize": (_st, [_vp]),
    "g1_event_elapsed_ms": (_st, [C.POINTER(C.c_float), _vp, _vp]),
    "g1_spmv_merge_workspace_bytes": (C.c_size_t, [_u32, _u32]),
    "g1_spmv_merge_partition": (_st, [_u32, _u32, _vp, _vp, C.c_size_t, _vp]),
    "g1_spmv_merge_minplus_u32": (_st, [_u32, _u32, _u32, _vp, _vp, _vp, _vp, _vp, _vp, C.c_size_t, _vp]),
    "g1_spmv_merge_plustimes_f32": (_st, [_u32, _u32, _u32, _vp, _vp, _vp, _vp, _vp, _vp, C.c_size_t, _vp]),
    "g1_spmv_merge_noalloc_max_workspace_bytes": (C.c_size_t, []),
    "g1_spmv_merge_minplus_u32_noalloc": (_st, [_u32, _u32, _u32, _vp, _vp, _vp, _vp, _vp, _vp]),
    "g1_spmv_csrvec_lanes": (_u32, [_u32, _u32]),
    "g1_csr_transpose_workspace_bytes": (C.c_size_t, [_u32, _u32, C.c_int32]),
    "g1_csr_transpose": (_st, [_u32, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_size_t, _vp]),
    "g1_step_workspace_bytes": (C.c_size_t, [_u32, _u32]),
    "g1_frontier_density": (_st, [_u32, _vp, _vp, _vp, _u32, _vp, _vp, C.c_size_t, _vp]),
    "g1_advance_sssp_u32": (_st, [_u32, _vp, _vp, _vp, _vp, _vp, _u32, _vp, _vp, _u32, _vp, _vp, _vp, C.c_size_t,
                                  _vp]),
    "g1_advance_bfs": (_st, [_u32, _vp, _vp, _vp, _vp, _u32, _vp, _u32, _vp, _vp, _vp, C.c_size_t, _vp]),
    "g1_filter_compact": (_st, [_vp, _vp, _u32, _vp, _vp, _vp, C.c_size_t, _vp]),
    "g1_frontier_queue_to_bitmap": (_st, [_u32, _vp, _vp, _u32, _vp, _vp]),
    "g1_frontier_bitmap_to_queue": (_st, [_u32, _vp, _vp, _vp, _vp, C.c_size_t, _vp]),
    "g1_apply_minplus_u32": (_st, [_u32, _vp, _vp, _vp, _vp, _vp, C.c_size_t, _vp]),
    "g1_pr_apply_f32": (_st, [_u32, _vp, C.c_float, _vp, _vp, _vp, _vp, _vp, _vp, C.c_size_t, _vp]),
    "g1_l1_delta_f64": (_st, [_u32, _vp, _vp, _vp, _vp, C.c_size_t, _vp]),
    "g1_spmm_workspace_bytes": (C.c_size_t, [_u32, _u32]),
    "g1_spmm_minplus_u32": (_st, [_u32, _u32, _u32, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_size_t, _vp]),
    "g1_spmm_plustimes_f32": (_st, [_u32, _u32, _u32, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_size_t, _vp]),
    "g1_spmv_csrvec_minplus_u32": (_st, [_u32, _u32, _u32, _vp, _vp, _vp, _vp, _vp, _u32, _vp]),
    "g1_spmv_csrvec_plustimes_f32": (_st, [_u32, _u32, _u32, _vp, _vp, _vp, _vp, _vp, _u32, _vp]),
    "g1_gen_rmat": (_st, [_u32, _u32, C.c_double, C.c_double, C.c_double, _u64, _i32, C.POINTER(HostCSR)]),
    "g1_gen_grid": (_st, [_u32, _u32, C.c_double, _u64, C.POINTER(HostCSR)]),
    "g1_csr_from_edges": (_st, [_u32, _u64, _vp, _vp, _vp, _i32, _i32, _i32, C.POINTER(HostCSR)]),
    "g1_assign_weights_u32": (_st, [C.POINTER(HostCSR), _u32, _u32, _u64, _i32]),
    "g1_host_csr_transpose": (_st, [C.POINTER(HostCSR), C.POINTER(HostCSR)]),
    "g1_host_csr_free": (None, [C.POINTER(HostCSR)]),
    "g1_graph_create": (_st, [C.POINTER(_vp), C.POINTER(HostCSR), _u32]),
    "g1_graph_destroy": (_st, [_vp]),
    "g1_graph_create_view": (_st, [_vp, C.POINTER(_vp)]),
    "g1_graph_info": (_st, [_vp, C.POINTER(_u32), C.POINTER(_u32), C.POINTER(_u64)]),
    "g1_graph_work_bytes": (_st, [_vp, _i32, C.POINTER(_u64)]),
    "g1_graph_prepare": (_st, [_vp, _i32, _vp]),
    "g1_graph_csr_dev": (_st, [_vp, _i32, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    "g1_profile_enable": (_st, [_vp, _i32]),
    "g1_profile_read": (_st, [_vp, C.POINTER(KernelProfile), _u32, C.POINTER(_u32)]),
    "g1_bfs": (_st, [_vp, _u32, _vp, _i32, C.POINTER(RunStats), _vp]),
    "g1_sssp": (_st, [_vp, _u32, _vp, _i32, C.POINTER(RunStats), _vp]),
    "g1_pagerank": (_st, [_vp, C.c_float, _u32, C.c_double, _i32, _i32, _vp, C.POINTER(RunStats), _vp]),
    "g1_ppr": (_st, [_vp, _u32, C.c_float, _u32, C.c_double, _vp, C.POINTER(RunStats), _vp]),
    "g1_batch_autotune": (_st, [_vp, _i32, _pu32, _u32, _i32, _pu32, _vp]),
    "g1_batch_set_merge_tile": (_st, [_vp, _u32]),
    "g1_road_batch": (_st, [_vp, _pu32, _pu32, _u32, _vp, _pu32, _vp]),
    "g1_sssp_batch": (_st, [_vp, _pu32, _u32, _u32, _i32, _vp, C.POINTER(BatchStats), _vp]),
    "g1_ppr_batch": (_st, [_vp, _pu32, _u32, C.c_float, _u32, C.c_double, _u32, _vp, C.POINTER(BatchStats), _vp]),
    "g1_server_create": (_st, [C.POINTER(ServerConfig), C.POINTER(_vp)]),
    "g1_server_destroy": (_st, [_vp]),
    "g1_server_add_graph": (_st, [_vp, _vp, C.POINTER(_u32)]),
    "g1_server_remove_graph": (_st, [_vp, _u32]),
    "g1_server_set_budget": (_st, [_vp, _u64]),
    "g1_server_submit": (_st, [_vp, C.POINTER(Query), C.POINTER(_u64)]),
    "g1_server_poll": (_st, [_vp, _u64, C.POINTER(QueryStatus)]),
    "g1_server_wait": (_st, [_vp, _u64, C.c_double, C.POINTER(QueryStatus)]),
    "g1_server_fetch": (_st, [_vp, _u64, _vp, _u64, _i32]),
    "g1_server_cancel": (_st, [_vp, _u64]),
    "g1_server_get_stats": (_st, [_vp, C.POINTER(ServerStats)]),
}


def declared_symbols() -> list[str]:
    """Names of every function the C header declares (parsed from include/graphone_hip.h)."""
    import re

    hdr = os.path.join(os.path.dirname(_HERE), "include", "graphone_hip.h")
    text = open(hdr).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(g1_[a-z0-9_]+)\s*\(", text)))


_lib = None


def lib() -> C.CDLL:
    """Loads the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        handle = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status: int, where: str) -> None:
    if status != G1_OK:
        L = lib()
        msg = L.g1_last_error_message().decode(errors="replace")
        raise G1Error(status, where, msg or L.g1_status_string(status).decode())


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)
