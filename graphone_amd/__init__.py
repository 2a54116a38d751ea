"""graphone_amd: GraphOne's graph-computation hot path on MI355X (gfx950).

Host-side Python view of the C-ABI in include/graphone_hip.h. Device memory, kernels and engine loops
live in libgraphone_hip.so; this module only moves arrays and forwards calls. Every compute entry point
raises if the HIP library is missing: there is no CPU fallback.

Mapping to the reference API (paths relative to the reference root):
  DeviceArray  ~ DenseVec / Buffer           include/GraphOne/vec/dense.h:7-42, base/buffer.h:12-202
  HostGraph    ~ GraphLoader + EdgeCache      include/GraphOne/loader/graph_loader.h, edge_cache.h
  Graph        ~ graph::build<arch, views>    include/GraphOne/graph/builder.h:9-46
  Graph.bfs/sssp/pagerank/ppr ~ Run<Functor>(comp, frontier) with the apps' functors
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import (G1_BATCH_MINPLUS, G1_BATCH_PPR, G1_ENGINE_ADVANCE, G1_ENGINE_BLAS_CSRVEC, G1_ENGINE_BLAS_MERGE,
                   G1_ENGINE_DELTA, G1_ENGINE_DIROPT, G1_INF, G1_PR_L1_F64, G1_PR_MAX_ABS, BatchStats, G1Error,
                   HostCSR, RunStats, call, lib)

INF = G1_INF

ENGINES = {
    "advance": G1_ENGINE_ADVANCE,
    "merge": G1_ENGINE_BLAS_MERGE,
    "csrvec": G1_ENGINE_BLAS_CSRVEC,
    "diropt": G1_ENGINE_DIROPT,
    "delta": G1_ENGINE_DELTA,
}


def _engine(e) -> int:
    if isinstance(e, str):
        return ENGINES[e]
    return int(e)


def _check_out(out, count: int, dtype, what: str) -> None:
    """Rejects a caller-supplied output array that is too small or of the wrong element type before any
    device work: the library writes `count` elements through the raw pointer."""
    if out is None:
        return
    if not isinstance(out, DeviceArray):
        raise TypeError(f"{what}: out must be a DeviceArray, got {type(out).__name__}")
    if out.dtype != np.dtype(dtype):
        raise ValueError(f"{what}: out has dtype {out.dtype}, {np.dtype(dtype)} needed")
    if out.count < count:
        raise ValueError(f"{what}: out holds {out.count} elements, {count} needed")


# ----------------------------------------------------------------------------------------- device memory

class DeviceArray:
    """A typed device (HBM) array owned by Python, allocated through g1_malloc."""

    def __init__(self, count: int, dtype):
        self.count = int(count)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.count * self.dtype.itemsize
        p = C.c_void_p()
        call("g1_malloc", C.byref(p), C.c_size_t(self.nbytes))
        self._ptr = p.value or 0

    @property
    def ptr(self) -> int:
        return self._ptr

    @classmethod
    def from_numpy(cls, arr: np.ndarray) -> "DeviceArray":
        arr = np.ascontiguousarray(arr)
        d = cls(arr.size, arr.dtype)
        if d.nbytes:
            call("g1_memcpy_h2d", C.c_void_p(d.ptr), arr.ctypes.data_as(C.c_void_p), C.c_size_t(d.nbytes), None)
        return d

    def numpy(self) -> np.ndarray:
        out = np.empty(self.count, dtype=self.dtype)
        if self.nbytes:
            call("g1_memcpy_d2h", out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), C.c_size_t(self.nbytes),
                 None)
        return out

    def copy_from(self, arr: np.ndarray) -> None:
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.size == self.count
        if self.nbytes:
            call("g1_memcpy_h2d", C.c_void_p(self.ptr), arr.ctypes.data_as(C.c_void_p), C.c_size_t(self.nbytes),
                 None)

    def fill(self, value) -> None:
        if self.dtype == np.uint32:
            call("g1_fill_u32", C.c_void_p(self.ptr), C.c_uint32(int(value)), C.c_size_t(self.count), None)
        elif self.dtype == np.float32:
            call("g1_fill_f32", C.c_void_p(self.ptr), C.c_float(float(value)), C.c_size_t(self.count), None)
        else:
            raise TypeError(f"fill: unsupported dtype {self.dtype}")

    def free(self) -> None:
        if self._ptr:
            lib().g1_free(C.c_void_p(self._ptr))
            self._ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize(stream=None) -> None:
    call("g1_stream_synchronize", stream)


class Stream:
    """A non-blocking HIP stream of the library (g1_stream_create); pass it as `stream=` to the engines."""

    def __init__(self):
        h = C.c_void_p()
        call("g1_stream_create", C.byref(h))
        self._as_parameter_ = h

    def synchronize(self) -> None:
        call("g1_stream_synchronize", self._as_parameter_)

    def close(self) -> None:
        if self._as_parameter_ and self._as_parameter_.value:
            lib().g1_stream_destroy(self._as_parameter_)
            self._as_parameter_ = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_info() -> dict:
    count, cus, hbm = C.c_int32(), C.c_int32(), C.c_uint64()
    name = C.create_string_buffer(64)
    call("g1_device_info", C.byref(count), C.byref(cus), C.byref(hbm), name, C.c_size_t(64))
    return {"devices": count.value, "cus": cus.value, "hbm_bytes": hbm.value, "arch": name.value.decode()}


# ----------------------------------------------------------------------------------------- host graphs

class HostGraph:
    """A host CSR (row_offsets u32[n+1], col_indices u32[m], optional u32 weights)."""

    def __init__(self, csr: HostCSR, owned_by_lib: bool, keep=()):
        self._csr = csr
        self._owned = owned_by_lib
        self._keep = keep  # numpy arrays backing a borrowed CSR

    # constructors ------------------------------------------------------------------------
    @classmethod
    def rmat(cls, scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             seed: int = 1, undirected: bool = False) -> "HostGraph":
        csr = HostCSR()
        call("g1_gen_rmat", scale, edge_factor, a, b, c, C.c_uint64(seed), int(undirected), C.byref(csr))
        return cls(csr, True)

    @classmethod
    def grid(cls, rows: int, cols: int, p_keep: float = 0.6, seed: int = 1) -> "HostGraph":
        csr = HostCSR()
        call("g1_gen_grid", rows, cols, C.c_double(p_keep), C.c_uint64(seed), C.byref(csr))
        return cls(csr, True)

    @classmethod
    def from_edges(cls, n: int, src, dst, weights=None, directed: bool = True, keep_self: bool = False,
                   keep_dup: bool = False) -> "HostGraph":
        src = np.ascontiguousarray(src, dtype=np.uint32)
        dst = np.ascontiguousarray(dst, dtype=np.uint32)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.uint32)
        csr = HostCSR()
        call("g1_csr_from_edges", n, C.c_uint64(src.size), src.ctypes.data_as(C.c_void_p),
             dst.ctypes.data_as(C.c_void_p), None if w is None else w.ctypes.data_as(C.c_void_p),
             int(directed), int(keep_self), int(keep_dup), C.byref(csr))
        return cls(csr, True)

    @classmethod
    def from_csr(cls, row_offsets, col_indices, weights=None) -> "HostGraph":
        ro = np.ascontiguousarray(row_offsets, dtype=np.uint32)
        ci = np.ascontiguousarray(col_indices, dtype=np.uint32)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.uint32)
        csr = HostCSR()
        csr.n = ro.size - 1
        csr.m = ci.size
        csr.row_offsets = ro.ctypes.data_as(C.POINTER(C.c_uint32))
        csr.col_indices = ci.ctypes.data_as(C.POINTER(C.c_uint32))
        csr.weights = w.ctypes.data_as(C.POINTER(C.c_uint32)) if w is not None else None
        return cls(csr, False, keep=(ro, ci, w))

    # views -------------------------------------------------------------------------------
    @property
    def n(self) -> int:
        return self._csr.n

    @property
    def m(self) -> int:
        return self._csr.m

    @property
    def row_offsets(self) -> np.ndarray:
        return np.ctypeslib.as_array(self._csr.row_offsets, shape=(self.n + 1,))

    @property
    def col_indices(self) -> np.ndarray:
        if self.m == 0:
            return np.zeros(0, dtype=np.uint32)
        return np.ctypeslib.as_array(self._csr.col_indices, shape=(self.m,))

    @property
    def weights(self):
        if not self._csr.weights:
            return None
        if self.m == 0:
            return np.zeros(0, dtype=np.uint32)
        return np.ctypeslib.as_array(self._csr.weights, shape=(self.m,))

    def out_degrees(self) -> np.ndarray:
        return np.diff(self.row_offsets.astype(np.int64)).astype(np.uint32)

    def assign_weights(self, lo: int, hi: int, seed: int = 2, symmetric: bool = False) -> "HostGraph":
        if not self._owned:
            raise ValueError("assign_weights needs a library-owned graph (use a generator or from_edges)")
        call("g1_assign_weights_u32", C.byref(self._csr), lo, hi, C.c_uint64(seed), int(symmetric))
        return self

    def transpose(self) -> "HostGraph":
        out = HostCSR()
        call("g1_host_csr_transpose", C.byref(self._csr), C.byref(out))
        return HostGraph(out, True)

    def __del__(self):
        try:
            if self._owned:
                lib().g1_host_csr_free(C.byref(self._csr))
                self._owned = False
        except Exception:
            pass


# ----------------------------------------------------------------------------------------- device graph

@dataclass
class RunResult:
    values: np.ndarray
    stats: dict


class Graph:
    """A device-resident graph (g1_graph) with its transpose, and the whole-run engines."""

    def __init__(self, host: HostGraph, transpose: bool = True, symmetric: bool = False):
        flags = (_lib.G1_GRAPH_TRANSPOSE if transpose else 0) | (_lib.G1_GRAPH_SYMMETRIC if symmetric else 0)
        h = C.c_void_p()
        call("g1_graph_create", C.byref(h), C.byref(host._csr), flags)
        self._h = h
        self.n = host.n
        self.m = host.m
        self.weighted = host.weights is not None

    @property
    def handle(self):
        return self._h

    def device_bytes(self) -> int:
        b = C.c_uint64()
        call("g1_graph_info", self._h, None, None, C.byref(b))
        return b.value

    def work_bytes(self, work: int) -> int:
        """Device bytes the engines' work areas in `work` (G1_WORK_* bits) would add to this handle now."""
        b = C.c_uint64()
        call("g1_graph_work_bytes", self._h, int(work), C.byref(b))
        return b.value

    def prepare(self, work: int, stream=None) -> None:
        """Allocates and builds the work areas in `work` now (G1_WORK_* bits)."""
        call("g1_graph_prepare", self._h, int(work), stream)

    def csr_dev(self, transpose: bool = False):
        ro, ci, w = C.c_void_p(), C.c_void_p(), C.c_void_p()
        call("g1_graph_csr_dev", self._h, int(transpose), C.byref(ro), C.byref(ci), C.byref(w))
        return ro.value, ci.value, w.value

    def profile(self, enable: bool = True) -> None:
        """Per-kernel HIP-event timing of every engine launch (see read_profile)."""
        call("g1_profile_enable", self._h, int(enable))

    def read_profile(self) -> list[dict]:
        """Kernel entries accumulated since the last read: launches, device ms, algorithmic bytes."""
        cap = 32
        arr = (_lib.KernelProfile * cap)()
        cnt = C.c_uint32()
        call("g1_profile_read", self._h, arr, cap, C.byref(cnt))
        out = []
        for i in range(min(cnt.value, cap)):
            e = arr[i]
            out.append({"name": e.name.decode(), "launches": e.launches, "total_ms": e.total_ms,
                        "algorithmic_bytes": e.algorithmic_bytes})
        return out

    def _run_u32(self, fn: str, src: int, engine, out: DeviceArray | None, stream):
        _check_out(out, self.n, np.uint32, fn)
        d = out if out is not None else DeviceArray(self.n, np.uint32)
        st = RunStats()
        call(fn, self._h, C.c_uint32(src), C.c_void_p(d.ptr), _engine(engine), C.byref(st), stream)
        return d, st.as_dict()

    def bfs(self, src: int, engine="advance", out: DeviceArray | None = None, stream=None,
            to_host: bool = True):
        d, st = self._run_u32("g1_bfs", src, engine, out, stream)
        return RunResult(d.numpy() if to_host else d, st)

    def sssp(self, src: int, engine="advance", out: DeviceArray | None = None, stream=None,
             to_host: bool = True):
        d, st = self._run_u32("g1_sssp", src, engine, out, stream)
        return RunResult(d.numpy() if to_host else d, st)

    def pagerank(self, alpha: float = 0.85, max_iter: int = 50, tol: float = 1e-7, mode: str = "l1",
                 engine="merge", out: DeviceArray | None = None, stream=None, to_host: bool = True):
        _check_out(out, self.n, np.float32, "g1_pagerank")
        r = out if out is not None else DeviceArray(self.n, np.float32)
        st = RunStats()
        cm = G1_PR_L1_F64 if mode == "l1" else G1_PR_MAX_ABS
        call("g1_pagerank", self._h, C.c_float(alpha), C.c_uint32(max_iter), C.c_double(tol), cm, _engine(engine),
             C.c_void_p(r.ptr), C.byref(st), stream)
        return RunResult(r.numpy() if to_host else r, st.as_dict())

    def ppr(self, seed: int, restart: float = 0.15, max_iter: int = 50, tol: float = 1e-7,
            out: DeviceArray | None = None, stream=None, to_host: bool = True):
        _check_out(out, self.n, np.float32, "g1_ppr")
        r = out if out is not None else DeviceArray(self.n, np.float32)
        st = RunStats()
        call("g1_ppr", self._h, C.c_uint32(seed), C.c_float(restart), C.c_uint32(max_iter), C.c_double(tol),
             C.c_void_p(r.ptr), C.byref(st), stream)
        return RunResult(r.numpy() if to_host else r, st.as_dict())

    # ---- batched multi-source engines (k sources as the columns of one SpMM)

    @staticmethod
    def _sources(sources) -> np.ndarray:
        return np.ascontiguousarray(np.asarray(sources, dtype=np.uint32).reshape(-1))

    def autotune(self, kind: str, sources, weighted: bool = True, stream=None) -> int:
        """Times every column width on the graph and keeps the best for kind 'minplus' or 'ppr'."""
        src = self._sources(sources)
        best = C.c_uint32()
        k = G1_BATCH_PPR if kind == "ppr" else G1_BATCH_MINPLUS
        call("g1_batch_autotune", self._h, k, src.ctypes.data_as(C.POINTER(C.c_uint32)), C.c_uint32(src.size),
             int(weighted), C.byref(best), stream)
        return best.value

    def set_merge_tile(self, tile_items: int) -> None:
        """Merge tile (rows + nonzeros: 512, 1024, 2048 or 4096) of the k = 64 min-plus batches."""
        call("g1_batch_set_merge_tile", self._h, C.c_uint32(tile_items))

    def sssp_batch(self, sources, k: int = 0, weighted: bool = True, out: DeviceArray | None = None, stream=None,
                   to_host: bool = True):
        """Distances (weighted) or BFS depths of every source: values[c] is the run from sources[c]."""
        src = self._sources(sources)
        _check_out(out, src.size * self.n, np.uint32, "g1_sssp_batch")
        d = out if out is not None else DeviceArray(max(1, src.size * self.n), np.uint32)
        st = BatchStats()
        call("g1_sssp_batch", self._h, src.ctypes.data_as(C.POINTER(C.c_uint32)), C.c_uint32(src.size),
             C.c_uint32(k), int(weighted), C.c_void_p(d.ptr), C.byref(st), stream)
        vals = d.numpy()[: src.size * self.n].reshape(src.size, self.n) if to_host else d
        return RunResult(vals, st.as_dict())

    def road_batch(self, sources, kinds=None, out: DeviceArray | None = None, stream=None, to_host: bool = True):
        """BFS ('bfs') / SSSP ('sssp') queries on a low-degree graph, one workgroup per query: values[c]
        equals bfs/sssp(sources[c]) bit for bit."""
        src = self._sources(sources)
        kk = None
        if kinds is not None:
            kinds = list(kinds)
            if len(kinds) != src.size:
                raise ValueError(f"g1_road_batch: {len(kinds)} kinds for {src.size} sources")
            kk = np.ascontiguousarray([1 if k in ("sssp", 1) else 0 for k in kinds], dtype=np.uint32)
        _check_out(out, src.size * self.n, np.uint32, "g1_road_batch")
        d = out if out is not None else DeviceArray(max(1, src.size * self.n), np.uint32)
        its = np.zeros(max(1, src.size), np.uint32)
        call("g1_road_batch", self._h, None if kk is None else kk.ctypes.data_as(C.POINTER(C.c_uint32)),
             src.ctypes.data_as(C.POINTER(C.c_uint32)), C.c_uint32(src.size), C.c_void_p(d.ptr),
             its.ctypes.data_as(C.POINTER(C.c_uint32)), stream)
        vals = d.numpy()[: src.size * self.n].reshape(src.size, self.n) if to_host else d
        return RunResult(vals, {"iterations": its[: src.size].tolist()})

    def ppr_batch(self, seeds, restart: float = 0.15, max_iter: int = 50, tol: float = 1e-7, k: int = 0,
                  out: DeviceArray | None = None, stream=None, to_host: bool = True):
        """Personalized PageRank of every seed: values[c] equals ppr(seeds[c]) bit for bit."""
        src = self._sources(seeds)
        _check_out(out, src.size * self.n, np.float32, "g1_ppr_batch")
        r = out if out is not None else DeviceArray(max(1, src.size * self.n), np.float32)
        st = BatchStats()
        call("g1_ppr_batch", self._h, src.ctypes.data_as(C.POINTER(C.c_uint32)), C.c_uint32(src.size),
             C.c_float(restart), C.c_uint32(max_iter), C.c_double(tol), C.c_uint32(k), C.c_void_p(r.ptr),
             C.byref(st), stream)
        vals = r.numpy()[: src.size * self.n].reshape(src.size, self.n) if to_host else r
        return RunResult(vals, st.as_dict())

    def view(self) -> "Graph":
        """A second handle on this resident graph (g1_graph_create_view) with its own work areas, for runs
        on another stream at the same time. Close it before this graph."""
        v = Graph.__new__(Graph)
        h = C.c_void_p()
        call("g1_graph_create_view", self._h, C.byref(h))
        v._h = h
        v.n, v.m, v.weighted = self.n, self.m, self.weighted
        v._base = self  # keeps the base alive while the view is
        return v

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().g1_graph_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------------------------------------- query server

QUERY_KINDS = {"bfs": _lib.G1_QUERY_BFS, "sssp": _lib.G1_QUERY_SSSP, "ppr": _lib.G1_QUERY_PPR}
QUERY_STATES = {_lib.G1_Q_QUEUED: "queued", _lib.G1_Q_RUNNING: "running", _lib.G1_Q_DONE: "done",
                _lib.G1_Q_FAILED: "failed", _lib.G1_Q_CANCELLED: "cancelled", _lib.G1_Q_FETCHED: "fetched"}


class Server:
    """The query server of the C-ABI (g1_server_*): resident graphs, per-graph request loops,
    HBM-budget admission and continuous batching of SSSP / personalized-PageRank queries."""

    def __init__(self, hbm_budget_bytes: int = 0, batch_k: int = 0, batch_min_sssp: int = 0,
                 batch_min_ppr: int = 0, ppr_restart: float = 0.0, single_workers: int = 0):
        cfg = _lib.ServerConfig()
        cfg.single_workers = int(single_workers)
        cfg.hbm_budget_bytes = int(hbm_budget_bytes)
        cfg.batch_k = int(batch_k)
        cfg.batch_min_sssp = int(batch_min_sssp)
        cfg.batch_min_ppr = int(batch_min_ppr)
        cfg.ppr_restart = float(ppr_restart)
        h = C.c_void_p()
        call("g1_server_create", C.byref(cfg), C.byref(h))
        self._h = h
        self._graphs: list = []

    def add_graph(self, graph: "Graph") -> int:
        gid = C.c_uint32()
        call("g1_server_add_graph", self._h, graph.handle, C.byref(gid))
        self._graphs.append(graph)  # the graph must outlive the server
        return gid.value

    def remove_graph(self, graph_id: int) -> None:
        """Unregisters a graph once its queued and running queries are final (g1_server_remove_graph)."""
        call("g1_server_remove_graph", self._h, C.c_uint32(graph_id))

    def set_budget(self, hbm_budget_bytes: int) -> None:
        call("g1_server_set_budget", self._h, C.c_uint64(hbm_budget_bytes))

    def submit(self, graph_id: int, kind: str, source: int, max_iter: int = 50, tol: float = 1e-7) -> int:
        q = _lib.Query()
        q.graph_id = int(graph_id)
        q.kind = QUERY_KINDS[kind] if isinstance(kind, str) else int(kind)
        q.source = int(source)
        q.max_iter = int(max_iter)
        q.tol = float(tol)
        qid = C.c_uint64()
        call("g1_server_submit", self._h, C.byref(q), C.byref(qid))
        return qid.value

    @staticmethod
    def _status(st) -> dict:
        d = st.as_dict()
        d["state_name"] = QUERY_STATES.get(d["state"], "?")
        d["latency_ms"] = d["t_done_ms"] - d["t_submit_ms"] if d["t_done_ms"] else None
        return d

    def poll(self, qid: int) -> dict:
        st = _lib.QueryStatus()
        call("g1_server_poll", self._h, C.c_uint64(qid), C.byref(st))
        return self._status(st)

    def wait(self, qid: int, timeout_ms: float = 60000.0) -> dict:
        st = _lib.QueryStatus()
        call("g1_server_wait", self._h, C.c_uint64(qid), C.c_double(timeout_ms), C.byref(st))
        return self._status(st)

    def fetch(self, qid: int, kind: str | None = None, out: DeviceArray | None = None):
        """Copies the result of a done query (to host numpy, or into `out` on the device) and frees it."""
        st = self.poll(qid)
        n = st["result_bytes"] // 4
        if out is not None:
            call("g1_server_fetch", self._h, C.c_uint64(qid), C.c_void_p(out.ptr), C.c_uint64(out.nbytes), 1)
            return out
        dtype = np.float32 if kind == "ppr" else np.uint32
        arr = np.empty(n, dtype=dtype)
        call("g1_server_fetch", self._h, C.c_uint64(qid), C.c_void_p(arr.ctypes.data), C.c_uint64(arr.nbytes), 0)
        return arr

    def cancel(self, qid: int) -> None:
        call("g1_server_cancel", self._h, C.c_uint64(qid))

    def stats(self) -> dict:
        st = _lib.ServerStats()
        call("g1_server_get_stats", self._h, C.byref(st))
        return st.as_dict()

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().g1_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------------------------------------- SpMV

def spmv(kind: str, semiring: str, row_offsets: DeviceArray, col_indices: DeviceArray, values, x: DeviceArray,
         y: DeviceArray, n_rows: int, n_cols: int, nnz: int, workspace: DeviceArray | None = None,
         lanes: int = 0, stream=None) -> None:
    """y = A (x) x on a CSR matrix already in device memory.

    kind: 'merge' (merge-path; workspace from merge_workspace() and merge_partition()) or 'csrvec'.
    semiring: 'minplus_u32' or 'plustimes_f32'. values may be None (unit weights).
    """
    vp = None if values is None else C.c_void_p(values.ptr)
    args = (C.c_uint32(n_rows), C.c_uint32(n_cols), C.c_uint32(nnz), C.c_void_p(row_offsets.ptr),
            C.c_void_p(col_indices.ptr), vp, C.c_void_p(x.ptr), C.c_void_p(y.ptr))
    if kind == "merge":
        assert workspace is not None
        call(f"g1_spmv_merge_{semiring}", *args, C.c_void_p(workspace.ptr), C.c_size_t(workspace.nbytes), stream)
    elif kind == "csrvec":
        call(f"g1_spmv_csrvec_{semiring}", *args, C.c_uint32(lanes), stream)
    else:
        raise ValueError(kind)


def merge_workspace(n_rows: int, nnz: int, row_offsets: DeviceArray, stream=None) -> DeviceArray:
    nbytes = lib().g1_spmv_merge_workspace_bytes(n_rows, nnz)
    ws = DeviceArray((nbytes + 3) // 4, np.uint32)
    call("g1_spmv_merge_partition", C.c_uint32(n_rows), C.c_uint32(nnz), C.c_void_p(row_offsets.ptr),
         C.c_void_p(ws.ptr), C.c_size_t(ws.nbytes), stream)
    return ws


__all__ = [
    "INF", "ENGINES", "DeviceArray", "HostGraph", "Graph", "RunResult", "G1Error", "spmv", "merge_workspace",
    "synchronize", "device_info",
]
