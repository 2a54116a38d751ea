"""ctypes view of the CPU oracle (oracle/liborc.so) and of the reference's own naive SpMV
(oracle/_ref/libref_naive.so, built from the reference tree when it is present).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke test and by bench.py's cpu_baseline leg,
always as the checker or the CPU baseline, never as a product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORC_PATH = os.path.join(_HERE, "liborc.so")
REF_PATH = os.path.join(_HERE, "_ref", "libref_naive.so")
INF = 0xFFFFFFFF

_orc = None
_ref = None


def build() -> None:
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def orc() -> C.CDLL:
    global _orc
    if _orc is None:
        if not os.path.exists(ORC_PATH):
            build()
        _orc = C.CDLL(ORC_PATH)
    return _orc


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref() -> C.CDLL:
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_PATH)
    return _ref


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _u32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.uint32)


def spmv_minplus_u32(ap, aj, ax, x, use_ref: bool = False) -> np.ndarray:
    ap, aj, ax, x = _u32(ap), _u32(aj), _u32(ax), _u32(x)
    n = ap.size - 1
    y = np.empty(n, dtype=np.uint32)
    fn = ref().ref_spmv_minplus_u32 if use_ref else orc().orc_spmv_minplus_u32
    if use_ref and ax is None:
        ax = np.ones(aj.size, dtype=np.uint32)
    fn(C.c_uint32(n), _p(ap), _p(aj), _p(ax), _p(x), _p(y))
    return y


def spmv_plustimes_f32(ap, aj, ax, x, use_ref: bool = False) -> np.ndarray:
    ap, aj = _u32(ap), _u32(aj)
    ax = None if ax is None else np.ascontiguousarray(ax, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    n = ap.size - 1
    y = np.empty(n, dtype=np.float32)
    fn = ref().ref_spmv_plustimes_f32 if use_ref else orc().orc_spmv_plustimes_f32
    if use_ref and ax is None:
        ax = np.ones(aj.size, dtype=np.float32)
    fn(C.c_uint32(n), _p(ap), _p(aj), _p(ax), _p(x), _p(y))
    return y


def spmv_plustimes_f64(ap, aj, ax, x) -> np.ndarray:
    ap, aj = _u32(ap), _u32(aj)
    ax = None if ax is None else np.ascontiguousarray(ax, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = ap.size - 1
    y = np.empty(n, dtype=np.float64)
    orc().orc_spmv_plustimes_f64(C.c_uint32(n), _p(ap), _p(aj), _p(ax), _p(x), _p(y))
    return y


def sssp_advance(rp, ci, w, src: int):
    rp, ci, w = _u32(rp), _u32(ci), _u32(w)
    n = rp.size - 1
    d = np.empty(n, dtype=np.uint32)
    st = np.zeros(3, dtype=np.uint64)
    rc = orc().orc_sssp_advance(C.c_uint32(n), _p(rp), _p(ci), _p(w), C.c_uint32(src), _p(d), _p(st))
    assert rc == 0
    return d, {"iterations": int(st[0]), "frontier_sum": int(st[1]), "edges_sum": int(st[2])}


def sssp_advance_omp(rp, ci, w, src: int):
    """orc_sssp_advance_omp: the advance loop on every host thread (see set_threads)."""
    rp, ci, w = _u32(rp), _u32(ci), _u32(w)
    n = rp.size - 1
    d = np.empty(n, dtype=np.uint32)
    st = np.zeros(3, dtype=np.uint64)
    rc = orc().orc_sssp_advance_omp(C.c_uint32(n), _p(rp), _p(ci), _p(w), C.c_uint32(src), _p(d), _p(st))
    assert rc == 0
    return d, {"iterations": int(st[0]), "frontier_sum": int(st[1]), "edges_sum": int(st[2])}


def set_threads(threads: int = 0) -> int:
    """Host threads of the oracle's OpenMP loops (0: every thread of the machine); returns the count."""
    orc().orc_set_threads(C.c_int(int(threads)))
    return int(orc().orc_get_threads())


def minplus_blas(tp, tj, tw, src: int):
    tp, tj, tw = _u32(tp), _u32(tj), _u32(tw)
    n = tp.size - 1
    d = np.empty(n, dtype=np.uint32)
    st = np.zeros(2, dtype=np.uint64)
    rc = orc().orc_minplus_blas(C.c_uint32(n), _p(tp), _p(tj), _p(tw), C.c_uint32(src), _p(d), _p(st))
    assert rc == 0
    return d, {"iterations": int(st[0]), "frontier_sum": int(st[1])}


def bfs_queue(rp, ci, src: int) -> np.ndarray:
    rp, ci = _u32(rp), _u32(ci)
    n = rp.size - 1
    d = np.empty(n, dtype=np.uint32)
    assert orc().orc_bfs_queue(C.c_uint32(n), _p(rp), _p(ci), C.c_uint32(src), _p(d)) == 0
    return d


def dijkstra(rp, ci, w, src: int) -> np.ndarray:
    rp, ci, w = _u32(rp), _u32(ci), _u32(w)
    n = rp.size - 1
    d = np.empty(n, dtype=np.uint32)
    assert orc().orc_dijkstra(C.c_uint32(n), _p(rp), _p(ci), _p(w), C.c_uint32(src), _p(d)) == 0
    return d


def floyd_warshall(rp, ci, w) -> np.ndarray:
    rp, ci, w = _u32(rp), _u32(ci), _u32(w)
    n = rp.size - 1
    d = np.empty(n * n, dtype=np.uint64)
    orc().orc_floyd_warshall(C.c_uint32(n), _p(rp), _p(ci), _p(w), _p(d))
    return d.reshape(n, n)


def pagerank_f64(rp, w, tp, tj, tw, alpha: float = 0.85, seed: int = INF, restart: float = 0.15,
                 max_iter: int = 50, tol: float = 1e-7, mode: int = 0):
    rp, w, tp, tj, tw = _u32(rp), _u32(w), _u32(tp), _u32(tj), _u32(tw)
    n = rp.size - 1
    out = np.empty(n, dtype=np.float64)
    last = C.c_double()
    f = orc().orc_pagerank_f64
    f.restype = C.c_int
    it = f(C.c_uint32(n), _p(rp), _p(w), _p(tp), _p(tj), _p(tw), C.c_double(alpha), C.c_uint32(seed),
           C.c_double(restart), C.c_uint32(max_iter), C.c_double(tol), C.c_int(mode), _p(out), C.byref(last))
    assert it >= 0
    return out, {"iterations": it, "last_delta": last.value}
