"""Pins the CPU oracle before it is trusted (no GPU needed).

The reference ships no golden vectors (its tests print without asserting and read absent datasets), so
each oracle function is checked against an independent answer chosen so that a plausible slip (a
dropped term, a wrong index, a transposed operand, a wrong fold identity) fails:
  - the reference's own SpMV_cpu_navie, compiled from the reference tree into oracle/_ref, where built;
  - scipy sparse products and numpy dense solves (library routines);
  - closed forms (grid Manhattan distances, path/star/complete-graph BFS depths, uniform PageRank on
    regular graphs, the two-vertex personalized PageRank);
  - brute force (Floyd-Warshall, dense min-plus products) on small graphs;
  - committed golden vectors in tests/golden/.
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle

INF = oracle.INF
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def random_csr(rng, n, avg, weighted=True, lo=1, hi=255, allow_self=False):
    rows = []
    for u in range(n):
        k = min(n, rng.poisson(avg))
        c = np.sort(rng.choice(n, k, replace=False)) if k else np.zeros(0, np.int64)
        if not allow_self:
            c = c[c != u]
        rows.append(c)
    rp = np.zeros(n + 1, np.uint32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.concatenate(rows).astype(np.uint32) if rp[-1] else np.zeros(0, np.uint32)
    w = rng.integers(lo, hi + 1, ci.size, dtype=np.uint32) if weighted else None
    return rp, ci, w


def transpose(rp, ci, w):
    n = rp.size - 1
    m = sp.csr_matrix((np.arange(1, ci.size + 1), ci, rp), shape=(n, n))
    t = m.T.tocsr()
    t.sort_indices()
    order = t.data - 1
    tw = None if w is None else w[order]
    return t.indptr.astype(np.uint32), t.indices.astype(np.uint32), tw


def grid_csr(rows, cols):
    src, dst = [], []
    for r in range(rows):
        for c in range(cols):
            v = r * cols + c
            for dr, dc in ((-1, 0), (0, -1), (0, 1), (1, 0)):
                rr, cc = r + dr, c + dc
                if 0 <= rr < rows and 0 <= cc < cols:
                    src.append(v)
                    dst.append(rr * cols + cc)
    n = rows * cols
    m = sp.csr_matrix((np.ones(len(src)), (src, dst)), shape=(n, n))
    m.sort_indices()
    return m.indptr.astype(np.uint32), m.indices.astype(np.uint32)


# ------------------------------------------------------------------------- SpMV fold

@pytest.mark.skipif(not oracle.ref_available(), reason="reference naive SpMV not built (no reference tree)")
def test_spmv_matches_reference_naive_build():
    rng = np.random.default_rng(0)
    for n, avg in ((1, 0), (50, 3), (400, 12)):
        rp, ci, w = random_csr(rng, n, avg, allow_self=True)
        x = rng.integers(0, 1 << 20, n, dtype=np.uint32)
        x[rng.random(n) < 0.4] = INF
        assert np.array_equal(oracle.spmv_minplus_u32(rp, ci, w, x),
                              oracle.spmv_minplus_u32(rp, ci, w, x, use_ref=True))
        xf = rng.random(n).astype(np.float32)
        wf = rng.random(ci.size).astype(np.float32)
        a = oracle.spmv_plustimes_f32(rp, ci, wf, xf)
        b = oracle.spmv_plustimes_f32(rp, ci, wf, xf, use_ref=True)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_spmv_minplus_vs_dense_bruteforce():
    rng = np.random.default_rng(1)
    rp, ci, w = random_csr(rng, 60, 5)
    x = rng.integers(0, 1000, 60, dtype=np.uint32)
    x[::3] = INF
    y = oracle.spmv_minplus_u32(rp, ci, w, x)
    for r in range(60):
        best = INF
        for k in range(rp[r], rp[r + 1]):
            if x[ci[k]] != INF:
                best = min(best, int(w[k]) + int(x[ci[k]]))
        assert y[r] == best
    # empty rows give the min-plus identity
    assert all(y[r] == INF for r in range(60) if rp[r] == rp[r + 1])


def test_spmv_plustimes_vs_scipy():
    rng = np.random.default_rng(2)
    rp, ci, _ = random_csr(rng, 300, 8, weighted=False)
    vals = rng.random(ci.size)
    x = rng.random(300)
    want = sp.csr_matrix((vals, ci, rp), shape=(300, 300)) @ x
    got = oracle.spmv_plustimes_f64(rp, ci, vals, x)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=0)
    # a transposed operand would give a different answer on a non-symmetric matrix
    wrong = sp.csr_matrix((vals, ci, rp), shape=(300, 300)).T @ x
    assert not np.allclose(got, wrong)


# ------------------------------------------------------------------------- SSSP / BFS

def test_dijkstra_vs_floyd_warshall():
    rng = np.random.default_rng(3)
    for n in (1, 2, 17, 48):
        rp, ci, w = random_csr(rng, n, 3)
        fw = oracle.floyd_warshall(rp, ci, w)
        for s in range(n):
            d = oracle.dijkstra(rp, ci, w, s)
            want = np.where(fw[s] == np.iinfo(np.uint64).max, INF, fw[s]).astype(np.uint32)
            assert np.array_equal(d, want)


def test_grid_closed_form():
    rp, ci = grid_csr(9, 13)
    w = np.ones(ci.size, np.uint32)
    src = 4 * 13 + 6
    r0, c0 = divmod(src, 13)
    want = np.array([abs(v // 13 - r0) + abs(v % 13 - c0) for v in range(9 * 13)], np.uint32)
    assert np.array_equal(oracle.dijkstra(rp, ci, w, src), want)
    assert np.array_equal(oracle.bfs_queue(rp, ci, src), want)
    assert np.array_equal(oracle.sssp_advance(rp, ci, w, src)[0], want)
    tp, tj, tw = transpose(rp, ci, w)
    assert np.array_equal(oracle.minplus_blas(tp, tj, tw, src)[0], want)
    assert np.array_equal(oracle.minplus_blas(tp, tj, None, src)[0], want)


def test_bfs_closed_forms():
    n = 40
    # directed path 0 -> 1 -> ... -> n-1
    rp = np.arange(n + 1, dtype=np.uint32)
    rp[-1] = n - 1
    ci = np.arange(1, n, dtype=np.uint32)
    d = oracle.bfs_queue(rp, ci, 5)
    assert np.array_equal(d[5:], np.arange(n - 5, dtype=np.uint32))
    assert np.all(d[:5] == INF)
    # star: center 0 <-> leaves
    src = [0] * (n - 1) + list(range(1, n))
    dst = list(range(1, n)) + [0] * (n - 1)
    m = sp.csr_matrix((np.ones(len(src)), (src, dst)), shape=(n, n))
    m.sort_indices()
    d = oracle.bfs_queue(m.indptr, m.indices, 7)
    assert d[7] == 0 and d[0] == 1 and np.all(np.delete(d, [0, 7]) == 2)
    # complete graph: every other vertex at depth 1
    full = sp.csr_matrix(np.ones((n, n)) - np.eye(n))
    d = oracle.bfs_queue(full.indptr, full.indices, 3)
    assert d[3] == 0 and np.all(np.delete(d, 3) == 1)


def test_restated_engines_equal_dijkstra():
    rng = np.random.default_rng(4)
    for n, avg in ((30, 2), (200, 4), (1500, 6)):
        rp, ci, w = random_csr(rng, n, avg)
        tp, tj, tw = transpose(rp, ci, w)
        for s in (0, n // 2, n - 1):
            want = oracle.dijkstra(rp, ci, w, s)
            assert np.array_equal(oracle.sssp_advance(rp, ci, w, s)[0], want)
            assert np.array_equal(oracle.minplus_blas(tp, tj, tw, s)[0], want)
            bfs = oracle.bfs_queue(rp, ci, s)
            assert np.array_equal(oracle.minplus_blas(tp, tj, None, s)[0], bfs)
            assert np.array_equal(oracle.dijkstra(rp, ci, np.ones_like(w), s), bfs)


def test_openmp_advance_equals_dijkstra_at_every_thread_count():
    """The all-core restated advance loop (orc_sssp_advance_omp) reaches Dijkstra's distances whatever
    the thread count, with the serial loop's edge total on a graph whose frontiers it cannot reorder
    (a path)."""
    rng = np.random.default_rng(5)
    try:
        for threads in (1, 3, 8):
            oracle.set_threads(threads)
            for n, avg in ((40, 2), (800, 5), (20000, 8)):
                rp, ci, w = random_csr(rng, n, avg)
                for s in (0, n // 3):
                    want = oracle.dijkstra(rp, ci, w, s)
                    got, st = oracle.sssp_advance_omp(rp, ci, w, s)
                    assert np.array_equal(got, want), (threads, n, s)
                    assert np.array_equal(oracle.sssp_advance_omp(rp, ci, None, s)[0], oracle.bfs_queue(rp, ci, s))
        n = 50
        rp = np.arange(n + 1, dtype=np.uint32)
        rp[-1] = n - 1
        ci = np.arange(1, n, dtype=np.uint32)
        d, st = oracle.sssp_advance_omp(rp, ci, None, 0)
        assert np.array_equal(d, np.arange(n, dtype=np.uint32))
        assert st == oracle.sssp_advance(rp, ci, None, 0)[1]
    finally:
        oracle.set_threads(0)


def test_pagerank_iterates_do_not_depend_on_threads():
    """orc_pagerank_f64 splits its rows over the host threads: one thread and many give the same bits and
    the same iteration count."""
    rng = np.random.default_rng(6)
    rp, ci, w = random_csr(rng, 150000, 6)
    tp, tj, tw = transpose(rp, ci, w)
    try:
        oracle.set_threads(1)
        a, sa = oracle.pagerank_f64(rp, w, tp, tj, tw, 0.85, max_iter=30, tol=1e-9)
        oracle.set_threads(7)
        b, sb = oracle.pagerank_f64(rp, w, tp, tj, tw, 0.85, max_iter=30, tol=1e-9)
    finally:
        oracle.set_threads(0)
    assert sa == sb
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_blas_engine_iteration_count_is_hop_depth():
    """Synchronous rounds: the BLAS loop runs (max hops of a shortest-path tree) + 1 iterations; on a
    path graph with unit weights that is the path length + 1."""
    n = 25
    rp = np.arange(n + 1, dtype=np.uint32)
    rp[-1] = n - 1
    ci = np.arange(1, n, dtype=np.uint32)
    tp, tj, tw = transpose(rp, ci, None)
    d, st = oracle.minplus_blas(tp, tj, None, 0)
    assert np.array_equal(d, np.arange(n, dtype=np.uint32))
    assert st["iterations"] == n
    assert st["frontier_sum"] == n


# ------------------------------------------------------------------------- PageRank

def _pr_solve(rp, ci, alpha, seed=None, restart=0.15):
    """Fixed point r = b + alpha' P^T r solved with numpy.linalg.solve (no dangling redistribution)."""
    n = rp.size - 1
    deg = np.diff(rp.astype(np.int64))
    A = np.zeros((n, n))
    for u in range(n):
        for k in range(rp[u], rp[u + 1]):
            A[ci[k], u] += 1.0 / deg[u]
    if seed is None:
        b = np.full(n, (1 - alpha) / n)
        M = np.eye(n) - alpha * A
    else:
        b = np.zeros(n)
        b[seed] = restart
        M = np.eye(n) - (1 - restart) * A
    return np.linalg.solve(M, b)


def test_pagerank_vs_linear_solve():
    rng = np.random.default_rng(5)
    rp, ci, _ = random_csr(rng, 120, 5, weighted=False)
    tp, tj, _ = transpose(rp, ci, None)
    r, st = oracle.pagerank_f64(rp, None, tp, tj, None, 0.85, max_iter=500, tol=1e-14)
    np.testing.assert_allclose(r, _pr_solve(rp, ci, 0.85), rtol=1e-9, atol=1e-13)
    assert st["last_delta"] < 1e-14


def test_pagerank_uniform_on_regular_graph():
    n, k = 64, 6  # circulant: u -> u+1..u+k, k-regular in and out
    rp = (np.arange(n + 1) * k).astype(np.uint32)
    ci = np.array([(u + j) % n for u in range(n) for j in range(1, k + 1)], np.uint32)
    order = np.argsort(ci.reshape(n, k), axis=1)
    ci = np.take_along_axis(ci.reshape(n, k), order, axis=1).ravel().astype(np.uint32)
    tp, tj, _ = transpose(rp, ci, None)
    r, _ = oracle.pagerank_f64(rp, None, tp, tj, None, 0.85, max_iter=100, tol=0)
    np.testing.assert_allclose(r, 1.0 / n, rtol=1e-12)
    assert abs(r.sum() - 1.0) < 1e-12


def test_pagerank_weighted_uses_out_weight_sums():
    rng = np.random.default_rng(6)
    rp, ci, w = random_csr(rng, 80, 4, lo=1, hi=9)
    tp, tj, tw = transpose(rp, ci, w)
    r, _ = oracle.pagerank_f64(rp, w, tp, tj, tw, 0.85, max_iter=1000, tol=1e-15)
    n = 80
    A = np.zeros((n, n))
    for u in range(n):
        s = w[rp[u]:rp[u + 1]].sum()
        for k in range(rp[u], rp[u + 1]):
            A[ci[k], u] += w[k] / s
    want = np.linalg.solve(np.eye(n) - 0.85 * A, np.full(n, 0.15 / n))
    np.testing.assert_allclose(r, want, rtol=1e-9, atol=1e-14)


def test_ppr_two_cycle_closed_form():
    # 0 <-> 1: r0 = c + (1-c) r1, r1 = (1-c) r0  =>  r0 = c / (1 - (1-c)^2), r1 = (1-c) r0
    rp = np.array([0, 1, 2], np.uint32)
    ci = np.array([1, 0], np.uint32)
    tp, tj, _ = transpose(rp, ci, None)
    c = 0.15
    r, _ = oracle.pagerank_f64(rp, None, tp, tj, None, seed=0, restart=c, max_iter=2000, tol=1e-16)
    r0 = c / (1 - (1 - c) ** 2)
    np.testing.assert_allclose(r, [r0, (1 - c) * r0], rtol=1e-12)


def test_ppr_vs_linear_solve():
    rng = np.random.default_rng(7)
    rp, ci, _ = random_csr(rng, 90, 4, weighted=False)
    tp, tj, _ = transpose(rp, ci, None)
    r, _ = oracle.pagerank_f64(rp, None, tp, tj, None, seed=3, restart=0.15, max_iter=1000, tol=1e-15)
    np.testing.assert_allclose(r, _pr_solve(rp, ci, 0.85, seed=3), rtol=1e-9, atol=1e-14)


def test_pagerank_reference_convergence_quirk():
    """Mode 1 reproduces apps/pagerank_blas.cu:63-78: the first test compares 1/N with a zero vector,
    so with tol >= 1/N the loop exits before any iteration and returns ranks[0] = 1/N."""
    rng = np.random.default_rng(8)
    rp, ci, _ = random_csr(rng, 50, 3, weighted=False)
    tp, tj, _ = transpose(rp, ci, None)
    r, st = oracle.pagerank_f64(rp, None, tp, tj, None, 0.85, max_iter=100, tol=1.0 / 50, mode=1)
    assert st["iterations"] == 0 and np.allclose(r, 1.0 / 50)
    r, st = oracle.pagerank_f64(rp, None, tp, tj, None, 0.85, max_iter=1000, tol=1e-9, mode=1)
    want = _pr_solve(rp, ci, 0.85)
    assert np.max(np.abs(r - want)) < 1e-7
    assert st["iterations"] > 5


def test_pagerank_reference_mode_iterates_are_the_power_iteration():
    """Mode 1 updates in place and alternates buffers, but its iterates must be exactly the power
    iteration of mode 0: after K iterations it returns ranks[opt_idx], the iterate K-1."""
    rng = np.random.default_rng(9)
    rp, ci, _ = random_csr(rng, 70, 4, weighted=False)
    tp, tj, _ = transpose(rp, ci, None)
    for k in (2, 3, 8, 9):
        r1, st1 = oracle.pagerank_f64(rp, None, tp, tj, None, 0.85, max_iter=k, tol=0.0, mode=1)
        r0, st0 = oracle.pagerank_f64(rp, None, tp, tj, None, 0.85, max_iter=k - 1, tol=0.0, mode=0)
        assert st1["iterations"] == k and st0["iterations"] == k - 1
        assert np.array_equal(r1, r0)


# ------------------------------------------------------------------------- golden vectors

def test_golden_vectors():
    path = os.path.join(GOLDEN, "oracle_golden.json")
    with open(path) as f:
        cases = json.load(f)
    for c in cases:
        rp = np.array(c["row_offsets"], np.uint32)
        ci = np.array(c["col_indices"], np.uint32)
        w = np.array(c["weights"], np.uint32) if c.get("weights") is not None else None
        if c["kind"] == "sssp":
            assert oracle.dijkstra(rp, ci, w, c["src"]).tolist() == c["expect"]
            assert oracle.sssp_advance(rp, ci, w, c["src"])[0].tolist() == c["expect"]
        elif c["kind"] == "bfs":
            assert oracle.bfs_queue(rp, ci, c["src"]).tolist() == c["expect"]
        elif c["kind"] == "spmv_minplus":
            x = np.array(c["x"], np.uint32)
            assert oracle.spmv_minplus_u32(rp, ci, w, x).tolist() == c["expect"]
