"""Parity at the configured sizes, on the GPU.

- RMAT scale 22 (edge factor 16, a/b/c/d .57/.19/.19/.05, seed 1, u32 weights [1, 255] seed 2), source =
  the vertex of highest out-degree: BFS depths on every engine against the oracle's queue BFS, SSSP
  distances on every engine against the oracle's Dijkstra, bit for bit. At this size the push engines run
  their multi-tile loops (tens of 2048-edge tiles per workgroup on the hub iterations), the dense SSSP
  iterations run without the winner test, the densest as a pull step, and the direction-optimizing BFS
  takes its bottom-up levels over several rounds of workgroups.
- Eight columns of a 64-source batch on the same graph against Dijkstra.
- RMAT scale 24 PageRank (unweighted, damping 0.85, 50 iterations or L1 < 1e-7) on both SpMV engines:
  L1 <= 1e-5 against the FP64 oracle, max-abs reported in the failure message.

Reference semantics: apps/sssp_adv.cu:48-72 (relax/filter), apps/sssp_blas.cu:48-84 (BLAS SSSP),
apps/pagerank_blas.cu:36-125 (PageRank).
"""
import numpy as np
import pytest

import oracle

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


@pytest.fixture(scope="module")
def rmat22(gpu):
    import graphone_amd as g1
    h = g1.HostGraph.rmat(22, 16, 0.57, 0.19, 0.19, seed=1).assign_weights(1, 255, seed=2)
    deg = h.out_degrees()
    src = int(np.argmax(deg))
    g = g1.Graph(h, transpose=True)
    print(f"rmat22: n={h.n} m={h.m} source={src} deg={int(deg[src])}", flush=True)
    want_bfs = oracle.bfs_queue(h.row_offsets, h.col_indices, src)
    want_sssp = oracle.dijkstra(h.row_offsets, h.col_indices, h.weights, src)
    yield h, g, src, want_bfs, want_sssp
    g.close()


def _diff(got, want):
    bad = np.flatnonzero(got != want)
    return f"{bad.size} differ, first {bad[:4].tolist()}: got {got[bad[:4]].tolist()} want {want[bad[:4]].tolist()}"


@pytest.mark.parametrize("engine", ["advance", "diropt", "merge", "csrvec"])
def test_rmat22_bfs_bitexact(rmat22, engine):
    h, g, src, want, _ = rmat22
    got = g.bfs(src, engine)
    assert np.array_equal(got.values, want), (engine, _diff(got.values, want))
    assert got.stats["reached"] == int(np.sum(want != oracle.INF))
    if engine == "diropt":
        assert got.stats["pull_iterations"] >= 1, got.stats


@pytest.mark.parametrize("engine", ["advance", "delta", "merge", "csrvec"])
def test_rmat22_sssp_bitexact(rmat22, engine):
    h, g, src, _, want = rmat22
    got = g.sssp(src, engine)
    assert np.array_equal(got.values, want), (engine, _diff(got.values, want))
    if engine == "advance":
        # the hub iterations: one pull step and the dense iterations without the winner test
        assert got.stats["pull_iterations"] >= 1, got.stats
        assert got.stats["dense_iterations"] >= 1, got.stats


@pytest.mark.parametrize("kind", ["bfs", "sssp"])
def test_rmat22_own_ids_bitexact(rmat22, monkeypatch, kind):
    """The two engines that run on the hub-first copy at this size (direction-optimizing BFS, delta-stepping
    SSSP), on the graph's own ids instead (G1_HUB=0): the same results and reached counts."""
    h, g, src, want_bfs, want_sssp = rmat22
    monkeypatch.setenv("G1_HUB", "0")
    got = g.bfs(src, "diropt") if kind == "bfs" else g.sssp(src, "delta")
    want = want_bfs if kind == "bfs" else want_sssp
    assert np.array_equal(got.values, want), (kind, _diff(got.values, want))
    assert got.stats["reached"] == int(np.sum(want != oracle.INF))


def test_rmat22_sssp_other_sources(rmat22):
    """Two more sources (a low-degree vertex and a vertex of median degree among the reached ones): other
    frontier shapes through the same engines."""
    h, g, src, want_bfs, _ = rmat22
    deg = h.out_degrees()
    reached = np.flatnonzero((want_bfs != oracle.INF) & (deg > 0))
    order = np.argsort(deg[reached], kind="stable")
    for s in (int(reached[order[0]]), int(reached[order[order.size // 2]])):
        want = oracle.dijkstra(h.row_offsets, h.col_indices, h.weights, s)
        for engine in ("advance", "delta"):
            got = g.sssp(s, engine)
            assert np.array_equal(got.values, want), (s, engine, _diff(got.values, want))
        wb = oracle.bfs_queue(h.row_offsets, h.col_indices, s)
        gb = g.bfs(s, "diropt")
        assert np.array_equal(gb.values, wb), (s, _diff(gb.values, wb))


def test_rmat22_batch_columns_vs_dijkstra(rmat22):
    """64 sources drawn as the batched workload draws them (uniform over vertices with out-degree > 0,
    without replacement); eight of the 64 batched columns against Dijkstra."""
    h, g, _, _, _ = rmat22
    rng = np.random.default_rng(3)
    deg = h.out_degrees()
    srcs = rng.choice(np.flatnonzero(deg > 0), size=64, replace=False).astype(np.uint32)
    res = g.sssp_batch(srcs, k=64)
    for c in range(0, 64, 8):
        want = oracle.dijkstra(h.row_offsets, h.col_indices, h.weights, int(srcs[c]))
        assert np.array_equal(res.values[c], want), (c, int(srcs[c]), _diff(res.values[c], want))


@pytest.fixture(scope="module")
def rmat24(gpu):
    import graphone_amd as g1
    h = g1.HostGraph.rmat(24, 16, 0.57, 0.19, 0.19, seed=11)
    t = h.transpose()
    oracle.set_threads(0)
    want, st = oracle.pagerank_f64(h.row_offsets, None, t.row_offsets, t.col_indices, None, 0.85,
                                   max_iter=50, tol=1e-7)
    del t
    g = g1.Graph(h)
    print(f"rmat24: n={h.n} m={h.m} oracle iterations={st['iterations']}", flush=True)
    yield h, g, want, st
    g.close()


@pytest.mark.parametrize("engine", ["merge", "csrvec"])
def test_rmat24_pagerank_vs_f64(rmat24, engine):
    h, g, want, st = rmat24
    got = g.pagerank(0.85, 50, 1e-7, "l1", engine)
    diff = np.abs(got.values.astype(np.float64) - want)
    l1, mx = float(diff.sum()), float(diff.max())
    print(f"rmat24 pagerank {engine}: L1 {l1:.3e} max-abs {mx:.3e} iterations {got.stats['iterations']}",
          flush=True)
    assert l1 <= 1e-5, (engine, l1, mx)
    assert abs(got.stats["iterations"] - st["iterations"]) <= 1, (got.stats, st)
