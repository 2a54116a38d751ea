"""Whole-run parity on the GPU: BFS depths and SSSP distances bit-exact against the oracle on every
engine, PageRank / personalized PageRank within L1 <= 1e-5 of the FP64 oracle.

Reference semantics: apps/sssp_adv.cu:17-72 (advance), apps/sssp_blas.cu:17-84 (BLAS),
apps/pagerank_blas.cu:16-125 (PageRank); the oracle restates them (oracle/g1_oracle.c).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _graphs():
    import graphone_amd as g1
    out = []
    h = g1.HostGraph.rmat(14, 16, seed=11).assign_weights(1, 255, seed=12)
    out.append(("rmat14", h, False))
    h = g1.HostGraph.grid(64, 64, 0.6, seed=3).assign_weights(1, 64, seed=4, symmetric=True)
    out.append(("grid64", h, True))
    h = g1.HostGraph.rmat(10, 4, seed=5, undirected=True).assign_weights(1, 9, seed=6, symmetric=True)
    out.append(("rmat10u", h, True))
    return out


@pytest.fixture(scope="module")
def graphs(gpu):
    import graphone_amd as g1
    res = []
    for name, h, sym in _graphs():
        res.append((name, h, g1.Graph(h, transpose=True, symmetric=sym)))
    return res


def _sources(h):
    deg = h.out_degrees()
    return [int(np.argmax(deg)), 1 % h.n, int(np.flatnonzero(deg == 0)[0]) if np.any(deg == 0) else 0]


@pytest.mark.parametrize("engine", ["advance", "diropt", "merge", "csrvec"])
def test_bfs_bitexact(graphs, engine):
    for name, h, g in graphs:
        for src in _sources(h):
            want = oracle.bfs_queue(h.row_offsets, h.col_indices, src)
            got = g.bfs(src, engine)
            assert np.array_equal(got.values, want), (name, src, engine)
            assert got.stats["reached"] == int(np.sum(want != oracle.INF))


@pytest.mark.parametrize("engine", ["advance", "delta", "merge", "csrvec"])
def test_sssp_bitexact(graphs, engine):
    for name, h, g in graphs:
        for src in _sources(h):
            want = oracle.dijkstra(h.row_offsets, h.col_indices, h.weights, src)
            got = g.sssp(src, engine)
            assert np.array_equal(got.values, want), (name, src, engine)


@pytest.mark.parametrize("width", ["1", "5", "64", "100000"])
@pytest.mark.parametrize("switch", ["0", "0.01", "1"])
def test_sssp_delta_schedules(graphs, monkeypatch, width, switch):
    """Delta-stepping with bucket widths from 1 (every edge heavy, one bucket per distance) to wider than any
    path (every edge light: one bucket), leaving the buckets for the advance loop after the first heavy
    phase, after one that relaxes 1 % of the edges, or never (up to its bucket cap): every schedule gives
    Dijkstra's distances, from the hub, a low-degree vertex and an isolated one."""
    monkeypatch.setenv("G1_SSSP_DELTA", width)
    monkeypatch.setenv("G1_SSSP_DELTA_SWITCH", switch)
    name, h, g = graphs[0]
    deg = h.out_degrees()
    pos = np.flatnonzero(deg > 0)
    for src in (int(np.argmax(deg)), int(pos[np.argmin(deg[pos])]), int(np.flatnonzero(deg == 0)[0])):
        want = oracle.dijkstra(h.row_offsets, h.col_indices, h.weights, src)
        got = g.sssp(src, "delta")
        assert np.array_equal(got.values, want), (name, src, width, switch)
        assert got.stats["reached"] == int(np.sum(want != oracle.INF))


@pytest.mark.parametrize("hi,width,pull,rows,edges", [
    (255, "16", "0.01", "1", "1"), (255, "16", "0.01", "1", "4"), (255, "16", "0.01", "1", "16"),
    (255, "16", "0.01", "0", "1"), (1000, "254", "0.01", "1", "1"), (1000, "300", "0.01", "1", "1"),
    (255, "16", "0", "1", "1")])
def test_sssp_delta_heavy_steps(gpu, monkeypatch, hi, width, pull, rows, edges):
    """The heavy phase as a pull step over a one-byte x (bucket width <= 254; over the rows at or past the
    bucket's end with 1, 4 or 16 edges per lane per round, or as the merge-path SpMV over every row), over the u32 x (a
    wider bucket), and as a push launch (pull steps off): Dijkstra's distances from the hub and from a
    low-degree vertex, with the pull step taken where it is enabled. Width 254 puts bucket offsets next to
    the byte's two markers (254: below the bucket, 255: past it)."""
    import graphone_amd as g1
    monkeypatch.setenv("G1_SSSP_DELTA", width)
    monkeypatch.setenv("G1_SSSP_PULL", pull)
    monkeypatch.setenv("G1_SSSP_DELTA_PULL_ROWS", rows)
    monkeypatch.setenv("G1_SSSP_DELTA_PULL_EDGES", edges)
    h = g1.HostGraph.rmat(14, 16, seed=41).assign_weights(1, hi, seed=42)
    g = g1.Graph(h)
    deg = h.out_degrees()
    pos = np.flatnonzero(deg > 0)
    for src in (int(np.argmax(deg)), int(pos[np.argmin(deg[pos])])):
        want = oracle.dijkstra(h.row_offsets, h.col_indices, h.weights, src)
        got = g.sssp(src, "delta")
        assert np.array_equal(got.values, want), (src, hi, width, pull, rows, edges)
        if src == int(np.argmax(deg)):
            assert (got.stats["pull_iterations"] >= 1) == (pull != "0"), got.stats


def test_sssp_delta_pull_workgroup_rows(gpu, monkeypatch):
    """The heavy pull over rows with its workgroup path taken by every row past the long-row bound (128
    in-edges here, G1_SSSP_DELTA_PULL_HUGE; the handle's long-row list is built with it) and pull steps from
    1 % of m: from low-degree sources a heavy pull then visits hubs that are not settled yet. Dijkstra's
    distances from those sources and from the hub."""
    import graphone_amd as g1
    monkeypatch.setenv("G1_SSSP_DELTA_PULL_HUGE", "1")
    monkeypatch.setenv("G1_SSSP_PULL", "0.01")
    h = g1.HostGraph.rmat(14, 16, seed=43).assign_weights(1, 255, seed=44)
    g = g1.Graph(h)
    deg = h.out_degrees()
    pos = np.flatnonzero(deg > 0)
    rng = np.random.default_rng(45)
    pulls = 0
    for src in [int(np.argmax(deg))] + [int(s) for s in rng.choice(pos[deg[pos] <= 2], 4, replace=False)]:
        want = oracle.dijkstra(h.row_offsets, h.col_indices, h.weights, src)
        got = g.sssp(src, "delta")
        assert np.array_equal(got.values, want), src
        pulls += got.stats["pull_iterations"]
    assert pulls >= 1


@pytest.mark.parametrize("sym", [False, True])
def test_hub_copy_runs_match_oracle(gpu, monkeypatch, sym):
    """Delta-stepping SSSP and direction-optimizing BFS on the hub-first copy (prepared with G1_WORK_HUB on
    a graph below the size the engines build it at by themselves) equal Dijkstra and the BFS queue in the
    graph's own ids, with the reached-vertex stats of the runs without it (a second handle, and G1_HUB=0 on
    the prepared one); directed with its transpose, and symmetric."""
    import graphone_amd as g1
    from graphone_amd import _lib
    h = g1.HostGraph.rmat(14, 16, seed=61, undirected=sym).assign_weights(1, 255, seed=62, symmetric=sym)
    g = g1.Graph(h, transpose=True, symmetric=sym)
    g0 = g1.Graph(h, transpose=True, symmetric=sym)
    want_bytes = g.work_bytes(_lib.G1_WORK_HUB)
    before = g.device_bytes()
    g.prepare(_lib.G1_WORK_HUB)
    grew = g.device_bytes() - before
    assert grew <= want_bytes <= grew + 4 * h.n, (want_bytes, grew)
    assert g.work_bytes(_lib.G1_WORK_HUB) == 0
    deg = h.out_degrees()
    rng = np.random.default_rng(63)
    srcs = [int(np.argmax(deg)), 1, int(np.flatnonzero(deg == 0)[0]) if np.any(deg == 0) else 2]
    srcs += [int(v) for v in rng.choice(h.n, 3, replace=False)]
    total = g.device_bytes()
    for src in srcs:
        want_d = oracle.dijkstra(h.row_offsets, h.col_indices, h.weights, src)
        want_b = oracle.bfs_queue(h.row_offsets, h.col_indices, src)
        d, d0 = g.sssp(src, "delta"), g0.sssp(src, "delta")
        b, b0 = g.bfs(src, "diropt"), g0.bfs(src, "diropt")
        assert np.array_equal(d.values, want_d), src
        assert np.array_equal(b.values, want_b), src
        for got, ref in ((d, d0), (b, b0)):
            assert got.stats["reached"] == ref.stats["reached"], src
            assert got.stats["reached_edges"] == ref.stats["reached_edges"], src
    assert g.device_bytes() == total  # the runs allocate nothing more
    g.profile(True)
    g.sssp(srcs[0], "delta")
    names = {e["name"] for e in g.read_profile()}
    g.profile(False)
    assert "sssp_delta_light" in names, names
    monkeypatch.setenv("G1_HUB", "0")  # the graph's own ids again
    assert np.array_equal(g.sssp(srcs[0], "delta").values,
                          oracle.dijkstra(h.row_offsets, h.col_indices, h.weights, srcs[0]))
    assert np.array_equal(g.bfs(srcs[0], "diropt").values, oracle.bfs_queue(h.row_offsets, h.col_indices, srcs[0]))


def test_hub_copy_edge_cases(gpu):
    """The hub-first copy of an unweighted graph: direction-optimizing BFS on it equals the BFS queue, and
    delta-stepping SSSP is refused as on the graph's own ids. Views and road-like graphs have no copy (the
    HUB set adds nothing and prepares nothing)."""
    import graphone_amd as g1
    from graphone_amd import _lib
    h = g1.HostGraph.rmat(13, 16, seed=81)
    g = g1.Graph(h, transpose=True)
    assert g.work_bytes(_lib.G1_WORK_HUB) > 0
    g.prepare(_lib.G1_WORK_HUB)
    assert g.work_bytes(_lib.G1_WORK_HUB) == 0
    deg = h.out_degrees()
    for src in (int(np.argmax(deg)), int(np.flatnonzero(deg == 0)[0]) if np.any(deg == 0) else 3):
        assert np.array_equal(g.bfs(src, "diropt").values, oracle.bfs_queue(h.row_offsets, h.col_indices, src)), src
    with pytest.raises(g1.G1Error):
        g.sssp(0, "delta")
    v = g.view()
    assert v.work_bytes(_lib.G1_WORK_HUB) == 0
    before = v.device_bytes()
    v.prepare(_lib.G1_WORK_HUB)
    assert v.device_bytes() == before
    src = int(np.argmax(deg))
    assert np.array_equal(v.bfs(src, "diropt").values, oracle.bfs_queue(h.row_offsets, h.col_indices, src))
    v.close()
    r = g1.Graph(g1.HostGraph.grid(32, 32, 0.7, seed=83).assign_weights(1, 9, seed=84, symmetric=True), symmetric=True)
    assert r.work_bytes(_lib.G1_WORK_HUB) == 0
    before = r.device_bytes()
    r.prepare(_lib.G1_WORK_HUB)
    assert r.device_bytes() == before


def test_hub_copy_built_by_the_engines(gpu):
    """From 2^20 vertices the two engines build the hub-first copy on their own first run (the handle's
    device bytes grow by what G1_WORK_HUB reports) and stay bit-exact: RMAT20 from the hub and from a
    low-degree vertex against Dijkstra and the BFS queue."""
    import graphone_amd as g1
    from graphone_amd import _lib
    h = g1.HostGraph.rmat(20, 8, 0.57, 0.19, 0.19, seed=71).assign_weights(1, 255, seed=72)
    assert h.n >= 1 << 20
    g = g1.Graph(h, transpose=True)
    deg = h.out_degrees()
    srcs = [int(np.argmax(deg)), int(np.flatnonzero(deg == 1)[0])]
    want_bytes, before = g.work_bytes(_lib.G1_WORK_HUB), g.device_bytes()
    b = g.bfs(srcs[0], "diropt")
    assert np.array_equal(b.values, oracle.bfs_queue(h.row_offsets, h.col_indices, srcs[0]))
    for src in srcs:
        want = oracle.dijkstra(h.row_offsets, h.col_indices, h.weights, src)
        got = g.sssp(src, "delta")
        assert np.array_equal(got.values, want), src
        assert got.stats["reached"] == int(np.count_nonzero(want != 0xFFFFFFFF)), src
        assert got.stats["reached_edges"] == int(deg[want != 0xFFFFFFFF].sum()), src
        assert np.array_equal(g.bfs(src, "diropt").values, oracle.bfs_queue(h.row_offsets, h.col_indices, src)), src
    grew = g.device_bytes() - before  # the copy and its work areas, the long-row list within its bound
    assert grew <= want_bytes <= grew + 4 * h.n, (want_bytes, grew)
    assert g.work_bytes(_lib.G1_WORK_HUB) == 0


def test_sssp_delta_zero_weights(gpu):
    """Weights 0..3 (a quarter of the edges weigh nothing, so improvements stay inside their bucket):
    delta-stepping from the hub and from vertex 1 equals Dijkstra."""
    import graphone_amd as g1
    h = g1.HostGraph.rmat(13, 16, 0.57, 0.19, 0.19, seed=31).assign_weights(0, 3, seed=32)
    g = g1.Graph(h)
    for src in (int(np.argmax(h.out_degrees())), 1):
        want = oracle.dijkstra(h.row_offsets, h.col_indices, h.weights, src)
        assert np.array_equal(g.sssp(src, "delta").values, want), src


def test_sssp_engines_match_restated_engines(graphs):
    """The push engine's final distances equal the restated advance loop; the pull engine's iteration
    count equals the restated BLAS loop's (synchronous rounds are deterministic)."""
    name, h, g = graphs[0]
    src = int(np.argmax(h.out_degrees()))
    d_adv, _ = oracle.sssp_advance(h.row_offsets, h.col_indices, h.weights, src)
    adv = g.sssp(src, "advance")
    assert np.array_equal(adv.values, d_adv)
    # from the hub, the densest frontier holds most of the graph's edges: it runs as a pull step, and the
    # hub's own frontier (>= 0.3 m out-edges) relaxes without the winner test
    assert adv.stats["pull_iterations"] >= 1, adv.stats
    assert adv.stats["dense_iterations"] >= 1, adv.stats
    t = h.transpose()
    d_blas, st = oracle.minplus_blas(t.row_offsets, t.col_indices, t.weights, src)
    got = g.sssp(src, "merge")
    assert np.array_equal(got.values, d_blas)
    assert got.stats["iterations"] == st["iterations"]
    assert got.stats["frontier_sum"] == st["frontier_sum"]


@pytest.mark.parametrize("engine", ["merge", "csrvec"])
def test_pagerank_l1(graphs, engine):
    for name, h, g in graphs:
        t = h.transpose()
        want, ost = oracle.pagerank_f64(h.row_offsets, None, t.row_offsets, t.col_indices, None, 0.85,
                                        max_iter=50, tol=1e-7)
        # unit weights: build an unweighted device graph
        import graphone_amd as g1
        hu = g1.HostGraph.from_csr(h.row_offsets, h.col_indices)
        gu = g1.Graph(hu)
        got = gu.pagerank(0.85, 50, 1e-7, "l1", engine)
        l1 = float(np.abs(got.values.astype(np.float64) - want).sum())
        assert l1 <= 1e-5, (name, l1)
        assert abs(got.stats["iterations"] - ost["iterations"]) <= 1, (name, got.stats, ost)


def test_pagerank_reference_convergence_rule(graphs):
    """G1_PR_MAX_ABS reproduces the reference's loop: first test against a zero vector, in-place
    update, result = ranks[opt_idx] (apps/pagerank_blas.cu:63-78,169)."""
    import graphone_amd as g1
    name, h, _ = graphs[1]
    hu = g1.HostGraph.from_csr(h.row_offsets, h.col_indices)
    gu = g1.Graph(hu)
    t = h.transpose()
    want, ost = oracle.pagerank_f64(h.row_offsets, None, t.row_offsets, t.col_indices, None, 0.85,
                                    max_iter=1000, tol=1e-6, mode=1)
    got = gu.pagerank(0.85, 1000, 1e-6, "max", "merge")
    assert got.stats["iterations"] == ost["iterations"]
    assert float(np.abs(got.values - want).max()) < 1e-6


def test_ppr_l1(graphs):
    import graphone_amd as g1
    name, h, _ = graphs[0]
    hu = g1.HostGraph.from_csr(h.row_offsets, h.col_indices)
    gu = g1.Graph(hu)
    t = h.transpose()
    seed = int(np.argmax(h.out_degrees()))
    want, ost = oracle.pagerank_f64(h.row_offsets, None, t.row_offsets, t.col_indices, None, seed=seed,
                                    restart=0.15, max_iter=50, tol=1e-7)
    got = gu.ppr(seed, 0.15, 50, 1e-7)
    l1 = float(np.abs(got.values.astype(np.float64) - want).sum())
    assert l1 <= 1e-5, l1


def test_bad_source_rejected(graphs):
    import graphone_amd as g1
    name, h, g = graphs[0]
    with pytest.raises(g1.G1Error):
        g.bfs(h.n + 5)


def test_grid213_sssp_from_vertex_1(gpu):
    """The sssp_adv workload: 213 x 213 grid, p = 0.6, symmetric weights in [1, 64], source 1; the
    advance engine (the road engine on this low-degree graph) and the merge-path BLAS engine both equal
    the oracle's Dijkstra bit for bit."""
    import graphone_amd as g1
    h = g1.HostGraph.grid(213, 213, 0.6, seed=7).assign_weights(1, 64, seed=8, symmetric=True)
    g = g1.Graph(h, symmetric=True)
    want = oracle.dijkstra(h.row_offsets, h.col_indices, h.weights, 1)
    for eng in ("advance", "merge", "csrvec"):
        assert np.array_equal(g.sssp(1, eng).values, want), eng


def test_road_sssp_past_16_bit_distances(gpu):
    """A small grid (its distances fit the on-chip road kernel's LDS) with weights up to 60000: distances
    pass 0xFFFF, the on-chip kernel hands the query to the global-memory road kernel, and the result still
    equals Dijkstra; the same grid with light weights stays on chip."""
    import graphone_amd as g1
    for hi in (60000, 9):
        h = g1.HostGraph.grid(40, 40, 0.7, seed=11).assign_weights(1, hi, seed=12, symmetric=True)
        want = oracle.dijkstra(h.row_offsets, h.col_indices, h.weights, 5)
        if hi > 65535 // 40:
            assert int(want[want != 0xFFFFFFFF].max()) > 0xFFFF
        got = g1.Graph(h, symmetric=True).sssp(5, "advance")
        assert np.array_equal(got.values, want), hi


def test_pagerank_rmat20_l1(gpu):
    """PageRank on unweighted RMAT scale 20 (16.1M edges) within L1 1e-5 of the FP64 oracle."""
    import graphone_amd as g1
    h = g1.HostGraph.rmat(20, 16, 0.57, 0.19, 0.19, seed=20)
    t = h.transpose()
    want, ost = oracle.pagerank_f64(h.row_offsets, None, t.row_offsets, t.col_indices, None, 0.85,
                                    max_iter=50, tol=1e-7)
    got = g1.Graph(h).pagerank(0.85, 50, 1e-7, "l1", "merge")
    diff = np.abs(got.values.astype(np.float64) - want)
    assert float(diff.sum()) <= 1e-5, (float(diff.sum()), float(diff.max()))
    assert abs(got.stats["iterations"] - ost["iterations"]) <= 1


def test_work_areas_sized_and_prepared_ahead(gpu):
    """g1_graph_work_bytes gives the bytes each engine's work areas add (exact, the long-row lists within
    their bound), g1_graph_prepare allocates them, and afterwards the engines run without allocating."""
    import graphone_amd as g1
    from graphone_amd import _lib
    h = g1.HostGraph.rmat(14, 16, seed=51).assign_weights(1, 255, seed=52)
    g = g1.Graph(h)
    n = h.n
    for work, bound in ((_lib.G1_WORK_PUSH, 0), (_lib.G1_WORK_PULL, 4 * n), (_lib.G1_WORK_RANK, 0),
                        (_lib.G1_WORK_RELABEL, 4 * n), (_lib.G1_WORK_DELTA, 0), (_lib.G1_WORK_HUB, 4 * n)):
        want = g.work_bytes(work)
        before = g.device_bytes()
        g.prepare(work)
        grew = g.device_bytes() - before
        assert grew <= want <= grew + bound, (work, want, grew)
        assert g.work_bytes(work) == 0, work
    assert g.work_bytes(_lib.G1_WORK_ALL & ~_lib.G1_WORK_ROAD) == 0
    total = g.device_bytes()
    src = int(np.argmax(h.out_degrees()))
    for eng in ("advance", "delta", "merge", "csrvec"):
        g.sssp(src, eng)
    g.bfs(src, "diropt")
    g.pagerank(0.85, 20, 1e-7, "l1", "merge")
    g.pagerank(0.85, 20, 1e-7, "l1", "csrvec")
    g.pagerank(0.85, 20, 1e-6, "max", "merge")
    g.ppr(src, 0.15, 20, 1e-7)
    assert g.device_bytes() == total
    # a road-like graph: the road engine's work
    r = g1.Graph(g1.HostGraph.grid(32, 32, 0.7, seed=53).assign_weights(1, 9, seed=54, symmetric=True), symmetric=True)
    want = r.work_bytes(_lib.G1_WORK_ROAD)
    before = r.device_bytes()
    r.prepare(_lib.G1_WORK_ROAD)
    assert r.device_bytes() - before == want
    r.sssp(3, "advance")
    assert r.device_bytes() - before == want
    with pytest.raises(g1.G1Error):
        g.prepare(128)


def test_views_run_concurrently(graphs):
    """Two views of one graph run SSSP (with its pull steps) and personalized PageRank on their own
    streams from two threads at once; every result equals the serial run (views own their work areas,
    the SpMV tile records included)."""
    import threading
    import graphone_amd as g1
    name, h, g = graphs[0]
    srcs = [int(np.argmax(h.out_degrees())), 1, 7, 42]
    want_d = [g.sssp(s).values for s in srcs]
    want_r = [g.ppr(s, 0.15, 30, 1e-7).values for s in srcs]
    views = [g.view(), g.view()]
    streams = [g1.Stream(), g1.Stream()]
    got = [[], []]

    def work(i):
        for _ in range(3):
            for s in srcs:
                if i == 0:
                    got[i].append(views[i].sssp(s, stream=streams[i]).values)
                else:
                    got[i].append(views[i].ppr(s, 0.15, 30, 1e-7, stream=streams[i]).values)

    ts = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    for k, d in enumerate(got[0]):
        assert np.array_equal(d, want_d[k % len(srcs)]), k
    for k, r in enumerate(got[1]):
        assert np.array_equal(r.view(np.uint32), want_r[k % len(srcs)].view(np.uint32)), k
    for v in views:
        v.close()
