"""The query server on the GPU: every served result equals its single-source run (and the oracle), the
same alone and in a mix of queries on several resident graphs; HBM-budget admission holds queries back
until results are fetched; cancellation.

Served BFS/SSSP results are bit-exact against the oracle's BFS queue / Dijkstra; served PPR results are
bit-identical to g1_ppr's (restart 0.15) and within L1 1e-5 of the FP64 oracle.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def graphs(gpu):
    import graphone_amd as g1
    hs = [g1.HostGraph.rmat(12, 16, seed=31).assign_weights(1, 255, seed=32),
          g1.HostGraph.grid(40, 40, 0.6, seed=33).assign_weights(1, 255, seed=34, symmetric=True)]
    return [(h, g1.Graph(h, symmetric=(i == 1))) for i, h in enumerate(hs)]


def _want(h, kind, src):
    if kind == "bfs":
        return oracle.bfs_queue(h.row_offsets, h.col_indices, src)
    if kind == "sssp":
        return oracle.dijkstra(h.row_offsets, h.col_indices, h.weights, src)
    return None


def _queries(graphs, count, seed):
    rng = np.random.default_rng(seed)
    kinds = rng.choice(["bfs", "sssp", "ppr"], size=count, p=[0.4, 0.4, 0.2])
    out = []
    for k in kinds:
        gid = int(rng.integers(0, len(graphs)))
        out.append((gid, str(k), int(rng.integers(0, graphs[gid][0].n))))
    return out


def _check(graphs, gid, kind, src, res):
    h, g = graphs[gid]
    if kind == "ppr":
        single = g.ppr(src, 0.15, 50, 1e-7).values
        assert np.array_equal(res.view(np.uint32), single.view(np.uint32)), (gid, src)
    else:
        assert np.array_equal(res, _want(h, kind, src)), (gid, kind, src)


@pytest.mark.parametrize("batch_k", [16, 64])
def test_served_results_match_single_source(graphs, batch_k):
    """batch_k 64 runs SSSP sessions on the k = 64 wavefront kernel, narrower ones on the group kernel."""
    import graphone_amd as g1
    srv = g1.Server(batch_k=batch_k, batch_min_sssp=2, batch_min_ppr=2)
    ids = [srv.add_graph(g) for _, g in graphs]
    qs = _queries(graphs, 48, seed=1)
    qids = [srv.submit(ids[gid], kind, src) for gid, kind, src in qs]
    paths = set()
    for (gid, kind, src), qid in zip(qs, qids):
        st = srv.wait(qid, 60000)
        assert st["state_name"] == "done", st
        paths.add(st["path"])
        _check(graphs, gid, kind, src, srv.fetch(qid, kind))
    assert paths == {0, 1}, "both the single-source and the batched path should have run"
    s = srv.stats()
    assert s["completed"] == len(qs) and s["failed"] == 0 and s["batched"] > 0
    srv.close()


def test_same_result_alone_and_in_the_mix(graphs):
    import graphone_amd as g1
    probes = [(0, "sssp", 5), (1, "bfs", 17), (0, "ppr", 9), (1, "ppr", 3)]
    alone = []
    for gid, kind, src in probes:
        srv = g1.Server(batch_k=8)
        ids = [srv.add_graph(g) for _, g in graphs]
        q = srv.submit(ids[gid], kind, src)
        assert srv.wait(q)["state_name"] == "done"
        alone.append(srv.fetch(q, kind))
        srv.close()
    srv = g1.Server(batch_k=8, batch_min_sssp=2, batch_min_ppr=2)
    ids = [srv.add_graph(g) for _, g in graphs]
    mix = _queries(graphs, 40, seed=2)
    order = list(mix[:20]) + probes + list(mix[20:])
    qids = [srv.submit(ids[gid], kind, src) for gid, kind, src in order]
    results = {}
    for (gid, kind, src), qid in zip(order, qids):
        assert srv.wait(qid)["state_name"] == "done"
        results[qid] = srv.fetch(qid, kind)
    for i, (gid, kind, src) in enumerate(probes):
        got = results[qids[20 + i]]
        assert np.array_equal(got.view(np.uint32), alone[i].view(np.uint32)), (gid, kind, src)
        _check(graphs, gid, kind, src, got)
    # PPR against the FP64 oracle
    h, _ = graphs[0]
    t = h.transpose()
    want, _ = oracle.pagerank_f64(h.row_offsets, h.weights, t.row_offsets, t.col_indices, t.weights, 0.85,
                                  seed=9, restart=0.15, max_iter=50, tol=1e-7)
    assert float(np.abs(results[qids[22]].astype(np.float64) - want).sum()) <= 1e-5
    srv.close()


def test_budget_admission_waits_for_fetch(graphs):
    import graphone_amd as g1
    h, g = graphs[0]
    graph_bytes = g.device_bytes()
    # room for the graph and three results only
    srv = g1.Server(hbm_budget_bytes=graph_bytes + 3 * h.n * 4 + 64, batch_min_sssp=1000, batch_min_ppr=1000)
    gid = srv.add_graph(g)
    qids = [srv.submit(gid, "bfs", s) for s in range(8)]
    for q in qids[:3]:
        assert srv.wait(q)["state_name"] == "done"
    assert srv.poll(qids[3])["state_name"] == "queued"
    assert srv.stats()["admission_waits"] >= 1
    for i, q in enumerate(qids):
        assert srv.wait(q)["state_name"] == "done"
        assert np.array_equal(srv.fetch(q, "bfs"), _want(h, "bfs", i))
    s = srv.stats()
    assert s["completed"] == 8 and s["hbm_used"] == graph_bytes
    srv.close()


def test_raising_the_budget_admits_waiting_queries(graphs):
    """A budget with room for one result holds the other queries back; g1_server_set_budget raising it
    admits them without any fetch, and every result is right."""
    import graphone_amd as g1
    h, g = graphs[0]
    graph_bytes = g.device_bytes()
    srv = g1.Server(hbm_budget_bytes=graph_bytes + h.n * 4 + 64, batch_min_sssp=1000, batch_min_ppr=1000)
    gid = srv.add_graph(g)
    qids = [srv.submit(gid, "sssp", s) for s in range(6)]
    assert srv.wait(qids[0])["state_name"] == "done"
    assert srv.poll(qids[1])["state_name"] == "queued"
    srv.set_budget(graph_bytes + 16 * h.n * 4)
    for i, q in enumerate(qids):
        assert srv.wait(q, 60000)["state_name"] == "done"
        assert np.array_equal(srv.fetch(q, "sssp"), _want(h, "sssp", i))
    srv.close()


def test_remove_graph(graphs):
    """Removing a graph waits for its queries, keeps their results fetchable, returns its bytes to the
    budget and refuses new queries on it; the other graph keeps serving."""
    import graphone_amd as g1
    (h0, g0), (h1, g1_) = graphs
    srv = g1.Server(batch_min_sssp=1000, batch_min_ppr=1000)
    bytes_a = g0.device_bytes()  # what add_graph reserves for it
    a = srv.add_graph(g0)
    b = srv.add_graph(g1_)
    used = srv.stats()["hbm_used"]
    qa = [srv.submit(a, "bfs", s) for s in range(5)]
    srv.remove_graph(a)
    for i, q in enumerate(qa):
        assert srv.poll(q)["state_name"] == "done"
        assert np.array_equal(srv.fetch(q, "bfs"), _want(h0, "bfs", i))
    assert srv.stats()["hbm_used"] == used - bytes_a
    with pytest.raises(g1.G1Error) as e:
        srv.submit(a, "bfs", 0)
    assert e.value.status == g1._lib.G1_ERR_NOT_FOUND
    with pytest.raises(g1.G1Error):
        srv.remove_graph(a)
    q = srv.submit(b, "sssp", 3)
    assert srv.wait(q)["state_name"] == "done"
    assert np.array_equal(srv.fetch(q, "sssp"), _want(h1, "sssp", 3))
    srv.close()


def test_cancel_and_errors(graphs):
    import graphone_amd as g1
    h, g = graphs[0]
    srv = g1.Server(hbm_budget_bytes=g.device_bytes() + h.n * 4 + 64)
    gid = srv.add_graph(g)
    a = srv.submit(gid, "bfs", 1)
    b = srv.submit(gid, "bfs", 2)  # waits: one result fits
    srv.cancel(b)
    assert srv.poll(b)["state_name"] == "cancelled"
    assert srv.wait(a)["state_name"] == "done"
    with pytest.raises(g1.G1Error):
        srv.fetch(b)
    srv.fetch(a)
    with pytest.raises(g1.G1Error):
        srv.fetch(a)  # already fetched
    with pytest.raises(g1.G1Error):
        srv.submit(gid, "sssp", h.n)
    with pytest.raises(g1.G1Error):
        srv.submit(7, "bfs", 0)
    srv.close()


def test_fresh_server_on_recycled_memory(graphs):
    """Device memory that held earlier data (here all ones) is handed to the next server's graph views:
    their work areas must be cleared on the stream that uses them, so every single-source thread's first
    query is right. Regression test for flags cleared on the null stream, which is not ordered before a
    non-blocking stream's kernels."""
    import graphone_amd as g1
    junk = g1.DeviceArray(64 << 20, np.uint32)
    junk.fill(0xFFFFFFFF)
    g1.synchronize()
    junk.free()
    srv = g1.Server(batch_k=8, batch_min_sssp=64, single_workers=4)
    ids = [srv.add_graph(g) for _, g in graphs]
    qs = [(gid, kind, 3 + 7 * i) for i, (gid, kind) in enumerate([(0, "sssp"), (0, "bfs")] * 6 + [(1, "sssp")] * 4)]
    qids = [srv.submit(ids[gid], kind, src) for gid, kind, src in qs]
    for (gid, kind, src), qid in zip(qs, qids):
        st = srv.wait(qid, 60000)
        assert st["state_name"] == "done", (gid, kind, src, st)
        _check(graphs, gid, kind, src, srv.fetch(qid, kind))
    srv.close()


def test_sssp_session_at_width_64(graphs):
    """A backlog of 40 SSSP queries opens a 64-column session (the k = 64 wavefront kernel, 32-bit
    distances, continuous batching); every served column equals the oracle's Dijkstra."""
    import graphone_amd as g1
    h, g = graphs[0]
    srv = g1.Server(batch_k=64, batch_min_sssp=2)
    gid = srv.add_graph(g)
    first = srv.submit(gid, "sssp", 1)  # a single-source run first: the graph's iteration count is known
    assert srv.wait(first, 60000)["state_name"] == "done"
    _check(graphs, 0, "sssp", 1, srv.fetch(first, "sssp"))
    srcs = [int(s) for s in np.random.default_rng(8).integers(0, h.n, size=40)]
    qids = [srv.submit(gid, "sssp", s) for s in srcs]
    batched = 0
    for s, q in zip(srcs, qids):
        st = srv.wait(q, 60000)
        assert st["state_name"] == "done", st
        batched += st["path"] == 1
        _check(graphs, 0, "sssp", s, srv.fetch(q, "sssp"))
    assert batched > 0
    srv.close()
