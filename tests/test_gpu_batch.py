"""Batched multi-source engines on the GPU: every column of a batched run equals the single-source run of
its source, bit for bit, at every column width, and the single-source runs equal the oracle.

Min-plus columns (SSSP distances, BFS depths) are compared with the oracle's Dijkstra / BFS queue and
with g1_sssp's BLAS engine (apps/sssp_blas.cu:17-84); personalized PageRank columns with g1_ppr
(apps/pagerank_blas.cu:16-125 with a restart vector) bit for bit, and with the FP64 oracle within L1 1e-5.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

WIDTHS = [8, 16, 32, 64]


@pytest.fixture(scope="module")
def graphs(gpu):
    import graphone_amd as g1
    out = []
    h = g1.HostGraph.rmat(13, 16, seed=21).assign_weights(1, 255, seed=22)
    out.append(("rmat13", h, g1.Graph(h)))
    h = g1.HostGraph.grid(48, 48, 0.6, seed=5).assign_weights(1, 64, seed=6, symmetric=True)
    out.append(("grid48", h, g1.Graph(h, symmetric=True)))
    return out


def _sources(h, count, seed):
    rng = np.random.default_rng(seed)
    deg = h.out_degrees()
    src = list(rng.integers(0, h.n, size=count - 3))
    src.append(int(np.argmax(deg)))                       # hub
    src.append(int(np.flatnonzero(deg == 0)[0]) if np.any(deg == 0) else 0)  # leaf / isolated
    src.append(src[0])                                      # duplicate source
    return [int(s) for s in src]


@pytest.mark.parametrize("k", WIDTHS)
def test_sssp_batch_columns_bitexact(graphs, k):
    for name, h, g in graphs:
        srcs = _sources(h, 19, seed=k)
        got = g.sssp_batch(srcs, k=k)
        assert got.values.shape == (len(srcs), h.n)
        assert got.stats["k"] == k and got.stats["tiles"] == -(-len(srcs) // k)
        for c, s in enumerate(srcs):
            want = oracle.dijkstra(h.row_offsets, h.col_indices, h.weights, s)
            assert np.array_equal(got.values[c], want), (name, k, c, s)


@pytest.mark.parametrize("k", [8, 64])
def test_bfs_batch_columns_bitexact(graphs, k):
    for name, h, g in graphs:
        srcs = _sources(h, 11, seed=100 + k)
        got = g.sssp_batch(srcs, k=k, weighted=False)
        for c, s in enumerate(srcs):
            want = oracle.bfs_queue(h.row_offsets, h.col_indices, s)
            assert np.array_equal(got.values[c], want), (name, k, c, s)


@pytest.mark.parametrize("tile", [512, 1024, 2048, 4096])
def test_minplus_merge_tiles_bitexact(graphs, tile):
    """The k = 64 min-plus kernel on each merge partition the autotuner may pick (rows + nonzeros per
    tile): SSSP and BFS columns equal the oracle's, and so does the 32-bit rerun of a batch whose
    distances outgrow 16 bits (a 400-vertex path: 3 tiles of 512, 1 of 4096)."""
    import graphone_amd as g1
    try:
        for name, h, g in graphs:
            g.set_merge_tile(tile)
            srcs = _sources(h, 64, seed=tile)
            got = g.sssp_batch(srcs, k=64)
            assert got.stats["merge_tile"] == tile
            bfs = g.sssp_batch(srcs, k=64, weighted=False)
            for c, s in enumerate(srcs):
                want = oracle.dijkstra(h.row_offsets, h.col_indices, h.weights, s)
                assert np.array_equal(got.values[c], want), (name, tile, c, s)
                want = oracle.bfs_queue(h.row_offsets, h.col_indices, s)
                assert np.array_equal(bfs.values[c], want), (name, tile, c, s)
    finally:
        for _, _, g in graphs:
            g.set_merge_tile(2048)
    n = 400
    src = np.arange(n - 1, dtype=np.uint32)
    h = g1.HostGraph.from_edges(n, src, src + 1, np.full(n - 1, 250, dtype=np.uint32), directed=False)
    g = g1.Graph(h)
    g.set_merge_tile(tile)
    srcs = [int(s) for s in np.linspace(0, n - 1, 64).astype(np.uint32)]
    got = g.sssp_batch(srcs, k=64)
    assert int(got.values.max()) > 0xFFFF
    for c, s in enumerate(srcs):
        assert np.array_equal(got.values[c], oracle.dijkstra(h.row_offsets, h.col_indices, h.weights, s)), (tile, c)
    with pytest.raises(g1.G1Error):
        g.set_merge_tile(3000)


def test_sssp_batch_equals_single_source_engine(graphs):
    name, h, g = graphs[0]
    srcs = _sources(h, 7, seed=3)
    got = g.sssp_batch(srcs, k=16)
    for c, s in enumerate(srcs):
        assert np.array_equal(got.values[c], g.sssp(s, "merge").values)


@pytest.mark.parametrize("k", WIDTHS)
def test_ppr_batch_columns_bit_identical(graphs, k):
    """A batched PPR column is the single-seed g1_ppr result bit for bit (same row-sum fold order, same
    per-column L1 tree and stop rule), including seeds that stop at different iterations."""
    for name, h, g in graphs:
        seeds = _sources(h, 13, seed=200 + k)
        got = g.ppr_batch(seeds, restart=0.15, max_iter=60, tol=1e-7, k=k)
        iters = set()
        for c, s in enumerate(seeds):
            single = g.ppr(s, restart=0.15, max_iter=60, tol=1e-7)
            iters.add(single.stats["iterations"])
            assert np.array_equal(got.values[c].view(np.uint32), single.values.view(np.uint32)), (name, k, c, s)
        if name == "rmat13":
            assert len(iters) > 1, "seeds should stop at different iterations"


def test_ppr_batch_matches_oracle(graphs):
    for name, h, g in graphs:
        t = h.transpose()
        seeds = _sources(h, 5, seed=9)
        got = g.ppr_batch(seeds, restart=0.15, max_iter=100, tol=1e-9, k=8)
        for c, s in enumerate(seeds):
            want, _ = oracle.pagerank_f64(h.row_offsets, h.weights, t.row_offsets, t.col_indices, t.weights,
                                          0.85, max_iter=100, tol=1e-9, seed=s, restart=0.15)
            l1 = float(np.abs(got.values[c].astype(np.float64) - want).sum())
            assert l1 <= 1e-5, (name, c, s, l1)


def test_autotune_and_default_width(graphs):
    name, h, g = graphs[0]
    srcs = _sources(h, 9, seed=1)
    k = g.autotune("minplus", srcs, weighted=True)
    assert k in WIDTHS
    got = g.sssp_batch(srcs)  # k = 0: the tuned width
    assert got.stats["k"] == k
    assert got.stats["merge_tile"] in (512, 1024, 2048, 4096)
    for c, s in enumerate(srcs):
        assert np.array_equal(got.values[c], oracle.dijkstra(h.row_offsets, h.col_indices, h.weights, s))
    kp = g.autotune("ppr", srcs)
    assert kp in WIDTHS
    assert g.ppr_batch(srcs[:3]).stats["k"] == kp


def test_batch_edge_cases(graphs):
    import graphone_amd as g1
    name, h, g = graphs[0]
    # one source, and an empty batch
    one = g.sssp_batch([5], k=64)
    assert np.array_equal(one.values[0], oracle.dijkstra(h.row_offsets, h.col_indices, h.weights, 5))
    empty = g.sssp_batch([], k=8)
    assert empty.values.shape == (0, h.n) and empty.stats["tiles"] == 0
    with pytest.raises(g1.G1Error):
        g.sssp_batch([h.n], k=8)
    with pytest.raises(g1.G1Error):
        g.sssp_batch([0, 1], k=12)
    with pytest.raises(g1.G1Error):
        g.ppr_batch([0], restart=0.0, k=8)


def test_road_batch_matches_oracle(graphs):
    """One workgroup per query on a low-degree graph (the grid): mixed BFS / SSSP queries, bit-exact
    against the oracle, including more queries than one launch takes."""
    import graphone_amd as g1
    name, h, g = graphs[1]
    assert name == "grid48"
    rng = np.random.default_rng(5)
    srcs = [int(s) for s in rng.integers(0, h.n, size=70)]
    kinds = ["sssp" if i % 3 else "bfs" for i in range(len(srcs))]
    got = g.road_batch(srcs, kinds)
    for c, (s, k) in enumerate(zip(srcs, kinds)):
        if k == "sssp":
            want = oracle.dijkstra(h.row_offsets, h.col_indices, h.weights, s)
        else:
            want = oracle.bfs_queue(h.row_offsets, h.col_indices, s)
        assert np.array_equal(got.values[c], want), (c, s, k)
    with pytest.raises(g1.G1Error):
        graphs[0][2].road_batch([0])  # RMAT: hubs far above the road engine's degree bound


def test_sssp_batch_distances_past_16_bits(gpu):
    """k = 64 keeps distances in 16 bits while they fit; on a long heavy path (distances up to
    ~100,000 > 0xFFFF) the batch must notice and redo the run with 32-bit storage, still bit-exact.
    A second, light path checks the 16-bit run itself (all distances < 0xFFFF)."""
    import graphone_amd as g1
    n = 400
    src = np.arange(n - 1, dtype=np.uint32)
    dst = src + 1
    for wv in (250, 3):
        h = g1.HostGraph.from_edges(n, src, dst, np.full(n - 1, wv, dtype=np.uint32), directed=False)
        g = g1.Graph(h)
        srcs = [int(s) for s in np.linspace(0, n - 1, 64).astype(np.uint32)]
        got = g.sssp_batch(srcs, k=64)
        for c, s in enumerate(srcs):
            want = oracle.dijkstra(h.row_offsets, h.col_indices, h.weights, s)
            assert np.array_equal(got.values[c], want), (wv, c, s)
        if wv == 250:
            assert int(got.values.max()) > 0xFFFF
