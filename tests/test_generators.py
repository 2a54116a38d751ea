"""Host generators and CSR construction (no GPU): the documented random streams, the CSR invariants of
the reference loader (include/GraphOne/loader/graph_loader.h:254-296, edge_cache.h:50-105), weights and
the transpose.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import graphone_amd as g1

M64 = (1 << 64) - 1


def mix64(z):
    z ^= z >> 30
    z = (z * 0xBF58476D1CE4E5B9) & M64
    z ^= z >> 27
    z = (z * 0x94D049BB133111EB) & M64
    z ^= z >> 31
    return z


def unit(r):
    return (r >> 11) * (1.0 / 9007199254740992.0)


def rmat_python(scale, ef, a, b, c, seed):
    """Pure-Python restatement of the R-MAT stream documented in graphone_amd/csrc/gen.cpp."""
    edges = set()
    for e in range((ef << scale)):
        state = mix64(seed ^ mix64((e + 0x632BE59BD9B4E019) & M64))
        s = d = 0
        for _ in range(scale):
            state = (state + 0x9E3779B97F4A7C15) & M64
            u = unit(mix64(state))
            if u < a:
                sb, db = 0, 0
            elif u < a + b:
                sb, db = 0, 1
            elif u < a + b + c:
                sb, db = 1, 0
            else:
                sb, db = 1, 1
            s, d = (s << 1) | sb, (d << 1) | db
        if s != d:
            edges.add((s, d))
    return sorted(edges)


def check_csr(h, allow_self=False):
    rp, ci = h.row_offsets.astype(np.int64), h.col_indices.astype(np.int64)
    assert rp[0] == 0 and rp[-1] == h.m and np.all(np.diff(rp) >= 0)
    assert np.all(ci < h.n)
    for v in range(h.n):
        row = ci[rp[v]:rp[v + 1]]
        assert np.all(np.diff(row) > 0), f"row {v} not strictly ascending"
        if not allow_self:
            assert not np.any(row == v)


@pytest.mark.parametrize("scale,ef,seed", [(4, 4, 1), (6, 8, 99), (7, 16, 12345)])
def test_rmat_matches_documented_stream(scale, ef, seed):
    h = g1.HostGraph.rmat(scale, ef, 0.57, 0.19, 0.19, seed)
    want = rmat_python(scale, ef, 0.57, 0.19, 0.19, seed)
    rp, ci = h.row_offsets, h.col_indices
    got = [(u, int(ci[k])) for u in range(h.n) for k in range(rp[u], rp[u + 1])]
    assert got == want
    assert h.n == 1 << scale
    check_csr(h)


def test_rmat_undirected_is_symmetric():
    h = g1.HostGraph.rmat(9, 8, seed=3, undirected=True)
    check_csr(h)
    m = sp.csr_matrix((np.ones(h.m), h.col_indices, h.row_offsets), shape=(h.n, h.n))
    assert (m != m.T).nnz == 0


def test_rmat_scale_16_properties():
    h = g1.HostGraph.rmat(16, 16, seed=1)
    check_csr(h)
    deg = h.out_degrees()
    assert 0.9 * (16 << 16) < h.m <= (16 << 16)
    assert deg.max() > 50 * deg.mean()  # power-law skew
    assert int(np.argmax(deg)) < 64  # a=0.57 puts the hubs at small ids


def test_grid_structure_and_keep_rate():
    h = g1.HostGraph.grid(50, 40, 0.6, seed=9)
    check_csr(h)
    rp, ci = h.row_offsets, h.col_indices
    for v in range(h.n):
        r, c = divmod(v, 40)
        for u in ci[rp[v]:rp[v + 1]]:
            ru, cu = divmod(int(u), 40)
            assert abs(ru - r) + abs(cu - c) == 1
    m = sp.csr_matrix((np.ones(h.m), ci, rp), shape=(h.n, h.n))
    assert (m != m.T).nnz == 0
    full = 2 * (50 * 39 + 49 * 40)
    assert abs(h.m / full - 0.6) < 0.03
    # p = 1 keeps every grid edge
    assert g1.HostGraph.grid(5, 7, 1.0, seed=1).m == 2 * (5 * 6 + 4 * 7)


def test_weights_range_and_symmetry():
    h = g1.HostGraph.grid(30, 30, 0.6, seed=1).assign_weights(1, 64, seed=2, symmetric=True)
    w = h.weights
    assert w.min() >= 1 and w.max() <= 64 and len(np.unique(w)) == 64
    m = sp.csr_matrix((w.astype(np.float64), h.col_indices, h.row_offsets), shape=(h.n, h.n))
    assert abs(m - m.T).max() == 0
    r = g1.HostGraph.rmat(10, 8, seed=4).assign_weights(1, 255, seed=5)
    assert r.weights.min() >= 1 and r.weights.max() <= 255
    again = g1.HostGraph.rmat(10, 8, seed=4).assign_weights(1, 255, seed=5)
    assert np.array_equal(r.weights, again.weights)


def test_transpose_matches_scipy():
    h = g1.HostGraph.rmat(11, 8, seed=6).assign_weights(1, 255, seed=7)
    t = h.transpose()
    check_csr(t)
    m = sp.csr_matrix((h.weights.astype(np.int64), h.col_indices, h.row_offsets), shape=(h.n, h.n))
    mt = m.T.tocsr()
    mt.sort_indices()
    assert np.array_equal(t.row_offsets, mt.indptr)
    assert np.array_equal(t.col_indices, mt.indices)
    assert np.array_equal(t.weights, mt.data)


def test_csr_from_edges_loader_rules():
    src = [0, 1, 1, 2, 2, 3]
    dst = [1, 0, 1, 3, 3, 0]
    w = [5, 6, 7, 8, 9, 10]
    # undirected (default of LoadEdgeOpts, opts.hpp:16): reverse edges added, self loop dropped,
    # duplicates dropped with the first occurrence keeping its weight
    h = g1.HostGraph.from_edges(4, src, dst, w, directed=False)
    check_csr(h)
    rows = {v: list(zip(h.col_indices[h.row_offsets[v]:h.row_offsets[v + 1]].tolist(),
                        h.weights[h.row_offsets[v]:h.row_offsets[v + 1]].tolist())) for v in range(4)}
    assert rows == {0: [(1, 5), (3, 10)], 1: [(0, 5)], 2: [(3, 8)], 3: [(0, 10), (2, 8)]}
    hd = g1.HostGraph.from_edges(4, src, dst, w, directed=True, keep_self=True, keep_dup=True)
    assert hd.m == 6
    with pytest.raises(g1.G1Error):
        g1.HostGraph.from_edges(2, [0], [5])
