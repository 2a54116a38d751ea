"""Step-level C-ABI on the GPU: an SSSP and a BFS loop driven one advance step at a time
(g1_advance_sssp_u32 / g1_advance_bfs, the reference's AdvanceGC::Forward, engine/advance_gc.h:180-244),
a BLAS-engine SSSP loop (SpMV + g1_apply_minplus_u32, engine/blas_gc.h:171-351), the frontier conversions,
the PageRank apply and L1 norm, and the semiring SpMM with caller-owned operands -- each against the oracle
or a numpy statement of the same step."""
import ctypes as C

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

INF = 0xFFFFFFFF


def _g1():
    import graphone_amd as g1
    return g1


def _dev(a):
    return _g1().DeviceArray.from_numpy(np.ascontiguousarray(a))


def _p(d):
    return None if d is None else C.c_void_p(d.ptr)


def _call(name, *args):
    _g1()._lib.call(name, *args)


def _ws(n, cap):
    g1 = _g1()
    nbytes = g1._lib.lib().g1_step_workspace_bytes(n, cap)
    return g1.DeviceArray((nbytes + 3) // 4, np.uint32), nbytes


def _u32(d):
    return int(d.numpy()[0])


def _graphs():
    g1 = _g1()
    yield "rmat14", g1.HostGraph.rmat(14, 16, seed=11).assign_weights(1, 255, seed=12)
    yield "grid48", g1.HostGraph.grid(48, 48, 0.6, seed=3).assign_weights(1, 64, seed=4, symmetric=True)
    yield "rmat10u", g1.HostGraph.rmat(10, 4, seed=5, undirected=True).assign_weights(1, 9, seed=6, symmetric=True)


def _csr_dev(h):
    return _dev(h.row_offsets), _dev(np.concatenate([h.col_indices, np.zeros(1, np.uint32)])), \
        _dev(np.concatenate([h.weights, np.zeros(1, np.uint32)]))


def _advance_loop(h, src, sssp):
    n = h.n
    rp, ci, w = _csr_dev(h)
    dist = _dev(np.where(np.arange(n) == src, 0, INF).astype(np.uint32))
    visited = _dev(np.zeros(n, np.uint32))
    q = [_dev(np.full(n, INF, np.uint32)), _dev(np.full(n, INF, np.uint32))]
    q[0].copy_from(np.concatenate([[src], np.full(n - 1, INF)]).astype(np.uint32))
    cnt = [_dev(np.array([1], np.uint32)), _dev(np.array([0], np.uint32))]
    ws, nbytes = _ws(n, n)
    it = 0
    sizes = []
    while True:
        a, b = it % 2, (it + 1) % 2
        if sssp:
            _call("g1_advance_sssp_u32", n, _p(rp), _p(ci), _p(w), _p(q[a]), _p(cnt[a]), n, _p(dist), _p(visited),
                  it + 1, _p(q[b]), _p(cnt[b]), _p(ws), C.c_size_t(nbytes), None)
        else:
            _call("g1_advance_bfs", n, _p(rp), _p(ci), _p(q[a]), _p(cnt[a]), n, _p(dist), it + 1, _p(q[b]),
                  _p(cnt[b]), _p(ws), C.c_size_t(nbytes), None)
        it += 1
        c = _u32(cnt[b])
        sizes.append(c)
        if c == 0:
            break
        if not sssp:
            # a BFS level's frontier is exactly the vertices at that depth, each once
            got = np.sort(q[b].numpy()[:c])
            d = dist.numpy()
            assert np.array_equal(got, np.flatnonzero(d == it).astype(np.uint32)), it
        assert it < 10 * n
    return dist.numpy(), sizes


def test_sssp_loop_through_advance_steps(gpu):
    for name, h in _graphs():
        deg = h.out_degrees()
        for src in (int(np.argmax(deg)), 1 % h.n):
            got, _ = _advance_loop(h, src, True)
            want = oracle.dijkstra(h.row_offsets, h.col_indices, h.weights, src)
            assert np.array_equal(got, want), (name, src)


def test_bfs_loop_through_advance_steps(gpu):
    for name, h in _graphs():
        src = int(np.argmax(h.out_degrees()))
        got, sizes = _advance_loop(h, src, False)
        want = oracle.bfs_queue(h.row_offsets, h.col_indices, src)
        assert np.array_equal(got, want), name


def test_blas_sssp_loop_through_spmv_and_apply(gpu):
    """The sparse BLAS engine's loop (engine/blas_gc.h:171-247 with apps/sssp_blas.cu's functor): y = A^T (x) x
    by the merge-path SpMV, then g1_apply_minplus_u32, until no vertex changes; the iteration count equals
    the oracle's restated loop."""
    g1 = _g1()
    for name, h in _graphs():
        n = h.n
        src = int(np.argmax(h.out_degrees()))
        t = h.transpose()
        trp, tci = _dev(t.row_offsets), _dev(np.concatenate([t.col_indices, np.zeros(1, np.uint32)]))
        tw = _dev(np.concatenate([t.weights, np.zeros(1, np.uint32)]))
        d = _dev(np.where(np.arange(n) == src, 0, INF).astype(np.uint32))
        x = _dev(np.where(np.arange(n) == src, 0, INF).astype(np.uint32))
        y = g1.DeviceArray(n, np.uint32)
        changed = g1.DeviceArray(1, np.uint32)
        sws = g1.merge_workspace(n, t.m, trp)
        ws, nbytes = _ws(n, 0)
        it = 0
        while True:
            g1.spmv("merge", "minplus_u32", trp, tci, tw, x, y, n, n, t.m, sws)
            _call("g1_apply_minplus_u32", n, _p(y), _p(d), _p(x), _p(changed), _p(ws), C.c_size_t(nbytes), None)
            it += 1
            if _u32(changed) == 0:
                break
        want, st = oracle.minplus_blas(t.row_offsets, t.col_indices, t.weights, src)
        assert np.array_equal(d.numpy(), want), name
        assert it == st["iterations"], (name, it, st)


def test_frontier_conversions_and_density(gpu):
    g1 = _g1()
    rng = np.random.default_rng(5)
    h = g1.HostGraph.rmat(12, 8, seed=9)
    n = h.n
    rp = _dev(h.row_offsets)
    deg = h.out_degrees().astype(np.int64)
    for cap, count in ((1, 1), (1000, 700), (5000, 5000), (n, 17)):
        qv = rng.integers(0, n, cap).astype(np.uint32)
        qv[rng.random(cap) < 0.2] = INF
        q = _dev(qv)
        cnt = _dev(np.array([count], np.uint32))
        ws, nbytes = _ws(n, cap)
        edges = g1.DeviceArray(1, np.uint32)
        _call("g1_frontier_density", n, _p(rp), _p(q), _p(cnt), cap, _p(edges), _p(ws), C.c_size_t(nbytes), None)
        valid = qv[:count][qv[:count] != INF]
        assert _u32(edges) == int(deg[valid].sum()), (cap, count)
        bm = g1.DeviceArray((n + 31) // 32, np.uint32)
        bm.fill(0xFFFFFFFF)  # cleared by the call
        _call("g1_frontier_queue_to_bitmap", n, _p(q), _p(cnt), cap, _p(bm), None)
        bits = np.unpackbits(bm.numpy().view(np.uint8), bitorder="little")[:n]
        assert np.array_equal(np.flatnonzero(bits), np.unique(valid)), (cap, count)
        out = g1.DeviceArray(n, np.uint32)
        ocnt = g1.DeviceArray(1, np.uint32)
        _call("g1_frontier_bitmap_to_queue", n, _p(bm), _p(out), _p(ocnt), _p(ws), C.c_size_t(nbytes), None)
        c = _u32(ocnt)
        assert np.array_equal(out.numpy()[:c], np.unique(valid).astype(np.uint32))
        _call("g1_filter_compact", _p(q), _p(cnt), cap, _p(out), _p(ocnt), _p(ws), C.c_size_t(nbytes), None)
        c = _u32(ocnt)
        assert np.array_equal(out.numpy()[:c], valid), (cap, count)


def test_step_rejects_small_workspace(gpu):
    g1 = _g1()
    ws = g1.DeviceArray(4, np.uint32)
    d = g1.DeviceArray(8, np.uint32)
    with pytest.raises(g1.G1Error):
        _call("g1_apply_minplus_u32", 100000, _p(d), _p(d), _p(d), _p(d), _p(ws), C.c_size_t(16), None)


def test_apply_minplus_and_pagerank_apply(gpu):
    g1 = _g1()
    rng = np.random.default_rng(8)
    for n in (1, 300, 100003):
        y = rng.integers(0, 1000, n).astype(np.uint32)
        y[rng.random(n) < 0.3] = INF
        d0 = rng.integers(0, 1200, n).astype(np.uint32)
        d0[rng.random(n) < 0.3] = INF
        dy, dd, dx = _dev(y), _dev(d0), g1.DeviceArray(n, np.uint32)
        ch = g1.DeviceArray(1, np.uint32)
        ws, nbytes = _ws(n, 0)
        _call("g1_apply_minplus_u32", n, _p(dy), _p(dd), _p(dx), _p(ch), _p(ws), C.c_size_t(nbytes), None)
        pred = (y != INF) & (y < d0)
        assert np.array_equal(dd.numpy(), np.where(pred, y, d0))
        assert np.array_equal(dx.numpy(), np.where(pred, y, INF).astype(np.uint32))
        assert _u32(ch) == int(pred.sum())
        # PageRank apply: float32 arithmetic, same bits as numpy's float32 add and multiply
        yf = rng.random(n, dtype=np.float32) * 1e-3
        iw = rng.random(n, dtype=np.float32)
        ro = rng.random(n, dtype=np.float32) * 1e-3
        base = np.float32(0.15 / max(n, 1))
        dyf, diw, dro = _dev(yf), _dev(iw), _dev(ro)
        rn, xf = g1.DeviceArray(n, np.float32), g1.DeviceArray(n, np.float32)
        l1 = g1.DeviceArray(1, np.float64)
        _call("g1_pr_apply_f32", n, _p(dyf), C.c_float(base), _p(diw), _p(dro), _p(rn), _p(xf), _p(l1), _p(ws),
              C.c_size_t(nbytes), None)
        want_r = (base + yf).astype(np.float32)
        assert np.array_equal(rn.numpy().view(np.uint32), want_r.view(np.uint32))
        assert np.array_equal(xf.numpy().view(np.uint32), (want_r * iw).astype(np.float32).view(np.uint32))
        want_l1 = float(np.abs(want_r.astype(np.float64) - ro.astype(np.float64)).sum())
        assert abs(float(l1.numpy()[0]) - want_l1) <= 1e-12 * max(want_l1, 1.0)
        l1b = g1.DeviceArray(1, np.float64)
        _call("g1_l1_delta_f64", n, _p(rn), _p(dro), _p(l1b), _p(ws), C.c_size_t(nbytes), None)
        assert l1b.numpy()[0] == l1.numpy()[0]  # same fixed order, same bits


def test_spmm_hub_row_spreads_over_tiles(gpu):
    """One row with 200,000 nonzeros (a hub) beside short rows: the row spans ~100 merge tiles, so its
    columns are folded across tiles by the fixup; min-plus exact, plus-times within 2e-6 of FP64."""
    g1 = _g1()
    rng = np.random.default_rng(17)
    n_cols, k = 250_000, 64
    degs = np.concatenate([[200_000], rng.integers(0, 5, 999)])
    rp = np.zeros(degs.size + 1, np.int64)
    rp[1:] = np.cumsum(degs)
    ci = np.concatenate([np.sort(rng.choice(n_cols, d, replace=False)) for d in degs]).astype(np.uint32)
    rp = rp.astype(np.uint32)
    n_rows, m = degs.size, int(rp[-1])
    w = rng.integers(1, 50, m).astype(np.uint32)
    X = rng.integers(0, 100_000, (n_cols, k)).astype(np.uint32)
    drp, dci, dw, dX = _dev(rp), _dev(np.concatenate([ci, np.zeros(1, np.uint32)])), _dev(w), _dev(X)
    dY = g1.DeviceArray(n_rows * k, np.uint32)
    wsb = g1._lib.lib().g1_spmm_workspace_bytes(n_rows, m)
    ws = g1.DeviceArray((wsb + 3) // 4, np.uint32)
    _call("g1_spmm_minplus_u32", n_rows, n_cols, m, k, _p(drp), _p(dci), _p(dw), _p(dX), _p(dY), None, _p(ws),
          C.c_size_t(wsb), None)
    Y = dY.numpy().reshape(n_rows, k)
    for j in (0, 31, 63):
        assert np.array_equal(Y[:, j], oracle.spmv_minplus_u32(rp, ci, w, X[:, j])), j
    Xf = rng.random((n_cols, k), dtype=np.float32)
    dXf, dYf, dwf = _dev(Xf), g1.DeviceArray(n_rows * k, np.float32), _dev(w.astype(np.float32))
    _call("g1_spmm_plustimes_f32", n_rows, n_cols, m, k, _p(drp), _p(dci), _p(dwf), _p(dXf), _p(dYf), None,
          _p(ws), C.c_size_t(wsb), None)
    Yf = dYf.numpy().reshape(n_rows, k).astype(np.float64)
    for j in (0, 63):
        want = oracle.spmv_plustimes_f64(rp, ci, w.astype(np.float64), Xf[:, j].astype(np.float64))
        assert np.allclose(Yf[:, j], want, rtol=2e-6, atol=1e-30), j


@pytest.mark.parametrize("k", [1, 7, 64, 100])
def test_spmm_minplus_and_plustimes(gpu, k):
    g1 = _g1()
    rng = np.random.default_rng(k)
    h = g1.HostGraph.rmat(11, 8, seed=k).assign_weights(1, 99, seed=k + 1)
    t = h.transpose()
    n, m = t.n, t.m
    trp, tci = _dev(t.row_offsets), _dev(np.concatenate([t.col_indices, np.zeros(1, np.uint32)]))
    tw = _dev(np.concatenate([t.weights, np.zeros(1, np.uint32)]))
    X = rng.integers(0, 5000, (n, k)).astype(np.uint32)
    X[rng.random((n, k)) < 0.4] = INF
    mask = np.zeros((k + 63) // 64, np.uint64)
    on = rng.random(k) < 0.7
    on[0] = True
    for j in np.flatnonzero(on):
        mask[j // 64] |= np.uint64(1) << np.uint64(j % 64)
    Y0 = rng.integers(0, 9, (n, k)).astype(np.uint32)
    dX, dY, dm = _dev(X), _dev(Y0), _dev(mask)
    wsb = g1._lib.lib().g1_spmm_workspace_bytes(n, m)
    ws = g1.DeviceArray((wsb + 3) // 4, np.uint32)
    _call("g1_spmm_minplus_u32", n, n, m, k, _p(trp), _p(tci), _p(tw), _p(dX), _p(dY), _p(dm), _p(ws),
          C.c_size_t(wsb), None)
    Y = dY.numpy().reshape(n, k)
    for j in range(k):
        want = oracle.spmv_minplus_u32(t.row_offsets, t.col_indices, t.weights, X[:, j]) if on[j] else Y0[:, j]
        assert np.array_equal(Y[:, j], want), j
    # plus-times, every column, against an FP64 product
    Xf = rng.random((n, k), dtype=np.float32)
    wf = t.weights.astype(np.float32)
    dXf, dYf, dwf = _dev(Xf), g1.DeviceArray(n * k, np.float32), _dev(np.concatenate([wf, np.zeros(1, np.float32)]))
    _call("g1_spmm_plustimes_f32", n, n, m, k, _p(trp), _p(tci), _p(dwf), _p(dXf), _p(dYf), None, _p(ws),
          C.c_size_t(wsb), None)
    Yf = dYf.numpy().reshape(n, k).astype(np.float64)
    for j in range(0, k, max(1, k // 8)):
        want = oracle.spmv_plustimes_f64(t.row_offsets, t.col_indices, wf.astype(np.float64), Xf[:, j].astype(np.float64))
        assert np.allclose(Yf[:, j], want, rtol=2e-6, atol=1e-30), j
