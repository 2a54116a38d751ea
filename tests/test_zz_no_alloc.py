"""g1_spmv_merge_minplus_u32_noalloc (the merge-path min-plus SpMV on the library's static scratch, no
allocation) against g1_spmv_merge_minplus_u32 with a caller workspace, bit for bit, on two calls in a row:
a larger matrix, then a smaller one (the second call runs on the scratch the first left behind)."""
import ctypes as C

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

INF = 0xFFFFFFFF


def _csr(rng, n_rows, n_cols, avg, hub_rows=(), hub_deg=0):
    degs = rng.poisson(avg, n_rows).astype(np.int64)
    degs[rng.random(n_rows) < 0.3] = 0
    for h in hub_rows:
        degs[h] = hub_deg
    degs = np.minimum(degs, n_cols)
    rp = np.zeros(n_rows + 1, dtype=np.int64)
    rp[1:] = np.cumsum(degs)
    cols = np.empty(rp[-1], dtype=np.uint32)
    for r in np.flatnonzero(degs):
        cols[rp[r]:rp[r + 1]] = np.sort(rng.choice(n_cols, degs[r], replace=False))
    return rp.astype(np.uint32), cols


def _both(g1, rp, cols, vals, x, n_cols):
    dev = g1.DeviceArray.from_numpy
    n_rows, nnz = rp.size - 1, cols.size
    d_rp, d_ci, d_x = dev(rp), dev(cols if nnz else np.zeros(1, np.uint32)), dev(x)
    d_v = None if vals is None else dev(vals)
    y_ws = g1.DeviceArray(n_rows, np.uint32)
    y_na = g1.DeviceArray(n_rows, np.uint32)
    y_na.fill(0x5A5A5A5A)
    ws = g1.merge_workspace(n_rows, nnz, d_rp)
    g1.spmv("merge", "minplus_u32", d_rp, d_ci, d_v, d_x, y_ws, n_rows, n_cols, nnz, ws)
    vp = None if d_v is None else C.c_void_p(d_v.ptr)
    g1._lib.call("g1_spmv_merge_minplus_u32_noalloc", C.c_uint32(n_rows), C.c_uint32(n_cols), C.c_uint32(nnz),
                 C.c_void_p(d_rp.ptr), C.c_void_p(d_ci.ptr), vp, C.c_void_p(d_x.ptr), C.c_void_p(y_na.ptr), None)
    g1.synchronize()
    return y_ws.numpy(), y_na.numpy()


def test_noalloc_matches_workspace_entry_point_twice(gpu):
    import graphone_amd as g1
    rng = np.random.default_rng(42)
    for n_rows, n_cols, avg, hubs, hub_deg in ((60000, 70000, 15, (3, 59999), 60000), (3000, 2000, 4, (7,), 2000)):
        rp, cols = _csr(rng, n_rows, n_cols, avg, hubs, hub_deg)
        vals = rng.integers(1, 256, cols.size, dtype=np.uint32)
        x = rng.integers(0, 1 << 20, n_cols, dtype=np.uint32)
        x[rng.random(n_cols) < 0.4] = INF
        print(f"noalloc spmv: n_rows={n_rows} nnz={cols.size}", flush=True)
        want = oracle.spmv_minplus_u32(rp, cols, vals, x)
        got_ws, got_na = _both(g1, rp, cols, vals, x, n_cols)
        assert np.array_equal(got_ws, want)
        assert np.array_equal(got_na, got_ws), (n_rows, int(np.sum(got_na != got_ws)))
