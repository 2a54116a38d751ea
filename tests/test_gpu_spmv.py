"""Semiring SpMV parity on the GPU (merge-path and CSR-vector) against the oracle's naive fold
(restating include/GraphOne/archi/blas/spmv/cpu/naive.hpp:5-20).

min-plus u32: bit-exact. plus-times f32 (Kahan): within 2e-6 relative of the FP64 fold per row, and the
result is deterministic (identical bits on a re-run).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

INF = 0xFFFFFFFF
PT_RTOL = 2e-6


def _random_csr(rng, n_rows, n_cols, avg, hub_rows=(), hub_deg=0, empty_frac=0.3):
    degs = rng.poisson(avg, n_rows).astype(np.int64)
    degs[rng.random(n_rows) < empty_frac] = 0
    for h in hub_rows:
        degs[h] = hub_deg
    degs = np.minimum(degs, n_cols)
    rp = np.zeros(n_rows + 1, dtype=np.int64)
    rp[1:] = np.cumsum(degs)
    cols = np.empty(rp[-1], dtype=np.uint32)
    for r in range(n_rows):
        if degs[r]:
            cols[rp[r]:rp[r + 1]] = np.sort(rng.choice(n_cols, degs[r], replace=False))
    return rp.astype(np.uint32), cols


def _dev(a):
    import graphone_amd as g1
    return g1.DeviceArray.from_numpy(a)


def _run(kind, semiring, rp, cols, vals, x, n_cols, lanes=0):
    import graphone_amd as g1
    n_rows = rp.size - 1
    d_rp, d_ci = _dev(rp), _dev(cols if cols.size else np.zeros(1, np.uint32))
    d_v = None if vals is None else _dev(vals)
    d_x = _dev(x)
    ydt = np.uint32 if semiring == "minplus_u32" else np.float32
    d_y = g1.DeviceArray(max(n_rows, 1), ydt)
    ws = g1.merge_workspace(n_rows, cols.size, d_rp) if kind == "merge" else None
    g1.spmv(kind, semiring, d_rp, d_ci, d_v, d_x, d_y, n_rows, n_cols, cols.size, ws, lanes)
    return d_y.numpy()[:n_rows]


CASES = [
    # (n_rows, n_cols, avg degree, hub rows, hub degree)
    (1, 1, 0, (), 0),
    (1, 5000, 0, (0,), 5000),
    (7, 7, 2, (), 0),
    (1000, 1000, 3, (), 0),
    (2047, 3000, 0.5, (), 0),
    (5000, 5000, 12, (17, 4000), 4999),
    (20000, 30000, 6, (0, 19999), 30000),
    (60000, 60000, 15, (5, 6, 7), 50000),
]


@pytest.mark.parametrize("kind", ["merge", "csrvec"])
@pytest.mark.parametrize("case", CASES)
def test_spmv_minplus_bitexact(gpu, kind, case):
    n_rows, n_cols, avg, hubs, hub_deg = case
    rng = np.random.default_rng(hash(case) & 0xFFFF)
    rp, cols = _random_csr(rng, n_rows, n_cols, avg, hubs, hub_deg)
    vals = rng.integers(1, 256, cols.size, dtype=np.uint32)
    x = rng.integers(0, 1 << 20, n_cols, dtype=np.uint32)
    x[rng.random(n_cols) < 0.5] = INF
    want = oracle.spmv_minplus_u32(rp, cols, vals, x)
    got = _run(kind, "minplus_u32", rp, cols, vals, x, n_cols)
    assert np.array_equal(got, want)
    # unit weights (BFS semiring)
    want1 = oracle.spmv_minplus_u32(rp, cols, None, x)
    got1 = _run(kind, "minplus_u32", rp, cols, None, x, n_cols)
    assert np.array_equal(got1, want1)


@pytest.mark.parametrize("kind", ["merge", "csrvec"])
@pytest.mark.parametrize("case", CASES)
def test_spmv_plustimes_vs_f64(gpu, kind, case):
    n_rows, n_cols, avg, hubs, hub_deg = case
    rng = np.random.default_rng((hash(case) + 1) & 0xFFFF)
    rp, cols = _random_csr(rng, n_rows, n_cols, avg, hubs, hub_deg)
    vals = rng.random(cols.size).astype(np.float32)
    x = (rng.random(n_cols) * 1e-3).astype(np.float32)
    want = oracle.spmv_plustimes_f64(rp, cols, vals.astype(np.float64), x.astype(np.float64))
    got = _run(kind, "plustimes_f32", rp, cols, vals, x, n_cols)
    err = np.abs(got.astype(np.float64) - want)
    assert np.all(err <= PT_RTOL * np.abs(want) + 1e-30), float(np.max(err / (np.abs(want) + 1e-30)))
    again = _run(kind, "plustimes_f32", rp, cols, vals, x, n_cols)
    assert np.array_equal(got.view(np.uint32), again.view(np.uint32))


@pytest.mark.parametrize("lanes", [1, 2, 4, 8, 16, 32, 64])
def test_csrvec_every_lane_width(gpu, lanes):
    rng = np.random.default_rng(lanes)
    rp, cols = _random_csr(rng, 3000, 3000, 9, (1,), 2999)
    vals = rng.integers(1, 256, cols.size, dtype=np.uint32)
    x = rng.integers(0, 1 << 20, 3000, dtype=np.uint32)
    want = oracle.spmv_minplus_u32(rp, cols, vals, x)
    assert np.array_equal(_run("csrvec", "minplus_u32", rp, cols, vals, x, 3000, lanes), want)


def test_merge_rejects_small_workspace(gpu):
    import graphone_amd as g1
    rp = np.array([0, 1], dtype=np.uint32)
    cols = np.array([0], dtype=np.uint32)
    d_rp, d_ci, d_x = _dev(rp), _dev(cols), _dev(np.zeros(1, np.uint32))
    d_y = g1.DeviceArray(1, np.uint32)
    tiny = g1.DeviceArray(1, np.uint32)
    with pytest.raises(g1.G1Error) as e:
        g1.spmv("merge", "minplus_u32", d_rp, d_ci, None, d_x, d_y, 1, 1, 1, tiny)
    assert e.value.status == 1
