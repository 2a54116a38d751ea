"""The reference's own applications (apps/sssp_adv.cu, sssp_blas.cu, sssp_blas2.cu, pagerank_blas.cu),
built against graphone_amd's graph_one C++ API with the documented four-rule port patch
(tests/dropin/Makefile), run on the GPU on .adj files and checked against the oracle; plus this
repository's own BFS app on the same API.

SSSP distances are exact: the apps compute in float, and every path sum here is an integer below 2^24.
PageRank runs the reference's own stopping rule (max |r' - r| <= 1e-6, apps/pagerank_blas.cu:63-78) and
prints "%f" (6 decimals), so it is compared with the oracle's same rule within 2e-6 per vertex.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "dropin", "_bin")
APP_BIN = os.path.join(os.path.dirname(HERE), "graphone_amd", "apps", "bin")
FLT_MAX = float(np.finfo(np.float32).max)


def _need(path):
    if not os.path.exists(path):
        pytest.skip(f"{path} not built (the drop-in apps build only where the reference tree is present)")
    return path


def _write_adj(path, src, dst, w=None):
    with open(path, "w") as f:
        f.write("% generated graph\n")
        for i in range(len(src)):
            if w is None:
                f.write(f"{src[i]} {dst[i]}\n")
            else:
                f.write(f"{src[i]}\t{dst[i]}\t{w[i]}\n")


def _read_out(path):
    vals = []
    with open(path) as f:
        lines = f.read().splitlines()
    assert lines[0].startswith("Elapsed time:")
    for i, line in enumerate(lines[1:]):
        v, x = line.split("-", 1)
        assert int(v) == i
        vals.append(float(x))
    return np.array(vals)


def _graph(seed):
    import graphone_amd as g1
    rng = np.random.default_rng(seed)
    # an undirected weighted graph (the loader's default): an RMAT core plus a few isolated ids
    h = g1.HostGraph.rmat(11, 6, seed=seed)
    rp, ci = h.row_offsets, h.col_indices
    src = np.repeat(np.arange(h.n, dtype=np.uint32), np.diff(rp.astype(np.int64)))
    keep = src < ci  # one direction per pair; the loader adds the reverse
    src, dst = src[keep], ci[keep]
    w = rng.integers(1, 64, src.size, dtype=np.uint32)
    return src, dst, w


def _run(args, tmp_path):
    out = str(tmp_path / "out.txt")
    subprocess.run(args + ["-o", out], check=True, capture_output=True, text=True, timeout=120)
    return _read_out(out)


@pytest.mark.parametrize("app", ["sssp_adv", "sssp_blas", "sssp_blas2"])
def test_reference_sssp_apps_bitexact(gpu, tmp_path, app):
    import graphone_amd as g1
    exe = _need(os.path.join(BIN, app))
    src, dst, w = _graph(3)
    adj = str(tmp_path / "g.adj")
    _write_adj(adj, src, dst, w)
    n = int(max(src.max(), dst.max())) + 1
    host = g1.HostGraph.from_edges(n, src, dst, w, directed=False)
    for s in (0, 1, int(np.argmax(host.out_degrees()))):
        got = _run([exe, "-i", adj, "--src", str(s)], tmp_path)
        want = oracle.dijkstra(host.row_offsets, host.col_indices, host.weights, s).astype(np.float64)
        want[want == oracle.INF] = FLT_MAX
        assert got.shape == want.shape
        assert np.array_equal(got, want), (app, s)


def test_reference_sssp_app_on_adaptive_frontier(gpu, tmp_path):
    """sssp_adv with its frontier declared as the README's DblBufFrontier: iterations whose frontier holds
    at least G1_DBL_PULL of the edges pull over the in-edges (built on the device: the app keeps only the
    out-edge view), the others push; the distances stay bit-exact and both directions are taken."""
    import graphone_amd as g1
    exe = _need(os.path.join(BIN, "sssp_adv_dbl"))
    src, dst, w = _graph(7)
    adj = str(tmp_path / "g.adj")
    _write_adj(adj, src, dst, w)
    n = int(max(src.max(), dst.max())) + 1
    host = g1.HostGraph.from_edges(n, src, dst, w, directed=False)
    env = dict(os.environ, G1_DBL_TRACE="1", G1_DBL_PULL="0.05")
    for s in (int(np.argmax(host.out_degrees())), 1):
        out = str(tmp_path / "out.txt")
        r = subprocess.run([exe, "-i", adj, "--src", str(s), "-o", out], capture_output=True, text=True,
                           timeout=120, env=env)
        assert r.returncode == 0, r.stdout + r.stderr
        got = _read_out(out)
        want = oracle.dijkstra(host.row_offsets, host.col_indices, host.weights, s).astype(np.float64)
        want[want == oracle.INF] = FLT_MAX
        assert np.array_equal(got, want), s
        line = [x for x in r.stderr.splitlines() if "DblBufFrontier" in x][-1]
        pushes, pulls = int(line.split()[2]), int(line.split()[5])
        assert pushes >= 1 and pulls >= 1, line


def test_reference_pagerank_app(gpu, tmp_path):
    import graphone_amd as g1
    exe = _need(os.path.join(BIN, "pagerank_blas"))
    src, dst, _ = _graph(4)
    adj = str(tmp_path / "g.adj")
    _write_adj(adj, src, dst)
    n = int(max(src.max(), dst.max())) + 1
    host = g1.HostGraph.from_edges(n, src, dst, None, directed=False)
    t = host.transpose()
    got = _run([exe, "-i", adj], tmp_path)
    want, st = oracle.pagerank_f64(host.row_offsets, None, t.row_offsets, t.col_indices, None, 0.85,
                                   max_iter=100000, tol=1e-6, mode=1)
    assert st["iterations"] > 5
    assert float(np.abs(got - want).max()) <= 2e-6


def test_reference_sssp_app_directed_and_errors(gpu, tmp_path):
    import graphone_amd as g1
    exe = _need(os.path.join(BIN, "sssp_adv"))
    src, dst, w = _graph(5)
    adj = str(tmp_path / "d.adj")
    _write_adj(adj, src, dst, w)
    n = int(max(src.max(), dst.max())) + 1
    host = g1.HostGraph.from_edges(n, src, dst, w, directed=True)
    got = _run([exe, "-i", adj, "--src", "2", "-d"], tmp_path)
    want = oracle.dijkstra(host.row_offsets, host.col_indices, host.weights, 2).astype(np.float64)
    want[want == oracle.INF] = FLT_MAX
    assert np.array_equal(got, want)
    # the loader's error behaviour: a wrong file extension exits with status 1 (LOG_ERROR)
    bad = str(tmp_path / "g.txt")
    _write_adj(bad, src, dst, w)
    r = subprocess.run([exe, "-i", bad, "--src", "0"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "extension" in r.stdout
    r = subprocess.run([exe, "-i", adj], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0  # --src is required


def test_bfs_app_on_graph_one_api(gpu, tmp_path):
    import graphone_amd as g1
    exe = _need(os.path.join(APP_BIN, "bfs_adv"))
    src, dst, _ = _graph(6)
    adj = str(tmp_path / "g.adj")
    _write_adj(adj, src, dst)
    n = int(max(src.max(), dst.max())) + 1
    host = g1.HostGraph.from_edges(n, src, dst, None, directed=False)
    s = int(np.argmax(host.out_degrees()))
    got = _run([exe, "-i", adj, "--src", str(s)], tmp_path)
    want = oracle.bfs_queue(host.row_offsets, host.col_indices, s).astype(np.int64)
    want[want == oracle.INF] = -1
    assert np.array_equal(got.astype(np.int64), want)
