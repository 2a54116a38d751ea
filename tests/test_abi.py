"""The C-ABI library loads and exports every symbol include/graphone_hip.h declares; argument errors are
reported with status codes before any GPU work (no GPU needed)."""
import ctypes as C
import subprocess

import numpy as np
import pytest

import graphone_amd as g1
from graphone_amd import _lib


def test_every_declared_symbol_is_exported():
    declared = _lib.declared_symbols()
    assert len(declared) >= 40
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    # and every binding of the Python layer names a declared symbol
    assert set(_lib._SIGNATURES) <= set(declared)


def test_status_strings_and_version():
    L = _lib.lib()
    assert L.g1_abi_version() == 1
    assert L.g1_status_string(0) == b"ok"
    assert L.g1_status_string(4) == b"HBM budget exceeded"


def test_generator_argument_errors():
    L = _lib.lib()
    csr = _lib.HostCSR()
    assert L.g1_gen_rmat(0, 16, 0.57, 0.19, 0.19, 1, 0, C.byref(csr)) == _lib.G1_ERR_INVALID_ARG
    assert b"scale" in L.g1_last_error_message()
    assert L.g1_gen_rmat(5, 16, 0.7, 0.3, 0.3, 1, 0, C.byref(csr)) == _lib.G1_ERR_INVALID_ARG
    assert L.g1_gen_grid(0, 5, 0.5, 1, C.byref(csr)) == _lib.G1_ERR_INVALID_ARG


def test_graph_create_validates_csr_on_host():
    bad = g1.HostGraph.from_csr(np.array([0, 2, 1], np.uint32), np.array([1, 0], np.uint32))
    with pytest.raises(g1.G1Error) as e:
        g1.Graph(bad)
    assert e.value.status == _lib.G1_ERR_INVALID_ARG
    bad_col = g1.HostGraph.from_csr(np.array([0, 1, 1], np.uint32), np.array([7], np.uint32))
    with pytest.raises(g1.G1Error):
        g1.Graph(bad_col)


def test_merge_workspace_size_formula():
    L = _lib.lib()
    assert L.g1_spmv_merge_workspace_bytes(0, 0) > 0
    a = L.g1_spmv_merge_workspace_bytes(4194304, 65000000)
    tiles = (4194304 + 65000000 + 2047) // 2048
    assert a >= (tiles + 1) * 8 + tiles * 32
    assert L.g1_spmv_csrvec_lanes(4194304, 65000000) == 8  # average 15: half, rounded up to 2^k
    assert L.g1_spmv_csrvec_lanes(45369, 108374) == 1  # integer average 2
    assert L.g1_spmv_csrvec_lanes(10, 5000) == 64


def test_server_argument_errors_without_a_gpu():
    """Failure paths of the query server that need no device: bad configuration, unknown graph or query,
    waiting on or fetching what does not exist."""
    with pytest.raises(g1.G1Error) as e:
        g1.Server(hbm_budget_bytes=1 << 20, batch_k=12)
    assert e.value.status == _lib.G1_ERR_INVALID_ARG
    with pytest.raises(g1.G1Error):
        g1.Server(hbm_budget_bytes=1 << 20, ppr_restart=1.5)
    srv = g1.Server(hbm_budget_bytes=1 << 20)
    with pytest.raises(g1.G1Error) as e:
        srv.submit(0, "bfs", 0)
    assert e.value.status == _lib.G1_ERR_NOT_FOUND
    for fn in (srv.poll, srv.cancel):
        with pytest.raises(g1.G1Error) as e:
            fn(12345)
        assert e.value.status == _lib.G1_ERR_NOT_FOUND
    with pytest.raises(g1.G1Error):
        srv.wait(12345, 1.0)
    s = srv.stats()
    assert s["hbm_budget"] == 1 << 20 and s["submitted"] == 0
    with pytest.raises(g1.G1Error) as e:
        srv.remove_graph(3)
    assert e.value.status == _lib.G1_ERR_NOT_FOUND
    with pytest.raises(g1.G1Error) as e:
        srv.set_budget(0)
    assert e.value.status == _lib.G1_ERR_INVALID_ARG
    srv.set_budget(2 << 20)
    assert srv.stats()["hbm_budget"] == 2 << 20
    srv.close()
    L = _lib.lib()
    assert L.g1_server_destroy(None) == _lib.G1_OK
    assert L.g1_server_get_stats(None, None) == _lib.G1_ERR_INVALID_ARG


def _fake_graph(n):
    """A Graph object with no device handle: the Python layer's argument checks run before any C call."""
    g = g1.Graph.__new__(g1.Graph)
    g._h = C.c_void_p()
    g.n, g.m, g.weighted = n, 0, True
    return g


def _fake_array(count, dtype):
    a = g1.DeviceArray.__new__(g1.DeviceArray)
    a.count, a.dtype, a._ptr = count, np.dtype(dtype), 0
    a.nbytes = count * a.dtype.itemsize
    return a


def test_output_buffers_checked_before_device_work():
    """An undersized or mistyped output array is refused on the host: the library would write n (or
    sources x n) elements through its pointer."""
    g = _fake_graph(100)
    for fn, args in ((g.bfs, (0,)), (g.sssp, (0,))):
        with pytest.raises(ValueError, match="99 elements, 100 needed"):
            fn(*args, out=_fake_array(99, np.uint32))
        with pytest.raises(ValueError, match="dtype"):
            fn(*args, out=_fake_array(100, np.float32))
    with pytest.raises(ValueError, match="needed"):
        g.pagerank(out=_fake_array(50, np.float32))
    with pytest.raises(ValueError, match="dtype"):
        g.pagerank(out=_fake_array(100, np.uint32))
    with pytest.raises(ValueError, match="needed"):
        g.ppr(0, out=_fake_array(10, np.float32))
    with pytest.raises(ValueError, match="300 needed"):
        g.sssp_batch([0, 1, 2], out=_fake_array(299, np.uint32))
    with pytest.raises(ValueError, match="200 needed"):
        g.ppr_batch([0, 1], out=_fake_array(100, np.float32))
    with pytest.raises(ValueError, match="200 needed"):
        g.road_batch([0, 1], out=_fake_array(199, np.uint32))
    with pytest.raises(ValueError, match="1 kinds for 2 sources"):
        g.road_batch([0, 1], kinds=["bfs"])
    with pytest.raises(TypeError):
        g.bfs(0, out=np.zeros(100, np.uint32))


def test_noalloc_spmv_rejects_oversize_before_device_work():
    """g1_spmv_merge_minplus_u32_noalloc refuses a matrix whose workspace exceeds its static scratch with a
    status code and a message, before touching any pointer or the device."""
    L = _lib.lib()
    cap = L.g1_spmv_merge_noalloc_max_workspace_bytes()
    assert cap == 8 << 20
    n_rows, nnz = 1 << 26, 400_000_000
    assert L.g1_spmv_merge_workspace_bytes(n_rows, nnz) > cap
    st = L.g1_spmv_merge_minplus_u32_noalloc(n_rows, n_rows, nnz, None, None, None, None, None, None)
    assert st == _lib.G1_ERR_INVALID_ARG
    assert b"scratch" in L.g1_last_error_message()


def test_header_constants_match_bindings():
    """Every integer #define G1_* of include/graphone_hip.h that the Python layer also names has the same
    value there (engine ids, status codes, modes)."""
    import os
    import re
    header = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "graphone_hip.h")
    defines = dict(re.findall(r"^#define (G1_[A-Z0-9_]+) (-?\d+)\s*$", open(header).read(), flags=re.M))
    named = {k: v for k, v in vars(_lib).items() if k.startswith("G1_") and isinstance(v, int)}
    shared = sorted(set(defines) & set(named))
    assert "G1_ENGINE_DELTA" in shared and "G1_ENGINE_ADVANCE" in shared
    for k in shared:
        assert int(defines[k]) == named[k], k


def test_work_area_calls_reject_bad_arguments():
    L = _lib.lib()
    b = C.c_uint64()
    assert L.g1_graph_work_bytes(None, _lib.G1_WORK_PUSH, C.byref(b)) == _lib.G1_ERR_INVALID_ARG
    assert L.g1_graph_prepare(None, _lib.G1_WORK_PUSH, None) == _lib.G1_ERR_INVALID_ARG
    assert b"null" in L.g1_last_error_message()
