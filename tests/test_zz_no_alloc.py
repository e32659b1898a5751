"""gpm_tc_edge_noalloc: the triangle count that allocates nothing (its
staging lives in a fixed device array) must equal gpm_tc's edge-parallel
count, call after call, including a second, smaller input after a larger
one (the shared array keeps the previous call's contents).

The test's own device buffers come from the HIP runtime libgpm is linked
against (hipMalloc through ctypes): a second HIP runtime in the process,
such as the one bundled with torch, does not see the device once libgpm's
runtime holds it."""
import ctypes as C

import numpy as np
import pytest

import graphs

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip(gpm):
    gpm.lib()
    h = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    h.hipGetErrorString.restype = C.c_char_p
    return h


def _ok(hip, st):
    assert st == 0, hip.hipGetErrorString(st)


class Dev:
    def __init__(self, hip, nbytes):
        self.hip, self.p = hip, C.c_void_p()
        _ok(hip, hip.hipMalloc(C.byref(self.p), C.c_size_t(max(nbytes, 4))))

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        if arr.nbytes:
            _ok(self.hip, self.hip.hipMemcpy(self.p, arr.ctypes.data_as(C.c_void_p),
                                             C.c_size_t(arr.nbytes), H2D))
        return self

    def get(self, dtype, count):
        out = np.zeros(count, dtype=dtype)
        _ok(self.hip, self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p,
                                         C.c_size_t(out.nbytes), D2H))
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _noalloc_count(gpm, hip, stream, d):
    off = Dev(hip, 4 * (d.n + 1)).put(d.offsets.astype(np.uint32))
    adj = Dev(hip, 4 * d.m).put(d.adj.astype(np.uint32))
    out = Dev(hip, 8).put(np.array([12345], dtype=np.uint64))
    try:
        gpm.tc_edge_noalloc(off.p.value, adj.p.value, d.n, d.m, out.p.value, stream.value or 0)
        _ok(hip, hip.hipStreamSynchronize(stream))
        return int(out.get(np.uint64, 1)[0])
    finally:
        for b in (off, adj, out):
            b.free()


def _edge_count(gpm, d):
    with gpm.DeviceGraph(d, oriented=True) as dg:
        return dg.tc(gpm.TC_EDGE)


def test_noalloc_matches_edge_kernel(gpm, hip):
    stream = C.c_void_p()
    _ok(hip, hip.hipStreamCreate(C.byref(stream)))
    big = gpm.orient(gpm.er_graph(16384, 131072, 20240601))  # triangle-counting workload
    small = gpm.orient(gpm.rmat_graph(10, 16, 3))
    tiny = gpm.orient(graphs.complete(7))
    try:
        for d in (big, small, tiny, big):
            want = _edge_count(gpm, d)
            print(f"noalloc n={d.n} m={d.m} want={want}", flush=True)
            assert _noalloc_count(gpm, hip, stream, d) == want
            assert _noalloc_count(gpm, hip, stream, d) == want  # a second call in a row
        assert _edge_count(gpm, tiny) == 35
    finally:
        hip.hipStreamDestroy(stream)


def test_noalloc_empty_and_limit(gpm, hip):
    empty = gpm.orient(graphs.from_edges(5, []))
    assert _noalloc_count(gpm, hip, C.c_void_p(), empty) == 0  # null stream
    with pytest.raises(gpm.GpmError):
        gpm.tc_edge_noalloc(1, 1, 10, (16 << 20) + 1, 1, 0)
