"""The reference's free functions DOBFS and DeltaStep
(variants/DrawingBase.h:62-67, C++ linkage), called through their mangled
names: results match the oracle, also after the caller changes its arrays in
place between calls (the library keeps the last graph resident, keyed on the
arrays' contents as well as their addresses; gapbs/bfs.cc:144 and
gapbs/sssp.cc:53 read the arrays on every call)."""
import ctypes

import numpy as np
import pytest

from icpp_2020_parhde_amd import _native
from icpp_2020_parhde_amd.graphs import Graph
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

UNREACHED_SSSP = float((1 << 30) - 1)


def _fn(name, argtypes):
    f = getattr(_native.parhde_lib(), name)
    f.argtypes = argtypes
    return f


def _dobfs():
    f = _fn("_Z5DOBFSPjS_lllPdii", [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long, ctypes.c_long,
                                     ctypes.c_long, ctypes.c_void_p, ctypes.c_int, ctypes.c_int])
    f.restype = None
    return f


def _deltastep():
    f = _fn("_Z9DeltaStepPKjPjPilllPdi", [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long,
                                          ctypes.c_long, ctypes.c_long, ctypes.c_void_p, ctypes.c_int32])
    f.restype = ctypes.c_void_p
    return f


def test_deltastep_weights_changed_in_place(gpu):
    g = Graph.grid3d(12, 10, 8, seed=41, wmax=50)
    row = np.ascontiguousarray(g.rowptr, np.uint32)
    col = np.ascontiguousarray(g.adj, np.uint32)
    w = np.ascontiguousarray(g.w, np.int32)
    ds = _deltastep()
    for src, new_w in ((5, None), (5, (w.astype(np.int64) * 7) % 50 + 1), (17, None)):
        if new_w is not None:
            w[:] = new_w  # same buffer, same address, new contents (both directions alike)
        d = np.full(g.n, -7.0)
        ds(row.ctypes.data, col.ctypes.data, w.ctypes.data, src, g.n, g.m, d.ctypes.data, 8)
        ref = O.dijkstra(Graph(g.n, g.m, row, col, w.copy()), src).astype(np.float64)
        assert np.array_equal(d, ref), (src, new_w is not None)


def test_dobfs_adjacency_changed_in_place(gpu):
    # the same degrees (one rowptr) and edge count: a 2000-cycle, then two
    # 1000-cycles written over its adjacency
    n = 2000
    a = Graph.cycle(n)
    u = np.arange(n)
    h = n // 2
    v2 = np.where(u < h, (u + 1) % h, h + (u - h + 1) % h)
    b = Graph.from_edges(n, u, v2)
    assert a.m == b.m and np.array_equal(a.rowptr, b.rowptr)
    row = np.ascontiguousarray(a.rowptr, np.uint32)
    col = np.ascontiguousarray(a.adj, np.uint32)
    bfs = _dobfs()
    for g in (a, b):
        if g is b:
            col[:] = np.ascontiguousarray(b.adj, np.uint32)
        d = np.full(n, -1.0)
        bfs(row.ctypes.data, col.ctypes.data, 3, n, g.m, d.ctypes.data, 15, 18)
        assert np.array_equal(d, O.dobfs(g, 3)), g is b
