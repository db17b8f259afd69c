"""Graph formats and generators (host library).  CPU only."""
import numpy as np
import pytest

from icpp_2020_parhde_amd.graphs import Graph


def test_mesh_counts():
    g = Graph.mesh(128, 128)
    assert (g.n, g.m) == (16384, 97282)
    assert g.validate()
    assert g.degrees.max() == 6


def test_grid3d_counts_and_weights():
    g = Graph.grid3d(16, 16, 16, seed=4, wmax=255)
    assert g.n == 4096 and g.m == 2 * 3 * 16 * 16 * 15
    assert g.validate()  # includes matching weights in both directions
    assert g.w.min() >= 1 and g.w.max() <= 255
    h = Graph.grid3d(16, 16, 16, seed=4, wmax=255)
    assert np.array_equal(g.w, h.w)
    k = Graph.grid3d(16, 16, 16, seed=5, wmax=255)
    assert not np.array_equal(g.w, k.w)


def test_rgg_deterministic_and_connected():
    g, xy = Graph.rgg(20000, 10.0, seed=2, with_points=True)
    h = Graph.rgg(20000, 10.0, seed=2)
    assert np.array_equal(g.adj, h.adj) and np.array_equal(g.rowptr, h.rowptr)
    assert g.validate()
    assert 0.9 * 20000 < g.n <= 20000
    assert 8.0 < g.m / g.n < 11.0
    r = np.sqrt(10.0 / (np.pi * 20000))
    for v in range(0, g.n, 997):
        nb = g.adj[g.rowptr[v]:g.rowptr[v + 1]]
        d = np.linalg.norm(xy[nb] - xy[v], axis=1)
        assert np.all(d < r)
        allv = np.flatnonzero(np.linalg.norm(xy - xy[v], axis=1) < r)
        assert set(allv.tolist()) - {v} == set(nb.tolist())


def test_rmat_deterministic():
    g = Graph.rmat(12, 16, seed=3)
    h = Graph.rmat(12, 16, seed=3)
    assert np.array_equal(g.adj, h.adj)
    assert g.validate()
    assert g.degrees.argmax() == 0  # a = 0.57 makes vertex 0 the hub


def test_holed_grid():
    g = Graph.holed_grid(60, 80, 8, 1234)
    assert g.validate() and g.n < 4800


def test_csr_roundtrip(tmp_path):
    g = Graph.grid3d(5, 6, 7, seed=1, wmax=9)
    p = tmp_path / "g.csr"
    g.write_csr(p)
    raw = np.fromfile(p, dtype=np.int64, count=6)
    assert raw.tolist() == [g.n, g.m, 1, 0, 1, 0]
    h = Graph.read_csr(p, weighted=True)
    assert np.array_equal(h.rowptr, g.rowptr) and np.array_equal(h.adj, g.adj) and np.array_equal(h.w, g.w)
    u = Graph.read_csr(p, weighted=False)
    assert u.w is None and np.array_equal(u.adj, g.adj)


def test_from_edges_dedupes_and_drops_loops():
    g = Graph.from_edges(4, [0, 1, 1, 2, 3], [1, 0, 1, 3, 2])
    assert g.m == 4 and g.validate()
    assert g.adj.tolist() == [1, 0, 3, 2]
