"""The multi-hop traversal (a wave expands its part of the frontier for
several hops between grid barriers; levels corrected by atomicMin when a
shorter path arrives later) against the oracle's BFS: the levels must be
bit-exact whatever the timing.  Forced on small graphs with PHDE_BFS_HOPS
(contexts read it when they are created); PHDE_BFS_QCAP shrinks its queue
so that the level-by-level repeat after an overflow is exercised too."""
import numpy as np
import pytest

from icpp_2020_parhde_amd import layout as L
from icpp_2020_parhde_amd.graphs import Graph
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

GRAPHS = {
    "path": lambda: Graph.path(3000),
    "mesh128": lambda: Graph.mesh(128, 128),
    "grid3d": lambda: Graph.grid3d(20, 18, 16, wmax=0),
    "rgg20k": lambda: Graph.rgg(20000, 10.0, seed=2),
    "holed": lambda: Graph.holed_grid(120, 90, 10, 77),
    "rmat14": lambda: Graph.rmat(14, 16, seed=3),
    "star": lambda: Graph.star(20000),
}


@pytest.fixture(scope="module")
def graphs():
    return {k: f() for k, f in GRAPHS.items()}


@pytest.mark.parametrize("hops", [2, 5, 16])
@pytest.mark.parametrize("name", list(GRAPHS))
def test_multihop_bfs_bit_exact(gpu, graphs, name, hops, monkeypatch):
    g = graphs[name]
    monkeypatch.setenv("PHDE_BFS_HOPS", str(hops))
    with L.Context(g) as ctx:
        for src in (0, g.n // 2, g.n - 1):
            want = O.seq_bfs(g, src)
            for _ in range(2):
                assert np.array_equal(ctx.bfs(src), want), (name, hops, src)


@pytest.mark.parametrize("name", ["path", "rgg20k", "rmat14"])
def test_multihop_queue_overflow_repeats_exactly(gpu, graphs, name, monkeypatch):
    g = graphs[name]
    monkeypatch.setenv("PHDE_BFS_HOPS", "4")
    monkeypatch.setenv("PHDE_BFS_QCAP", "3")
    with L.Context(g) as ctx:
        for src in (0, g.n - 1):
            assert np.array_equal(ctx.bfs(src), O.seq_bfs(g, src)), (name, src)


def test_multihop_layout_and_depth_match_level_by_level(gpu, graphs, monkeypatch):
    """k-centers pivots, the layout and the traversal depths (the stats'
    level count) do not depend on how the levels were computed."""
    g = graphs["holed"]
    with L.Context(g) as ctx:
        X0, st0 = ctx.layout(s=12, num_coord=2)
    monkeypatch.setenv("PHDE_BFS_HOPS", "6")
    with L.Context(g) as ctx:
        X1, st1 = ctx.layout(s=12, num_coord=2)
    assert np.array_equal(X0, X1)
    assert st0.levels_total == st1.levels_total


SEL_GRAPHS = {
    "mesh": lambda: Graph.mesh(60, 70),
    "rgg": lambda: Graph.rgg(8000, 10.0, seed=11),
    "holed": lambda: Graph.holed_grid(80, 90, 6, 12),
    "rmat": lambda: Graph.rmat(12, 16, seed=13),
}


@pytest.mark.parametrize("hops", [3, 12])
@pytest.mark.parametrize("name", list(SEL_GRAPHS))
def test_multihop_greedy_selection_bit_exact(gpu, name, hops, monkeypatch):
    """The greedy pivot selection's bounded BFS run several hops per grid
    barrier picks exactly the oracle's pivots (and so does the level-by-level
    repeat after a forced queue overflow)."""
    from oracle import learned as OL
    g = SEL_GRAPHS[name]()
    rng = np.random.default_rng(9)
    score = rng.random(g.n).astype(np.float32)
    d1 = O.seq_bfs(g, 0)
    second = int(np.argmax(d1))
    d2 = O.seq_bfs(g, second)
    monkeypatch.setenv("PHDE_BFS_HOPS", str(hops))
    for k, r in [(64, 3), (20, 1), (40, 9)]:
        want = OL.greedy_select(g, score, d1, d2, 0, second, k, r)
        with L.Context(g) as ctx:
            got = ctx.select_pivots(score, k, r, 0)
        assert got.tolist() == want.tolist(), (name, hops, k, r)
    monkeypatch.setenv("PHDE_BFS_QCAP", "3")
    with L.Context(g) as ctx:
        got = ctx.select_pivots(score, 40, 9, 0)
    assert got.tolist() == OL.greedy_select(g, score, d1, d2, 0, second, 40, 9).tolist()


@pytest.mark.parametrize("hops", [4, 12])
@pytest.mark.parametrize("name", ["mesh", "rgg", "rmat", "star"])
@pytest.mark.parametrize("k", [1, 7, 64])
def test_multihop_msbfs_equals_k_bfs(gpu, name, k, hops, monkeypatch):
    """k concurrent multi-hop searches, one column each, equal k BFS."""
    g = {"mesh": lambda: Graph.mesh(70, 90), "rgg": lambda: Graph.rgg(30000, 10.0, seed=8),
         "rmat": lambda: Graph.rmat(13, 16, seed=4), "star": lambda: Graph.star(10000)}[name]()
    rng = np.random.default_rng(k)
    src = rng.choice(g.n, size=k, replace=False)
    monkeypatch.setenv("PHDE_BFS_HOPS", str(hops))
    with L.Context(g) as ctx:
        D = ctx.msbfs(src)
    for j, s in enumerate(src):
        assert np.array_equal(D[:, j], O.seq_bfs(g, int(s))), (name, k, j)


def test_multihop_msbfs_duplicates_disconnected_and_overflow(gpu, monkeypatch):
    g = Graph.from_edges(6, [0, 1, 3], [1, 2, 4])
    monkeypatch.setenv("PHDE_BFS_HOPS", "3")
    with L.Context(g) as ctx:
        D = ctx.msbfs([0, 0, 4, 5])
    assert D[:, 0].tolist() == [0, 1, 2, -1, -1, -1]
    assert D[:, 1].tolist() == D[:, 0].tolist()
    assert D[:, 2].tolist() == [-1, -1, -1, 1, 0, -1]
    assert D[:, 3].tolist() == [-1, -1, -1, -1, -1, 0]
    # a queue too small for the searches: the bit-parallel sweep repeats them
    g = Graph.rgg(30000, 10.0, seed=8)
    src = np.random.default_rng(3).choice(g.n, size=40, replace=False)
    monkeypatch.setenv("PHDE_BFS_QCAP", "50")
    with L.Context(g) as ctx:
        D = ctx.msbfs(src)
    for j, s in enumerate(src):
        assert np.array_equal(D[:, j], O.seq_bfs(g, int(s))), j


@pytest.fixture(scope="module")
def wgraphs():
    from test_gpu_traversal import weighted_rmat
    return {"grid": Graph.grid3d(24, 20, 18, seed=4, wmax=255), "rmat": weighted_rmat(12, 21),
            "path": Graph.from_edges(2000, list(range(1999)), list(range(1, 2000)), [1 + (i * 7) % 13 for i in range(1999)])}


@pytest.mark.parametrize("hops", [2, 8, 20])
@pytest.mark.parametrize("name", ["grid", "rmat", "path"])
@pytest.mark.parametrize("delta", [1, 37, 4096, 0])
def test_multihop_sssp_bit_exact(gpu, wgraphs, name, delta, hops, monkeypatch):
    """Delta-stepping rounds that relax several hops per launch within the
    bucket give Dijkstra's distances exactly."""
    g = wgraphs[name]
    monkeypatch.setenv("PHDE_BFS_HOPS", "12")
    monkeypatch.setenv("PHDE_SSSP_HOPS", str(hops))
    with L.Context(g) as ctx:
        for src in (0, g.n // 2, g.n - 1):
            want = O.dijkstra(g, src)
            got = ctx.sssp(src, delta).astype(np.int64)
            assert np.array_equal(got, want), (name, delta, hops, src)


def test_multihop_wparhde_layout_matches_one_hop(gpu, wgraphs, monkeypatch):
    g = wgraphs["grid"]
    with L.Context(g) as ctx:
        X0, st0 = ctx.layout(s=10, num_coord=3, algorithm=L.ALGO_WPARHDE)
    monkeypatch.setenv("PHDE_BFS_HOPS", "12")
    with L.Context(g) as ctx:
        X1, st1 = ctx.layout(s=10, num_coord=3, algorithm=L.ALGO_WPARHDE)
    assert st0.pivots == st1.pivots
    assert np.array_equal(X0, X1)


@pytest.mark.parametrize("name", ["rgg", "rmat"])
def test_multihop_msbfs_two_blocks_per_cu(gpu, name, monkeypatch):
    """The concurrent searches on two blocks per CU (the setting of large
    graphs) give the same levels."""
    g = {"rgg": lambda: Graph.rgg(30000, 10.0, seed=8), "rmat": lambda: Graph.rmat(13, 16, seed=4)}[name]()
    src = np.random.default_rng(5).choice(g.n, size=52, replace=False)
    monkeypatch.setenv("PHDE_BFS_HOPS", "12")
    monkeypatch.setenv("PHDE_MS_HOP_BLOCKS_PER_CU", "2")
    with L.Context(g) as ctx:
        D = ctx.msbfs(src)
    for j, s in enumerate(src):
        assert np.array_equal(D[:, j], O.seq_bfs(g, int(s))), (name, j)
