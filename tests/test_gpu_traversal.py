"""Delta-stepping SSSP, 64-wide multi-source BFS, WParHDE and the exact /
generalised eigen modes on the GPU, against the oracle and library routines."""
import numpy as np
import pytest
import scipy.linalg

from icpp_2020_parhde_amd import layout as L
from icpp_2020_parhde_amd.graphs import Graph
from oracle import pyoracle as O

from test_gpu_core import aligned_rel_err, FP64_TOL

pytestmark = pytest.mark.gpu
INF = 2**30 - 1


def weighted_rmat(scale, seed):
    g = Graph.rmat(scale, 8, seed=seed)
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(g.n), np.diff(g.rowptr.astype(np.int64)))
    m = rows < g.adj
    u, v = rows[m], g.adj[m].astype(np.int64)
    return Graph.from_edges(g.n, u, v, rng.integers(1, 1000, size=len(u)))


@pytest.fixture(scope="module")
def wgraphs():
    return {"grid": Graph.grid3d(24, 20, 18, seed=4, wmax=255), "rmat": weighted_rmat(12, 21)}


@pytest.mark.parametrize("name", ["grid", "rmat"])
@pytest.mark.parametrize("delta", [1, 37, 256, 4096, 1 << 24, 0])
def test_sssp_bit_exact(gpu, wgraphs, name, delta):
    g = wgraphs[name]
    with L.Context(g) as ctx:
        for src in (0, g.n // 2, g.n - 1):
            want = O.dijkstra(g, src)
            got = ctx.sssp(src, delta).astype(np.int64)
            assert np.array_equal(got, want), (name, delta, src)


def test_sssp_unreached(gpu):
    g = Graph.from_edges(5, [0, 2], [1, 3], [7, 9])
    with L.Context(g) as ctx:
        assert ctx.sssp(0, 4).tolist() == [0, 7, INF, INF, INF]
        assert ctx.sssp(3, 4).tolist() == [INF, INF, 9, 0, INF]


@pytest.mark.parametrize("name", ["mesh", "rgg", "rmat", "star"])
@pytest.mark.parametrize("k", [1, 7, 64])
def test_msbfs_equals_k_bfs(gpu, name, k):
    g = {"mesh": lambda: Graph.mesh(70, 90), "rgg": lambda: Graph.rgg(30000, 10.0, seed=8),
         "rmat": lambda: Graph.rmat(13, 16, seed=4), "star": lambda: Graph.star(10000)}[name]()
    rng = np.random.default_rng(k)
    src = rng.choice(g.n, size=k, replace=False)
    with L.Context(g) as ctx:
        D = ctx.msbfs(src)
    for j, s in enumerate(src):
        assert np.array_equal(D[:, j], O.seq_bfs(g, int(s))), (name, k, j)


def test_msbfs_duplicate_sources_and_disconnected(gpu):
    g = Graph.from_edges(6, [0, 1, 3], [1, 2, 4])
    with L.Context(g) as ctx:
        D = ctx.msbfs([0, 0, 4, 5])
    assert D[:, 0].tolist() == [0, 1, 2, -1, -1, -1]
    assert D[:, 1].tolist() == D[:, 0].tolist()
    assert D[:, 2].tolist() == [-1, -1, -1, 1, 0, -1]
    assert D[:, 3].tolist() == [-1, -1, -1, -1, -1, 0]


@pytest.mark.parametrize("s,c", [(10, 3), (20, 2)])
def test_wparhde_parity(gpu, wgraphs, s, c):
    g = wgraphs["grid"]
    ref = O.wparhde(g, s, c, delta=10000)
    with L.Context(g) as ctx:
        X, st = ctx.layout(s=s, num_coord=c, algorithm=L.ALGO_WPARHDE)
        assert st.pivots == ref["pivots"]
        err = aligned_rel_err(X, ref["coords"])
        print(f"wparhde s={s} c={c} rel err {err:.2e}")
        assert err < FP64_TOL
        X2, _ = ctx.layout(s=s, num_coord=c, algorithm=L.ALGO_WPARHDE, delta=64)
        assert aligned_rel_err(X2, X) < 1e-12  # delta changes nothing but speed


def _ref_exact(g, s, generalized):
    r = O.parhde(g, s, 2, lazy=True, want_S=True)
    S = r["S"]
    A = r["SLS"].reshape(s, s)  # row i = S_i^T L S
    A = 0.5 * (A + A.T)
    if generalized:
        d = np.diff(g.rowptr.astype(np.int64)).astype(np.float64)
        B = S.T @ (d[:, None] * S)
        _, V = scipy.linalg.eigh(A, B)
    else:
        _, V = np.linalg.eigh(A)
    X = S @ V[:, :2]
    return X / np.linalg.norm(X, axis=0)


@pytest.mark.parametrize("mode", [L.EIG_JACOBI, L.EIG_GENERALIZED])
def test_exact_eigen_modes(gpu, mode):
    g = Graph.rgg(20000, 10.0, seed=2)
    s = 30
    want = _ref_exact(g, s, mode == L.EIG_GENERALIZED)
    with L.Context(g) as ctx:
        X, st = ctx.layout(s=s, num_coord=2, eig_mode=mode)
    err = aligned_rel_err(X, want)
    print(f"eig mode {mode}: rel err {err:.2e}")
    assert err < 1e-7
