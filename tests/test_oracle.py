"""Pins the CPU oracle (oracle/liboracle.so) before it is trusted as the
GPU path's checker: golden vectors captured from the reference, the
reference's own compiled power method, closed forms and library routines.
CPU only."""
import json
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components, shortest_path

from icpp_2020_parhde_amd.graphs import Graph
from oracle import pyoracle as O

GOLDEN = Path(__file__).resolve().parent / "golden"


def mesh_dist(rows, cols, src):
    """Closed form on the triangulated mesh with (i,j)-(i+1,j+1) diagonals."""
    si, sj = divmod(src, cols)
    i, j = np.divmod(np.arange(rows * cols), cols)
    di, dj = i - si, j - sj
    same = di * dj >= 0
    return np.where(same, np.maximum(np.abs(di), np.abs(dj)), np.abs(di) + np.abs(dj))


def csr_matrix(g, weighted=False):
    data = g.w.astype(np.float64) if (weighted and g.w is not None) else np.ones(g.m)
    return sp.csr_matrix((data, g.adj.astype(np.int64), g.rowptr.astype(np.int64)), shape=(g.n, g.n))


@pytest.fixture(scope="module")
def mesh128():
    return Graph.mesh(128, 128)


# ------------------------------------------------------------------ traversals

@pytest.mark.parametrize("src", [0, 127, 5332, 16383, 8000])
def test_bfs_mesh_closed_form(mesh128, src):
    want = mesh_dist(128, 128, src)
    assert np.array_equal(O.seq_bfs(mesh128, src), want)
    assert np.array_equal(O.dobfs(mesh128, src).astype(np.int64), want)


def test_bfs_grid3d_is_l1():
    g = Graph.grid3d(9, 7, 5, wmax=0)
    src = 123
    x, y, z = np.unravel_index(np.arange(g.n), (9, 7, 5))
    sx, sy, sz = np.unravel_index(src, (9, 7, 5))
    want = np.abs(x - sx) + np.abs(y - sy) + np.abs(z - sz)
    assert np.array_equal(O.dobfs(g, src).astype(np.int64), want)


@pytest.mark.parametrize("gen", ["rgg", "rmat", "holed", "star", "path", "cycle"])
def test_bfs_matches_scipy(gen):
    g = {"rgg": lambda: Graph.rgg(3000, 8.0, seed=11), "rmat": lambda: Graph.rmat(11, 8, seed=5),
         "holed": lambda: Graph.holed_grid(40, 50, 6, 3), "star": lambda: Graph.star(500),
         "path": lambda: Graph.path(700), "cycle": lambda: Graph.cycle(501)}[gen]()
    A = csr_matrix(g)
    for src in (0, g.n // 3, g.n - 1):
        want = shortest_path(A, method="D", unweighted=True, indices=src)
        want = np.where(np.isinf(want), -1, want).astype(np.int64)
        assert np.array_equal(O.seq_bfs(g, src).astype(np.int64), want)
        got = O.dobfs(g, src)  # unreached left at -1 by the caller's fill
        assert np.array_equal(got.astype(np.int64), want)
        # direction switches forced either way give the same levels
        assert np.array_equal(O.dobfs(g, src, alpha=1, beta=1).astype(np.int64), want)
        assert np.array_equal(O.dobfs(g, src, alpha=10**9, beta=18).astype(np.int64), want)


@pytest.mark.parametrize("delta", [1, 7, 300, 10000])
def test_deltastep_matches_dijkstra(delta):
    g = Graph.grid3d(8, 9, 10, seed=4, wmax=255)
    A = csr_matrix(g, weighted=True)
    for src in (0, 333, g.n - 1):
        want = shortest_path(A, method="D", indices=src).astype(np.int64)
        got = O.deltastep(g, src, delta).astype(np.int64)
        assert np.array_equal(got, want)
        assert np.array_equal(O.dijkstra(g, src), want)


def test_deltastep_unreached_value():
    u = np.array([0, 2]); v = np.array([1, 3]); w = np.array([5, 6])
    g = Graph.from_edges(4, u, v, w)
    d = O.deltastep(g, 0, 10)
    assert d.tolist() == [0, 5, 2**30 - 1, 2**30 - 1]


# ------------------------------------------------------------------ k-centers

def brute_kcenters(g, first, count):
    """Farthest-first traversal with lowest-index ties, from a dense matrix."""
    D = shortest_path(csr_matrix(g), unweighted=True)
    md = np.full(g.n, np.inf)
    src, out = first, []
    for _ in range(count):
        out.append(src)
        md = np.minimum(md, D[src])
        src = int(np.argmax(md))  # numpy argmax returns the first maximiser
    return out


def test_kcenters_golden_mesh128(mesh128):
    gold = json.loads((GOLDEN / "mesh128_kcenters.json").read_text())
    r = O.parhde(mesh128, 10, 2, lazy=False)
    assert r["rc"] == 0
    assert r["pivots"][:12] == gold["pivots"]
    D = [O.seq_bfs(mesh128, p).max() for p in gold["pivots"]]
    assert D == gold["eccentricities"]
    assert r["accepted"][gold["discarded_candidate"] - 1] == 0
    assert r["normdist"][gold["discarded_candidate"] - 1] < 1e-9
    assert r["eig_iters"] == gold["power_iterations"]["10"]


@pytest.mark.parametrize("gen", ["rgg", "rmat", "holed"])
def test_kcenters_brute_force(gen):
    g = {"rgg": lambda: Graph.rgg(1500, 9.0, seed=3), "rmat": lambda: Graph.rmat(10, 8, seed=9),
         "holed": lambda: Graph.holed_grid(30, 30, 4, 8)}[gen]()
    r = O.parhde(g, 8, 2, lazy=False)
    assert r["rc"] == 0
    assert r["pivots"] == brute_kcenters(g, 0, len(r["pivots"]))


def test_lazy_equals_eager(mesh128):
    a = O.parhde(mesh128, 10, 2, lazy=True)
    b = O.parhde(mesh128, 10, 2, lazy=False)
    assert a["pivots"] == b["pivots"][: len(a["pivots"])]
    assert np.array_equal(a["coords"], b["coords"])
    assert a["n_bfs"] == 11 and b["n_bfs"] == 30


# ------------------------------------------------------------------ D-orthogonalisation

def _laplacian(g):
    A = csr_matrix(g, weighted=True)
    d = np.asarray(A.sum(axis=1)).ravel()
    return sp.diags(d) - A, d


@pytest.mark.parametrize("s", [10, 25])
def test_cgs_is_the_d_projection(mesh128, s):
    """Each accepted column is the normalised residual of its candidate after
    the D-orthogonal projection onto span(1, earlier columns): a closed form
    CGS with D-orthogonal earlier columns must reproduce."""
    r = O.parhde(mesh128, s, 2, lazy=True, want_S=True)
    S = r["S"]
    _, d = _laplacian(mesh128)
    acc = [i for i, a in enumerate(r["accepted"]) if a == 1]
    assert len(acc) == s
    B = np.ones((mesh128.n, 1)) / np.sqrt(mesh128.n)
    for j, cand in enumerate(acc):
        a = O.seq_bfs(mesh128, r["pivots"][cand]).astype(np.float64)
        a /= np.linalg.norm(a)
        G = B.T @ (d[:, None] * B)
        res = a - B @ np.linalg.solve(G, B.T @ (d * a))
        res /= np.linalg.norm(res)
        assert np.max(np.abs(res - S[:, j])) < 1e-9, j
        B = np.hstack([B, S[:, j:j + 1]])
    # invariants: S^T D 1 = 0, S^T D S diagonal, unit columns
    StD1 = S.T @ d  # CGS1 loses orthogonality slowly: bound relative to ||d||
    assert np.max(np.abs(StD1)) < 1e-10 * np.linalg.norm(d)
    M = S.T @ (d[:, None] * S)
    off = M - np.diag(np.diag(M))
    assert np.max(np.abs(off)) < 1e-8 * np.max(np.diag(M))
    assert np.allclose(np.linalg.norm(S, axis=0), 1.0, atol=1e-12)


def test_sls_is_dense_algebra(mesh128):
    s = 10
    r = O.parhde(mesh128, s, 2, lazy=True, want_S=True)
    L, _ = _laplacian(mesh128)
    S = r["S"]
    M = S.T @ (L @ S)  # M[i,k] = S_i^T L S_k
    C = r["SLS"]  # reference storage C[k + i*s] = S_i^T (LS)_k
    assert np.allclose(C.reshape(s, s), M, rtol=0, atol=1e-12 * np.abs(M).max())


def test_coordinates_are_S_times_V(mesh128):
    s = 10
    r = O.parhde(mesh128, s, 2, lazy=True, want_S=True)
    X = r["S"] @ r["eigvec"]
    assert np.allclose(X, r["coords_raw"], atol=1e-14)
    assert np.allclose(r["coords"], X / np.linalg.norm(X, axis=0), atol=1e-14)


# ------------------------------------------------------------------ eigensolver

@pytest.mark.parametrize("name", ["mesh128_s10", "mesh128_s50", "spd3", "spd8", "spd20", "spd64"])
def test_power_method_matches_reference_fixture(name):
    f = np.load(GOLDEN / f"power_eig_{name}.npz")
    K, ev, it, rc = O.power_eig(f["K_in"])
    assert rc == 0 and int(f["rc"]) == 0
    assert np.array_equal(K, f["K_out"])
    assert np.array_equal(ev, f["evals"])


def test_power_method_live_reference():
    if O.ref_lib() is None:
        pytest.skip("reference tree not present: fixtures cover this")
    rng = np.random.default_rng(3)
    for dim in (2, 5, 17, 33, 50):
        A = rng.standard_normal((dim, dim))
        A = A @ A.T + np.diag(np.arange(1, dim + 1) * 3.0)
        Kin = np.ascontiguousarray(A.T.reshape(-1))
        a = O.power_eig(Kin)
        b = O.ref_power_eig(Kin)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_power_method_approaches_exact_spectrum():
    rng = np.random.default_rng(5)
    Q, _ = np.linalg.qr(rng.standard_normal((6, 6)))
    lam = np.array([1.0, 2.0, 4.0, 8.0, 16.0, 32.0])
    A = Q @ np.diag(lam) @ Q.T
    K, ev, it, rc = O.power_eig(np.ascontiguousarray(A.T.reshape(-1)))
    assert rc == 0
    assert np.allclose(ev, lam, rtol=2e-3)  # ascending after the reversal


# ------------------------------------------------------------------ other algorithms

def test_pivotmds_matches_numpy(mesh128):
    s, c = 8, 2
    r = O.pivotmds(mesh128, s, c, want_S=True)
    assert r["rc"] == 0
    D = np.stack([O.seq_bfs(mesh128, p) for p in r["pivots"]], axis=1).astype(np.float64)
    D2 = D * D
    C = -0.5 * (D2 - D2.mean(axis=0)[None, :] - D2.mean(axis=1)[:, None] + D2.mean())
    assert np.allclose(C, r["S"], atol=1e-9 * np.abs(C).max())
    ev = np.linalg.eigvalsh(C.T @ C)
    # the power method stops at min |cos| >= 0.999: estimates are loose
    assert np.allclose(np.sort(r["eigval"]), ev, rtol=1e-2, atol=1e-6 * ev.max())


def test_phde_column_centering(mesh128):
    s = 6
    r = O.phde(mesh128, s, 2, want_S=True)
    D = np.stack([O.seq_bfs(mesh128, p) for p in r["pivots"]], axis=1).astype(np.float64)
    assert np.allclose(r["S"], D - D.mean(axis=0), atol=1e-9)


def test_wparhde_invariants():
    g = Graph.grid3d(12, 11, 10, seed=4, wmax=255)
    s = 8
    r = O.wparhde(g, s, 3, delta=500, want_S=True)
    assert r["rc"] == 0
    S = r["S"]
    _, d = _laplacian(g)
    assert np.max(np.abs(S.T @ d)) < 1e-8
    M = S.T @ (d[:, None] * S)
    assert np.max(np.abs(M - np.diag(np.diag(M)))) < 1e-7 * np.diag(M).max()
    # pivots are farthest-first under the weighted metric
    A = csr_matrix(g, weighted=True)
    Dm = shortest_path(A, method="D", indices=r["pivots"])
    md = np.full(g.n, np.inf)
    for k, p in enumerate(r["pivots"][:-1]):
        md = np.minimum(md, Dm[k])
        assert r["pivots"][k + 1] == int(np.argmax(md))
    # delta-independent
    r2 = O.wparhde(g, s, 3, delta=10000)
    assert r2["pivots"] == r["pivots"]


def test_hall_energy_closed_form():
    n = 50
    g = Graph.path(n)
    y = np.arange(n, dtype=np.float64)
    deg = np.full(n, 2.0)
    deg[0] = deg[-1] = 1
    assert np.isclose(O.hall_energy(g, y), (n - 1) / np.sum(deg * y * y), rtol=1e-14)
    g = Graph.cycle(64)
    t = 2 * np.pi * np.arange(64) / 64
    Y = np.stack([np.cos(t), np.sin(t)], axis=1)
    # cycle eigenvectors: energy = (1 - cos(2 pi / n)) exactly
    assert np.isclose(O.hall_energy(g, Y), 1 - np.cos(2 * np.pi / 64), rtol=1e-12)


def test_lcc_matches_scipy():
    g = Graph.rgg(4000, 4.0, seed=9, lcc=False)
    ncomp, lab = connected_components(csr_matrix(g), directed=False)
    sizes = np.bincount(lab)
    keep = np.flatnonzero(lab == np.argmax(sizes))  # argmax: first component on ties
    h = g.lcc()
    assert h.n == len(keep)
    sub = csr_matrix(g)[keep][:, keep]
    assert np.array_equal(h.rowptr.astype(np.int64), sub.indptr)
    assert np.array_equal(h.adj.astype(np.int64), sub.indices)


def _mds_sequential(g, pivots, dbl):
    """The reference's PivotMDS / PHDE arithmetic with every product and sum
    rounded on its own and every sum a sequential chain (np.cumsum runs in
    index order): doubleCentering (PivotMDS.cpp:30-66) or columnCentering
    (PHDE.cpp:30-47), then C^T C with j ascending (DrawingBase.cpp:221-244)."""
    D = np.stack([O.seq_bfs(g, p) if g.w is None else O.dijkstra(g, p) for p in pivots],
                 axis=1).astype(np.float64)
    n, s = D.shape
    if dbl:
        sq = D * D
        colsum = np.array([np.cumsum(sq[:, i])[-1] for i in range(s)])
        rowsum = np.zeros(n)
        for i in range(s):  # row j accumulates column by column
            rowsum = rowsum + sq[:, i]
        total = 0.0
        for i in range(s):
            total += colsum[i]
        rN, rK = 1.0 / n, 1.0 / s
        C = -0.5 * (((sq - rN * colsum[None, :]) - (rK * rowsum)[:, None]) + (rK * rN) * total)
    else:
        C = D - np.array([np.cumsum(D[:, i])[-1] for i in range(s)])[None, :] / n
    K = np.empty((s, s))
    for i in range(s):
        for k in range(s):
            K[i, k] = np.cumsum(C[:, i] * C[:, k])[-1]
    return D, C, K


@pytest.mark.parametrize("algo", ["pivotmds", "phde"])
@pytest.mark.parametrize("weighted", [False, True])
def test_mds_matrix_bit_exact(algo, weighted):
    """The s x s matrix the power method sees is bit-exact against the
    sequential restatement, also where squared distances pass 2^53 and the
    column sums round (weights up to 2^22)."""
    g = Graph.grid3d(12, 11, 10, seed=8, wmax=1 << 22) if weighted else Graph.holed_grid(90, 80, 6, 12)
    s = 10 if weighted else 16
    fn = O.pivotmds if algo == "pivotmds" else O.phde
    r = fn(g, s, 2, want_S=True)
    assert r["rc"] == 0
    D, C, K = _mds_sequential(g, r["pivots"], algo == "pivotmds")
    assert np.array_equal(r["S"], C)
    assert np.array_equal(r["SLS"].reshape(s, s), K)
    if weighted:  # the case where the order of the roundings matters
        assert float(np.max(np.sum(D * D, axis=0))) > 2.0 ** 53
