"""Parity of the HIP path (through the C ABI) with the oracle and the
reference's golden vectors.  Bit-exact for traversals, pivots and the power
method; FP64 layouts within 1e-8 relative of the oracle after sign
alignment."""
import json
from pathlib import Path

import numpy as np
import pytest

from icpp_2020_parhde_amd import layout as L
from icpp_2020_parhde_amd.graphs import Graph
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"
FP64_TOL = 1e-8  # relative, after eigenvector sign alignment
FP32_TOL = 1e-5  # FP32 S and LS, FP64 accumulations and eigensolve, against the FP64 oracle


def aligned_rel_err(X, Y):
    errs = []
    for k in range(Y.shape[1]):
        x, y = X[:, k], Y[:, k]
        if np.dot(x, y) < 0:
            x = -x
        errs.append(np.max(np.abs(x - y)) / np.max(np.abs(y)))
    return max(errs)


def bfs_certificate(g, src, d):
    """A level function is the BFS one iff d[src] = 0, |d_u - d_v| <= 1 on
    every edge and every other reached vertex has a neighbour one level down."""
    d = d.astype(np.int64)
    rows = np.repeat(np.arange(g.n), np.diff(g.rowptr.astype(np.int64)))
    cols = g.adj.astype(np.int64)
    du, dv = d[rows], d[cols]
    ok = d[src] == 0
    reached = (du >= 0) & (dv >= 0)
    ok &= bool(np.all((du >= 0) == (dv >= 0)))  # components are closed
    ok &= bool(np.all(np.abs(du[reached] - dv[reached]) <= 1))
    parent = np.zeros(g.n, bool)
    np.logical_or.at(parent, rows[reached & (dv == du - 1)], True)
    need = (d > 0)
    ok &= bool(np.all(parent[need]))
    return ok


GRAPHS = {
    "mesh128": lambda: Graph.mesh(128, 128),
    "grid3d": lambda: Graph.grid3d(20, 18, 16, wmax=0),
    "rgg20k": lambda: Graph.rgg(20000, 10.0, seed=2),
    "rmat14": lambda: Graph.rmat(14, 16, seed=3),
    "holed": lambda: Graph.holed_grid(120, 90, 10, 77),
    "path": lambda: Graph.path(3000),
    "star": lambda: Graph.star(20000),
}


@pytest.fixture(scope="module")
def graphs():
    return {k: f() for k, f in GRAPHS.items()}


@pytest.mark.parametrize("name", list(GRAPHS))
def test_bfs_bit_exact(gpu, graphs, name):
    g = graphs[name]
    with L.Context(g) as ctx:
        for src in (0, g.n // 2, g.n - 1):
            want = O.seq_bfs(g, src)
            assert np.array_equal(ctx.bfs(src), want), (name, src)
            # forced bottom-up heavy and top-down only schedules
            assert np.array_equal(ctx.bfs(src, alpha=1, beta=1), want), (name, src, "bu")
            assert np.array_equal(ctx.bfs(src, alpha=1 << 30, beta=18), want), (name, src, "td")


def test_bfs_disconnected_and_tiny(gpu):
    g = Graph.from_edges(7, [0, 1, 3, 4], [1, 2, 4, 5])
    with L.Context(g) as ctx:
        assert ctx.bfs(0).tolist() == [0, 1, 2, -1, -1, -1, -1]
        assert ctx.bfs(4).tolist() == [-1, -1, -1, 1, 0, 1, -1]
        assert ctx.bfs(6).tolist() == [-1] * 6 + [0]


def test_bfs_hub_expansion(gpu):
    """Degree above the grid-wide expansion threshold (8192)."""
    g = Graph.star(50000)
    with L.Context(g) as ctx:
        d = ctx.bfs(0)
        assert d[0] == 0 and np.all(d[1:] == 1)
        d = ctx.bfs(17)
        assert d[17] == 0 and d[0] == 1 and np.all(np.delete(d, [0, 17]) == 2)


@pytest.mark.parametrize("name", ["rgg1m", "rmat18"])
def test_bfs_certificate_large(gpu, name):
    g = Graph.rgg(1 << 20, 10.0, seed=2) if name == "rgg1m" else Graph.rmat(18, 16, seed=3)
    with L.Context(g) as ctx:
        for src in (0, g.n - 1):
            assert bfs_certificate(g, src, ctx.bfs(src)), (name, src)


# ------------------------------------------------------------------ eigensolvers

@pytest.mark.parametrize("name", ["mesh128_s10", "mesh128_s50", "spd3", "spd8", "spd20", "spd64"])
def test_power_method_bit_exact_vs_reference(gpu, name):
    f = np.load(GOLDEN / f"power_eig_{name}.npz")
    dim = int(round(len(f["K_in"]) ** 0.5))
    K = f["K_in"].reshape(dim, dim).T  # column-major buffer -> matrix
    with L.Context(Graph.path(4)) as ctx:
        Kout, ev, it = ctx.power_eig(K)
    assert np.array_equal(Kout, f["K_out"]), name
    assert np.array_equal(ev, f["evals"]), name


@pytest.mark.parametrize("dim", [2, 7, 33, 64])
def test_jacobi_matches_eigh(gpu, dim):
    rng = np.random.default_rng(dim)
    A = rng.standard_normal((dim, dim))
    A = A + A.T
    with L.Context(Graph.path(4)) as ctx:
        ev, V = ctx.sym_eig(A)
        want = np.linalg.eigvalsh(A)
        assert np.allclose(ev, want, atol=1e-10 * np.abs(want).max())
        assert np.allclose(A @ V, V * ev, atol=1e-9 * np.abs(want).max())
        # generalised problem
        B = rng.standard_normal((dim, dim))
        B = B @ B.T + dim * np.eye(dim)
        ev, V = ctx.sym_eig(A, B)
        import scipy.linalg
        want = scipy.linalg.eigh(A, B, eigvals_only=True)
        assert np.allclose(ev, want, atol=1e-9 * np.abs(want).max())
        assert np.allclose(V.T @ B @ V, np.eye(dim), atol=1e-8)


# ------------------------------------------------------------------ layouts

def test_layout_mesh128_golden(gpu):
    gold = json.loads((GOLDEN / "mesh128_kcenters.json").read_text())
    g = Graph.mesh(128, 128)
    ref = O.parhde(g, 10, 2, lazy=True)
    with L.Context(g) as ctx:
        X, st = ctx.layout(s=10, num_coord=2)
        assert st.pivots == gold["pivots"][: st.n_traversals]
        assert st.n_traversals == 11 and st.n_discarded == 1 and st.n_accepted == 10
        assert st.eig_iters == gold["power_iterations"]["10"]
        assert aligned_rel_err(X, ref["coords"]) < FP64_TOL
        Xe, ste = ctx.layout(s=10, num_coord=2, eager=1)
        assert ste.n_traversals == 30 and ste.pivots[:12] == gold["pivots"]
        assert np.array_equal(Xe, X)  # eager and lazy: identical outputs


@pytest.mark.parametrize("name,s", [("mesh128", 50), ("rgg20k", 50), ("rmat14", 20), ("rmat14", 50),
                                    ("holed", 30)])
def test_layout_fp64_parity(gpu, graphs, name, s):
    g = graphs[name]
    ref = O.parhde(g, s, 2, lazy=True)
    with L.Context(g) as ctx:
        X, st = ctx.layout(s=s, num_coord=2)
    assert st.pivots == ref["pivots"]
    assert st.n_accepted == ref["n_accepted"] == s
    assert st.eig_iters == ref["eig_iters"]
    err = aligned_rel_err(X, ref["coords"])
    print(f"{name} s={s} fp64 coord rel err {err:.2e}")
    assert err < FP64_TOL


@pytest.mark.parametrize("name,s", [("mesh128", 50), ("rgg20k", 50), ("rmat14", 20), ("rmat14", 50)])
def test_layout_fp32_vs_fp64_oracle(gpu, graphs, name, s):
    g = graphs[name]
    ref = O.parhde(g, s, 2, lazy=True)
    with L.Context(g) as ctx:
        X, st = ctx.layout(s=s, num_coord=2, precision=L.PREC_FP32)
    assert st.pivots == ref["pivots"]
    err = aligned_rel_err(X, ref["coords"])
    print(f"{name} s={s} fp32 coord rel err {err:.2e}")
    assert err < FP32_TOL


@pytest.mark.parametrize("algo", [L.ALGO_PIVOTMDS, L.ALGO_PHDE])
def test_mds_parity(gpu, graphs, algo):
    g = graphs["mesh128"]
    s = 12
    ref = (O.pivotmds if algo == L.ALGO_PIVOTMDS else O.phde)(g, s, 2)
    with L.Context(g) as ctx:
        X, st = ctx.layout(s=s, num_coord=2, algorithm=algo)
    assert st.pivots == ref["pivots"]
    err = aligned_rel_err(X, ref["coords"])
    print(f"algo {algo} coord rel err {err:.2e}")
    assert err < FP64_TOL


def test_layout_errors(gpu):
    with L.Context(Graph.path(6)) as ctx:  # too few independent pivots
        with pytest.raises(L.PhdeError) as e:
            ctx.layout(s=10, num_coord=2)
        assert e.value.code == -4
    g = Graph.from_edges(8, [0, 1, 4, 5], [1, 2, 5, 6])
    with L.Context(g) as ctx:
        with pytest.raises(L.PhdeError) as e:
            ctx.layout(s=2, num_coord=2)
        assert e.value.code == -7
    w = Graph.grid3d(4, 4, 4, wmax=9)
    with L.Context(w) as ctx:
        with pytest.raises(L.PhdeError) as e:
            ctx.layout(s=4, num_coord=2, algorithm=L.ALGO_PARHDE)
        assert e.value.code == -6


def test_hall_energy(gpu):
    g = Graph.cycle(64)
    t = 2 * np.pi * np.arange(64) / 64
    Y = np.stack([np.cos(t), np.sin(t)], axis=1)
    with L.Context(g) as ctx:
        assert np.isclose(ctx.hall_energy(Y), 1 - np.cos(2 * np.pi / 64), rtol=1e-12)
