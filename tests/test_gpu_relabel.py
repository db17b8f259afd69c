"""The context's internal vertex numbering (csrc/relabel.hip): with the
renumbering forced on small graphs, every entry point must still answer in
the caller's numbering — bit-exact levels, distances, pivots and selections
against the oracle and against the same context without renumbering, and
coordinates within the FP64 tolerance."""
import os
from contextlib import contextmanager

import numpy as np
import pytest

from icpp_2020_parhde_amd import layout as L
from icpp_2020_parhde_amd.graphs import Graph
from oracle import learned as OL
from oracle import pyoracle as O

from test_gpu_core import aligned_rel_err, FP32_TOL, FP64_TOL

pytestmark = pytest.mark.gpu
INF = 2**30 - 1


@contextmanager
def context(g, relabel):
    old = os.environ.get("PHDE_RELABEL")
    os.environ["PHDE_RELABEL"] = "1" if relabel else "0"
    try:
        ctx = L.Context(g)
    finally:
        if old is None:
            del os.environ["PHDE_RELABEL"]
        else:
            os.environ["PHDE_RELABEL"] = old
    with ctx:
        _, r = ctx.vertex_order()
        assert r == relabel
        yield ctx


def two_components():
    a = Graph.mesh(20, 30)
    rows = np.repeat(np.arange(a.n), np.diff(a.rowptr.astype(np.int64)))
    m = rows < a.adj
    u, v = rows[m], a.adj[m].astype(np.int64)
    return Graph.from_edges(2 * a.n + 3, np.concatenate([u, u + a.n + 3]), np.concatenate([v, v + a.n + 3]))


GRAPHS = {
    "rgg": lambda: Graph.rgg(20000, 10.0, seed=21),
    "holed": lambda: Graph.holed_grid(90, 80, 6, 12),
    "rmat": lambda: Graph.rmat(12, 16, seed=13),
    "two": two_components,
}


def test_order_is_the_level_sort(gpu):
    g = GRAPHS["rgg"]()
    d1 = O.seq_bfs(g, 0)
    d2 = O.seq_bfs(g, int(np.argmax(d1)))
    with context(g, True) as ctx:
        order, _ = ctx.vertex_order()
    assert np.array_equal(order, np.lexsort((np.arange(g.n), d2, d1)))
    with context(g, False) as ctx:
        order, _ = ctx.vertex_order()
    assert np.array_equal(order, np.arange(g.n))


def test_auto_rule_keeps_small_graphs(gpu):
    g = GRAPHS["rgg"]()
    old = os.environ.pop("PHDE_RELABEL", None)
    try:
        with L.Context(g) as ctx:  # n < 65,536: never renumbered by default
            assert ctx.vertex_order()[1] is False
    finally:
        if old is not None:
            os.environ["PHDE_RELABEL"] = old


@pytest.mark.parametrize("name", list(GRAPHS))
def test_bfs_and_msbfs(gpu, name):
    g = GRAPHS[name]()
    rng = np.random.default_rng(3)
    src = [0, g.n - 1] + [int(x) for x in rng.choice(g.n, 5, replace=False)]
    with context(g, True) as ctx:
        for s in src:
            want = O.seq_bfs(g, s)
            got = ctx.bfs(s).astype(np.int64)
            want = np.where(want < 0, -1, want)
            assert np.array_equal(got, want), (name, s)
        D = ctx.msbfs(src)
    for j, s in enumerate(src):
        want = O.seq_bfs(g, s)
        assert np.array_equal(D[:, j].astype(np.int64), np.where(want < 0, -1, want)), (name, j)


def test_sssp(gpu):
    g = Graph.grid3d(20, 18, 16, seed=4, wmax=255)
    with context(g, True) as ctx:
        for s in (0, 77, g.n - 1):
            assert np.array_equal(ctx.sssp(s).astype(np.int64), O.dijkstra(g, s)), s
    gu = Graph.from_edges(5, [0, 2], [1, 3], [7, 9])
    with context(gu, True) as ctx:
        assert ctx.sssp(0, 4).tolist() == [0, 7, INF, INF, INF]
        assert ctx.sssp(3, 4).tolist() == [INF, INF, 9, 0, INF]


@pytest.mark.parametrize("name,s", [("rgg", 30), ("holed", 20)])
def test_parhde_layout(gpu, name, s):
    g = GRAPHS[name]()
    ref = O.parhde(g, s, 2, lazy=True)
    for prec, tol in ((L.PREC_FP64, FP64_TOL), (L.PREC_FP32, FP32_TOL)):
        with context(g, True) as ctx:
            X, st = ctx.layout(s=s, num_coord=2, precision=prec)
        assert st.pivots == ref["pivots"]
        assert aligned_rel_err(X, ref["coords"]) < tol, prec


def test_wparhde_pivotmds_phde(gpu):
    g = Graph.grid3d(16, 14, 12, seed=5, wmax=200)
    ref = O.wparhde(g, 10, 3)
    with context(g, True) as ctx:
        X, st = ctx.layout(s=10, num_coord=3, algorithm=L.ALGO_WPARHDE, delta=0)
    assert st.pivots == ref["pivots"]
    assert aligned_rel_err(X, ref["coords"]) < FP64_TOL
    # PivotMDS / PHDE through the Jacobi eigensolver: renumbered and not agree
    h = GRAPHS["holed"]()
    for algo, fn in ((L.ALGO_PIVOTMDS, O.pivotmds), (L.ALGO_PHDE, O.phde)):
        ref = fn(h, 12, 2)
        out = {}
        for r in (False, True):
            with context(h, r) as ctx:
                out[r] = ctx.layout(s=12, num_coord=2, algorithm=algo, eig_mode=L.EIG_JACOBI)
        assert out[True][1].pivots == out[False][1].pivots == ref["pivots"], algo
        assert aligned_rel_err(out[True][0], out[False][0]) < FP64_TOL, algo


@pytest.mark.parametrize("algo", [L.ALGO_PIVOTMDS, L.ALGO_PHDE])
def test_mds_power_method_renumbered(gpu, algo):
    """The reference's power method stops at |<e_i,t_i>| >= 0.999 and C^T C
    has a near-null direction, so PivotMDS / PHDE coordinates follow the
    rounding of C^T C: the library forms it in the caller's row order with
    the reference's sequential sums, and matches the oracle on renumbered
    graphs (a forced one and one the context renumbers by its own rule)."""
    fn = O.pivotmds if algo == L.ALGO_PIVOTMDS else O.phde
    cases = [(GRAPHS["holed"](), True, 16), (Graph.rgg(100000, 10.0, seed=31), None, 20)]
    for g, force, s in cases:
        ref = fn(g, s, 2)
        assert ref["rc"] == 0
        if force:
            with context(g, True) as ctx:
                X, st = ctx.layout(s=s, num_coord=2, algorithm=algo)
        else:
            with L.Context(g) as ctx:
                assert g.n >= 65536 and ctx.vertex_order()[1]  # renumbered by the context's rule
                X, st = ctx.layout(s=s, num_coord=2, algorithm=algo)
        assert st.pivots == ref["pivots"], algo
        assert st.eig_iters == ref["eig_iters"], (algo, st.eig_iters, ref["eig_iters"])
        err = aligned_rel_err(X, ref["coords"])
        print(f"algo {algo} n={g.n} s={s}: coordinate rel err {err:.2e}")
        assert err < FP64_TOL, (algo, g.n, err)


def test_mds_weighted_large_distances(gpu):
    """Weighted PivotMDS whose squared distances pass 2^53: the column sums
    of squares round, and the library then sums them in the reference's
    sequential order."""
    g = Graph.grid3d(12, 11, 10, seed=8, wmax=1 << 22)
    ref = O.pivotmds(g, 10, 2)
    with context(g, True) as ctx:
        X, st = ctx.layout(s=10, num_coord=2, algorithm=L.ALGO_PIVOTMDS)
    assert st.pivots == ref["pivots"]
    assert aligned_rel_err(X, ref["coords"]) < FP64_TOL


@pytest.mark.parametrize("strategy", [L.PIVOTS_RANDFINE, L.PIVOTS_RANDCOARSE])
def test_random_pivots(gpu, strategy):
    g = GRAPHS["rgg"]()
    out = {}
    for r in (False, True):
        with context(g, r) as ctx:
            out[r] = ctx.layout(s=20, num_coord=2, strategy=strategy, seed=9)
    (X0, s0), (X1, s1) = out[False], out[True]
    assert s1.pivots == s0.pivots
    assert aligned_rel_err(X1, X0) < FP64_TOL


def test_kcenters_and_dedupe(gpu):
    g = GRAPHS["holed"]()
    ref = {}
    for r in (False, True):
        with context(g, r) as ctx:
            ref[r] = ctx.kcenters(25, first=7)
    assert np.array_equal(ref[True][0], ref[False][0])
    assert np.array_equal(ref[True][1], ref[False][1])
    # more sources than distinct argmaxes: the dedupe branch picks the last
    # vertex (caller's numbering) at the maximum distance
    p = Graph.path(7)
    with context(p, False) as ctx:
        a = ctx.kcenters(9)
    with context(p, True) as ctx:
        b = ctx.kcenters(9)
    assert a[0].tolist() == b[0].tolist() and np.array_equal(a[1], b[1])


def test_learned_entry_points(gpu):
    g = GRAPHS["rgg"]()
    rng = np.random.default_rng(4)
    params = (rng.standard_normal(OL.NPARAMS) * 0.3).astype(np.float32)
    target = rng.random(g.n).astype(np.float32)
    res = {}
    with L.PivotNet(seed=1) as net:
        net.params = params
        for r in (False, True):
            with context(g, r) as ctx:
                x, srcs = ctx.pivot_features(5)
                ctx.set_pivot_net(net)
                y = net.scores(ctx)
                picks = ctx.select_pivots(y, 40, 3, first=5)
                ctx.set_training_data(x, target)
                loss, grad = net.gradient(ctx)
                X, st = ctx.layout(s=20, num_coord=2, strategy=L.PIVOTS_LEARNED)
                e = ctx.hall_energy(X)
                res[r] = (x, srcs, y, picks, loss, grad, X, st, e)
    a, b = res[False], res[True]
    assert np.array_equal(a[0], b[0]) and a[1] == b[1]  # features: same numbers
    assert np.array_equal(a[2], b[2])  # scores: per-row sums in the same order
    assert a[3].tolist() == b[3].tolist()  # selection: ties on the caller's ids
    assert b[4] == pytest.approx(a[4], rel=1e-6)
    assert np.allclose(b[5], a[5], rtol=1e-4, atol=1e-6 * np.abs(a[5]).max())
    assert b[7].pivots == a[7].pivots
    assert aligned_rel_err(b[6], a[6]) < FP64_TOL
    assert b[8] == pytest.approx(a[8], rel=1e-12)
    assert b[7].hall_energy == pytest.approx(a[7].hall_energy, rel=1e-12)


@pytest.mark.parametrize("name,s", [("path7", 7), ("cycle8", 5), ("mesh3x3", 8), ("star9", 10)])
def test_dedupe_branch_matches_oracle(gpu, name, s):
    """More traversals than vertices: every vertex becomes a source and the
    dedupe rule (ParHDE.cpp:142-155) repeats; the pivot sequence is the
    oracle's (or_parhde), and the layout stops with the reference's
    too-few-pivots error (ParHDE.cpp:281-285)."""
    g = {"path7": lambda: Graph.path(7), "cycle8": lambda: Graph.cycle(8),
         "mesh3x3": lambda: Graph.mesh(3, 3), "star9": lambda: Graph.star(9)}[name]()
    ref = O.parhde(g, s, 2)
    assert ref["n_bfs"] > g.n and ref["rc"] != 0
    for r in (False, True):
        with context(g, r) as ctx:
            src, _ = ctx.kcenters(ref["n_bfs"])
            assert src.tolist() == ref["pivots"]
            with pytest.raises(L.PhdeError) as e:
                ctx.layout(s=s, num_coord=2)
            assert e.value.code == -4
