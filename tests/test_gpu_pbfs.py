"""The partitioned traversal (csrc/pbfs.hip) of unweighted graphs with a
deep BFS: exact levels against the oracle's textbook BFS from several
sources, the k-centers chain and its distance field identical to the same
context without it (PHDE_PBFS=0), and layouts against the oracle.  Small
graphs are renumbered on request (PHDE_RELABEL=1, which also marks them
deep) and cut with small part budgets (PHDE_PBFS_BYTES), so that a few
thousand vertices make dozens to hundreds of parts and most levels cross a
part boundary."""
import os
from contextlib import contextmanager

import numpy as np
import pytest

from icpp_2020_parhde_amd import layout as L
from icpp_2020_parhde_amd.graphs import Graph
from oracle import pyoracle as O

from test_gpu_core import FP64_TOL, aligned_rel_err

pytestmark = pytest.mark.gpu


@contextmanager
def context(g, part_bytes=None, pbfs=True, obfs=True, din=True):
    env = {"PHDE_RELABEL": "1", "PHDE_PBFS": "1" if pbfs else "0", "PHDE_OBFS": "1" if obfs else "0",
           "PHDE_OBFS_DIN": "1" if din else "0"}
    if part_bytes:
        env["PHDE_PBFS_BYTES"] = str(part_bytes)
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        ctx = L.Context(g)
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v
    with ctx:
        yield ctx


def two_paths(n):
    i = np.arange(n - 1)
    return Graph.from_edges(2 * n, np.concatenate([i, i + n]), np.concatenate([i + 1, i + n + 1]))


GRAPHS = {
    "holed": lambda: Graph.holed_grid(120, 90, 10, 77),
    "rgg": lambda: Graph.rgg(20000, 10.0, seed=2),
    "mesh": lambda: Graph.mesh(96, 80),
    "path": lambda: Graph.path(3000),
    "two_paths": lambda: two_paths(400),
}
BUDGETS = [None, 4096, 600]


# boundary-distance method (step 1 from the in-part distance tables, or by
# a BFS inside the source's part), or rounds in LDS at every run
MODES = ["obfs", "obfs_bfs1", "pbfs"]


def check_mode(ctx, mode):
    parts = ctx.num_parts()
    assert parts >= 1
    if mode != "pbfs":
        assert ctx.num_boundary() >= 0
    else:
        assert ctx.num_boundary() == -1
    return parts


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("budget", BUDGETS)
@pytest.mark.parametrize("name", list(GRAPHS))
def test_levels_exact(gpu, name, budget, mode):
    g = GRAPHS[name]()
    rng = np.random.default_rng(7)
    srcs = [0, g.n - 1, g.n // 2] + [int(x) for x in rng.choice(g.n, 4, replace=False)]
    with context(g, budget, obfs=mode != "pbfs", din=mode != "obfs_bfs1") as ctx:
        parts = check_mode(ctx, mode)
        for s in srcs:
            want = O.seq_bfs(g, s)
            assert np.array_equal(ctx.bfs(s), want), (name, budget, parts, s)
        nbd = ctx.num_boundary()
    print(f"{name} budget {budget} {mode}: {parts} parts, {nbd} boundary vertices, {len(srcs)} sources exact")


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("budget", BUDGETS)
def test_kcenters_same_as_without(gpu, budget, mode):
    g = GRAPHS["holed"]()
    with context(g, budget, obfs=mode != "pbfs", din=mode != "obfs_bfs1") as ctx:
        check_mode(ctx, mode)
        a = ctx.kcenters(30, first=11)
    with context(g, pbfs=False) as ctx:
        assert ctx.num_parts() == 0
        b = ctx.kcenters(30, first=11)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("budget", [None, 4096])
@pytest.mark.parametrize("name,s", [("holed", 30), ("rgg", 50), ("mesh", 20)])
def test_layout_parity(gpu, name, s, budget, mode):
    g = GRAPHS[name]()
    ref = O.parhde(g, s, 2, lazy=True)
    with context(g, budget, obfs=mode != "pbfs", din=mode != "obfs_bfs1") as ctx:
        check_mode(ctx, mode)
        X, st = ctx.layout(s=s, num_coord=2)
        Xe, ste = ctx.layout(s=s, num_coord=2, eager=1)
    assert st.pivots == ref["pivots"]
    assert aligned_rel_err(X, ref["coords"]) < FP64_TOL
    assert np.array_equal(X, Xe) and ste.n_traversals == 3 * s


@pytest.mark.parametrize("mode", MODES)
def test_unreached_component(gpu, mode):
    g = two_paths(400)
    with context(g, 600, obfs=mode != "pbfs", din=mode != "obfs_bfs1") as ctx:
        check_mode(ctx, mode)
        d = ctx.bfs(3)
        assert d[400:].tolist() == [-1] * 400 and d[:400].tolist() == [abs(i - 3) for i in range(400)]
        with pytest.raises(L.PhdeError) as e:
            ctx.layout(s=4, num_coord=2)
        assert e.value.code == -7


@pytest.mark.parametrize("budget", BUDGETS)
@pytest.mark.parametrize("name", list(GRAPHS))
@pytest.mark.parametrize("k", [1, 7, 64])
def test_multi_source_columns_exact(gpu, name, budget, k):
    # the k searches of one multi-source call run together by the
    # boundary-distance method: every column equals its own textbook BFS,
    # with repeated sources and sources in one part
    g = GRAPHS[name]()
    rng = np.random.default_rng(100 + k)
    src = [int(x) for x in rng.choice(g.n, size=k, replace=k > g.n)]
    if k >= 7:
        src[1] = src[0]          # a repeated source
        src[2] = (src[0] + 1) % g.n  # a neighbour in the id order, most often in the same part
    with context(g, budget) as ctx:
        check_mode(ctx, "obfs")
        D = ctx.msbfs(src)
    for j, s in enumerate(src):
        assert np.array_equal(D[:, j], O.seq_bfs(g, s)), (name, budget, k, j)
