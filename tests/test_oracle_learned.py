"""Pins of the learned-pivot oracle (oracle/learned.py) on the CPU:
bfloat16 rounding against torch's conversion, the analytic gradient against
central finite differences, the greedy suppression against a brute-force
statement from all-pairs distances, and the given-sources ParHDE oracle
against the k-centers one."""
import numpy as np
import pytest
from scipy.sparse.csgraph import shortest_path

from icpp_2020_parhde_amd.graphs import Graph
from oracle import learned as OL
from oracle import pyoracle as O


def test_bf16_round_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 20000),
                        np.float32([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 0.0, 65504.0])])
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(OL.bf16_round(x), want)
    # ties to even, by hand
    assert OL.bf16_round(np.float32(1.0 + 2 ** -8)) == 1.0
    assert OL.bf16_round(np.float32(1.0 + 3 * 2 ** -8)) == 1.0 + 2 ** -6


def _toy(seed=0):
    g = Graph.rgg(90, 6.0, seed=seed)
    d1 = O.seq_bfs(g, 0)
    d2 = O.seq_bfs(g, int(np.argmax(d1)))
    x = OL.features(g, d1, d2)
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(OL.NPARAMS) * 0.3
    t = rng.random(g.n)
    return g, x, t, p


@pytest.mark.parametrize("seed", [0, 1])
def test_gradient_matches_finite_differences(seed):
    g, x, t, p = _toy(seed)
    loss, grad, _ = OL.loss_and_grad(g, x, t, p, emulate=False)
    rng = np.random.default_rng(100 + seed)
    # a few coordinates of every parameter block
    offs, o = [], 0
    for _, shape in OL.LAYOUT:
        k = int(np.prod(shape))
        offs += list(o + rng.choice(k, size=min(k, 6), replace=False))
        o += k
    eps = 1e-6
    for i in offs:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        lp = np.mean((OL.forward(g, x, pp, emulate=False) - t) ** 2)
        lm = np.mean((OL.forward(g, x, pm, emulate=False) - t) ** 2)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (i, fd, grad[i])
    assert loss == pytest.approx(np.mean((OL.forward(g, x, p, emulate=False) - t) ** 2))


def test_emulated_gradient_close_to_exact():
    # BF16 storage perturbs the gradient by about 2^-8 relative, not more
    g, x, t, p = _toy(3)
    _, ge, _ = OL.loss_and_grad(g, x, t, p, emulate=True)
    _, gx, _ = OL.loss_and_grad(g, x, t, p, emulate=False)
    assert np.linalg.norm(ge - gx) <= 0.05 * np.linalg.norm(gx)


def _brute_select(g, score, first, second, k, r, batches=None):
    """The round-wise selection restated from all-pairs distances."""
    D = shortest_path(OL.adjacency(g)[0], unweighted=True).astype(np.int64)
    r = max(1, min(r, 2046))
    hi = OL.orderable(score) >> np.uint64(11)
    picks, rounds = [first, second], []
    ends = list(batches or [k])
    while len(picks) < k:
        end = next(e for e in ends if e > len(picks))
        sd = np.minimum(D[first], D[second])
        for q in rounds:
            sd = np.where(D[q] <= r, np.minimum(sd, D[q]), sd)
        rnd = []
        while len(rnd) < OL.ROUND and len(picks) < end:
            best = None
            for v in range(g.n):
                if any(max(abs(D[first, v] - D[first, c]), abs(D[second, v] - D[second, c])) <= r for c in rnd):
                    continue
                key = (min(int(sd[v]), r + 1), int(hi[v]), -v)
                if best is None or key > best[0]:
                    best = (key, v)
            if best is None or (rnd and best[0][0] != r + 1):
                break
            rnd.append(best[1])
            picks.append(best[1])
        rounds += rnd
    return np.array(picks)


@pytest.mark.parametrize("name,r", [("mesh", 3), ("rgg", 2), ("holed", 4), ("rmat", 1), ("path", 3)])
def test_greedy_select_matches_brute_force(name, r):
    g = {"mesh": lambda: Graph.mesh(12, 15), "rgg": lambda: Graph.rgg(300, 8.0, seed=4),
         "holed": lambda: Graph.holed_grid(20, 20, 3, 7), "rmat": lambda: Graph.rmat(8, 8, seed=2),
         "path": lambda: Graph.path(40)}[name]()
    rng = np.random.default_rng(5)
    for score in (rng.random(g.n).astype(np.float32), np.zeros(g.n, np.float32)):
        d1 = O.seq_bfs(g, 0)
        second = int(np.argmax(d1))
        d2 = O.seq_bfs(g, second)
        k = min(20, g.n)
        got = OL.greedy_select(g, score, d1, d2, 0, second, k, r)
        assert got.tolist() == _brute_select(g, score, 0, second, k, r).tolist()


def test_greedy_select_resumed_in_batches():
    # the layout resumes the selection when columns are still missing: a
    # batch boundary closes the current round
    g = Graph.rgg(400, 8.0, seed=12)
    rng = np.random.default_rng(2)
    score = rng.random(g.n).astype(np.float32)
    d1 = O.seq_bfs(g, 0)
    second = int(np.argmax(d1))
    d2 = O.seq_bfs(g, second)
    for batches in ([7, 13, 24], [24], [5, 6, 24]):
        got = OL.greedy_select(g, score, d1, d2, 0, second, 24, 2, batches=batches)
        assert got.tolist() == _brute_select(g, score, 0, second, 24, 2, batches=batches).tolist()


def test_greedy_select_path_by_hand():
    # path 0..19, first 0, second 19, r = 3, constant scores: the first
    # vertex more than 3 hops from every pick, ties to the lower index
    g = Graph.path(20)
    d1 = O.seq_bfs(g, 0)
    d2 = O.seq_bfs(g, 19)
    got = OL.greedy_select(g, np.zeros(20, np.float32), d1, d2, 0, 19, 5, 3)
    assert got.tolist() == [0, 19, 4, 8, 12]


def test_coverage_target():
    assert OL.coverage_target([0, 2, 4]).tolist() == [1.0, 0.5, 0.0]


def test_parhde_given_sources_equals_kcenters():
    g = Graph.mesh(30, 40)
    r = O.parhde(g, 10, 2)
    r2 = O.parhde_sources(g, 10, r["pivots"], 2)
    assert r2["rc"] == 0 and r2["pivots"] == r["pivots"]
    assert np.array_equal(r2["coords"], r["coords"])
