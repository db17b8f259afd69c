"""CPU restatement of the learned pivot proposal (icpp_2020_parhde_amd/csrc/learned.hip).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product
never imports it.

The reference has no learned part (its strategies are k-centers and random
pivots, ParHDE.cpp:118-172), so nothing here restates reference code; this
module restates the GPU design so that the GPU kernels can be checked:

  features(g, d1, d2)        x_v = (log(1+deg)/4, d1/ecc1, d2/ecc2, 1), FP32 as on the GPU
  forward / loss_and_grad    the 3-layer message-passing network and its
                             backward pass; with emulate=True the values are
                             rounded to bfloat16 at the points where the GPU
                             stores BF16 (activations, neighbour means, MFMA
                             operands), everything else in float64
  greedy_select              the greedy graph non-maximum suppression in
                             rounds, exact integer semantics (bit-exact target
                             for the GPU)
  coverage_target            the training target 1 - md / max(md) from the
                             k-centers distance-to-nearest-pivot field

Pins: loss_and_grad(emulate=False) against central finite differences of
forward (tests/test_oracle_learned.py); greedy_select against a brute-force
restatement from all-pairs distances; bf16_round against the exhaustive
definition on sampled bit patterns (ties to even).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import dijkstra

H = 64
LAYOUT = [("W1", (8, 64)), ("b1", (64,)), ("W2", (128, 64)), ("b2", (64,)), ("W3", (128, 64)),
          ("b3", (64,)), ("wo", (64,)), ("bo", (1,))]
NPARAMS = sum(int(np.prod(s)) for _, s in LAYOUT)


def unpack(p):
    out, o = {}, 0
    for name, shape in LAYOUT:
        k = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + k], dtype=np.float64).reshape(shape)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1) for name, _ in LAYOUT])


def bf16_round(x):
    """Round to the nearest bfloat16 (ties to even); returns float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def adjacency(g):
    n = g.n
    A = sp.csr_matrix((np.ones(g.m), g.adj.astype(np.int64), g.rowptr.astype(np.int64)), shape=(n, n))
    deg = np.diff(g.rowptr.astype(np.int64)).astype(np.float64)
    inv = np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0)
    return A, deg, inv


def features(g, d1, d2):
    """FP32 features exactly as learned.hip's feature_kernel computes them
    (up to log1pf's last bit)."""
    deg = np.diff(g.rowptr.astype(np.int64)).astype(np.float32)
    e1, e2 = int(d1.max()), int(d2.max())
    inv1 = np.float32(1.0) / np.float32(e1) if e1 > 0 else np.float32(0)
    inv2 = np.float32(1.0) / np.float32(e2) if e2 > 0 else np.float32(0)
    x = np.empty((g.n, 4), dtype=np.float32)
    x[:, 0] = np.log1p(deg) * np.float32(0.25)
    x[:, 1] = d1.astype(np.float32) * inv1
    x[:, 2] = d2.astype(np.float32) * inv2
    x[:, 3] = 1.0
    return x


def forward(g, x, p, emulate=True, cache=False):
    """Scores y (n,) of the network; with cache=True also the intermediates."""
    R = bf16_round if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    P = unpack(p)
    A, deg, inv = adjacency(g)
    x = np.asarray(x, dtype=np.float64)
    xa = (A @ x) * inv[:, None]
    z1 = np.concatenate([x, xa], axis=1)
    h1 = R(np.maximum(z1 @ P["W1"] + P["b1"], 0))
    z2 = np.concatenate([h1, R((A @ h1) * inv[:, None])], axis=1)
    h2 = R(np.maximum(z2 @ R(P["W2"]) + P["b2"], 0))
    z3 = np.concatenate([h2, R((A @ h2) * inv[:, None])], axis=1)
    h3 = R(np.maximum(z3 @ R(P["W3"]) + P["b3"], 0))
    y = h3 @ P["wo"] + P["bo"][0]
    if not cache:
        return y
    return y, dict(A=A, inv=inv, z1=z1, h1=h1, z2=z2, h2=h2, z3=z3, h3=h3, P=P, R=R)


def loss_and_grad(g, x, t, p, emulate=True):
    """Mean squared error of forward() against t and its gradient (flat)."""
    y, c = forward(g, x, p, emulate, cache=True)
    R, P, A, inv = c["R"], c["P"], c["A"], c["inv"]
    n = g.n
    diff = y - np.asarray(t, dtype=np.float64)
    loss = float(np.mean(diff * diff))
    gy = 2.0 * diff / n
    G = {"wo": c["h3"].T @ gy, "bo": np.array([gy.sum()])}
    d = gy[:, None] * P["wo"][None, :] * (c["h3"] > 0)
    G["b3"] = d.sum(0)
    db = R(d)
    G["W3"] = c["z3"].T @ db
    gz = db @ R(P["W3"]).T
    gh = gz[:, :64] + A @ (gz[:, 64:] * inv[:, None])
    d = gh * (c["h2"] > 0)
    G["b2"] = d.sum(0)
    db = R(d)
    G["W2"] = c["z2"].T @ db
    gz = db @ R(P["W2"]).T
    gh = gz[:, :64] + A @ (gz[:, 64:] * inv[:, None])
    d = gh * (c["h1"] > 0)
    G["b1"] = d.sum(0)
    G["W1"] = c["z1"].T @ d
    return loss, pack(G), y


def orderable(score):
    u = np.asarray(score, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return np.where(u & 0x80000000, (~u) & 0xFFFFFFFF, u | 0x80000000)


ROUND = 8  # picks per round (learned.hip kSelRound)


def greedy_select(g, score, d1, d2, first, second, k, r, batches=None):
    """Pivots of learned.hip's select_kernel.  Picks 0 and 1 are given (the
    feature traversals' sources).  The rest come in rounds of up to ROUND:
    each pick maximises (min(sd, r+1), top 21 bits of the score's order key,
    -index) over the vertices v with max(|d1_v - d1_c|, |d2_v - d2_c|) > r
    for every earlier pick c of the same round, where sd is the distance to
    the picks of earlier rounds (exact within r hops; the two given picks
    exactly).  A round ends when the best such vertex has sd <= r (its first
    pick is always taken), or at the end of a batch (`batches`: increasing
    pick counts at which the selection was resumed, default [k]).  After a
    round, sd = min(sd, distance to the round's picks where that is <= r)."""
    A, _, _ = adjacency(g)
    r = max(1, min(int(r), 2046))
    d1 = np.asarray(d1, dtype=np.int64)
    d2 = np.asarray(d2, dtype=np.int64)
    sd = np.minimum(d1, d2)
    hi = orderable(score) >> np.uint64(11)
    idx = np.arange(g.n, dtype=np.uint64)
    picks = [int(first), int(second)]
    for end in (batches or [k]):
        while len(picks) < end:
            rnd = []
            while len(rnd) < ROUND and len(picks) < end:
                far = np.ones(g.n, dtype=bool)
                for c in rnd:
                    far &= np.maximum(np.abs(d1 - d1[c]), np.abs(d2 - d2[c])) > r
                if not far.any():
                    break
                sdc = np.minimum(sd, r + 1).astype(np.uint64)
                key = (((sdc << np.uint64(21)) | hi) << np.uint64(32)) | (np.uint64(0xFFFFFFFF) - idx)
                key = np.where(far, key, np.uint64(0))
                best = key.max()
                if rnd and int(best >> np.uint64(53)) != r + 1:
                    break
                p = int(np.uint64(0xFFFFFFFF) - (best & np.uint64(0xFFFFFFFF)))
                rnd.append(p)
                picks.append(p)
            dp = dijkstra(A, unweighted=True, indices=rnd, limit=r + 0.5, min_only=True)
            reach = np.isfinite(dp)
            sd[reach] = np.minimum(sd[reach], dp[reach].astype(np.int64))
    return np.array(picks, dtype=np.int64)


def coverage_target(min_dist):
    md = np.asarray(min_dist, dtype=np.float64)
    top = md.max()
    return (1.0 - md / top if top > 0 else np.ones_like(md)).astype(np.float32)
