"""Training and evaluation of the learned pivot proposal.

The network (csrc/learned.hip) scores every vertex from features of two
traversals; the layout's greedy suppression turns the scores into pivots.
It is trained here, on the GPU, with the library's own forward, backward and
Adam kernels (phde_net_*), against the coverage field of the plain k-centers
pivots that the library computes for each training graph (phde_kcenters):

    target_v = 1 - md_v / max(md),  md = hops from v to the nearest of the s
                                         k-centers pivots

Training graphs are seeded and reproducible: seeds 1000-2999 cycle through
random geometric graphs (n in [10k, 200k], average degree 10), holed grids
and RMAT graphs (scale 14-17, a=.57 b=.19 c=.19), each reduced to its
largest component.  Graphs are packed into disjoint unions of about two
million vertices per training batch (the neighbour mean never crosses a
component, so a union trains exactly like its parts).

Command line:
    python -m icpp_2020_parhde_amd.learned train --out weights.bin [--graphs 2000]
    python -m icpp_2020_parhde_amd.learned eval --weights weights.bin
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path
from typing import List, Optional

import numpy as np

from . import layout as L
from .graphs import Graph

DEFAULT_WEIGHTS = Path(__file__).resolve().parent / "weights" / "pivot_net.bin"
TRAIN_SEEDS = range(1000, 3000)
S_PIVOTS = 50


def _mix(seed: int, salt: int) -> int:
    z = (seed * 0x9E3779B97F4A7C15 + salt * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z ^= z >> 31
    z = (z * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return z ^ (z >> 29)


def training_graph(seed: int) -> Graph:
    """The training graph of `seed` (its kind cycles with seed % 3)."""
    kind = seed % 3
    if kind == 0:
        n = 10_000 + _mix(seed, 1) % 190_001
        return Graph.rgg(n, 10.0, seed=seed)
    if kind == 1:
        rows = 100 + _mix(seed, 2) % 351
        cols = 100 + _mix(seed, 3) % 351
        holes = 3 + _mix(seed, 4) % 18
        return Graph.holed_grid(rows, cols, holes, seed)
    scale = 14 + _mix(seed, 5) % 4
    return Graph.rmat(scale, 16, seed=seed)


def union(graphs: List[Graph]) -> Graph:
    """Disjoint union (vertex ids offset in order)."""
    n_off = np.cumsum([0] + [g.n for g in graphs])
    m_off = np.cumsum([0] + [g.m for g in graphs])
    rowptr = np.empty(n_off[-1] + 1, dtype=np.uint64)
    adj = np.empty(m_off[-1], dtype=np.uint64)
    for g, no, mo in zip(graphs, n_off, m_off):
        rowptr[no:no + g.n] = g.rowptr[:-1].astype(np.uint64) + mo
        adj[mo:mo + g.m] = g.adj.astype(np.uint64) + no
    rowptr[-1] = m_off[-1]
    if m_off[-1] >= 2 ** 32:
        raise ValueError("union too large for 32-bit CSR")
    return Graph(int(n_off[-1]), int(m_off[-1]), rowptr.astype(np.uint32), adj.astype(np.uint32))


def label(g: Graph, s: int = S_PIVOTS, first: int = 0):
    """(features n x 4, coverage target n) of one graph, computed on the GPU."""
    with L.Context(g) as ctx:
        feat, _ = ctx.pivot_features(first)
        _, md = ctx.kcenters(s, first)
    top = float(md.max())
    target = (1.0 - md.astype(np.float64) / top if top > 0 else np.ones(g.n)).astype(np.float32)
    return feat, target


def log(*a):
    print(*a, flush=True)


def train(out: str, n_graphs: int = 2000, epochs: int = 24, lr: float = 2e-3, batch_vertices: int = 2_000_000,
          seed: int = 7, log_every: int = 50) -> dict:
    t0 = time.time()
    data = []
    seeds = list(TRAIN_SEEDS)[:n_graphs]
    for i, sd in enumerate(seeds):
        g = training_graph(sd)
        feat, target = label(g)
        data.append((g, feat, target))
        if (i + 1) % log_every == 0 or i + 1 == len(seeds):
            log(f"[train] labelled {i + 1}/{len(seeds)} graphs ({time.time() - t0:.0f}s, last n={g.n})")
    # pack into batches of about batch_vertices vertices (fixed order)
    batches, cur = [], []
    for item in data:
        cur.append(item)
        if sum(x[0].n for x in cur) >= batch_vertices:
            batches.append(cur)
            cur = []
    if cur:
        batches.append(cur)
    ctxs = []
    for b in batches:
        ug = union([x[0] for x in b])
        ctx = L.Context(ug)
        ctx.set_training_data(np.concatenate([x[1] for x in b]), np.concatenate([x[2] for x in b]))
        ctxs.append(ctx)
    log(f"[train] {len(data)} graphs, {sum(c.n for c in ctxs)} vertices in {len(ctxs)} batches "
        f"({time.time() - t0:.0f}s)")
    net = L.PivotNet(seed=seed)
    rng = np.random.default_rng(seed)
    history = []
    for ep in range(epochs):
        # cosine decay from lr to lr / 20
        lr_ep = lr / 20 + (lr - lr / 20) * 0.5 * (1 + np.cos(np.pi * ep / max(1, epochs - 1)))
        losses = []
        for j in rng.permutation(len(ctxs)):
            losses.append(net.train_step(ctxs[j], float(lr_ep)))
        history.append(float(np.mean(losses)))
        log(f"[train] epoch {ep + 1}/{epochs} lr {lr_ep:.2e} loss {history[-1]:.5f} ({time.time() - t0:.0f}s)")
        Path(out).parent.mkdir(parents=True, exist_ok=True)
        net.save(out)
    for c in ctxs:
        c.close()
    return {"graphs": len(data), "batches": len(ctxs), "loss": history, "seconds": time.time() - t0}


def layout_gap(g: Graph, net: Optional[L.PivotNet], s: int = S_PIVOTS, **kw) -> dict:
    """k-centers layout vs learned-pivot layout on g: times and Hall energies.
    With net=None the selection runs on constant scores (no network)."""
    with L.Context(g) as ctx:
        _, st_k = ctx.layout(s=s, num_coord=2)
        Xk, st_k = ctx.layout(s=s, num_coord=2)
        e_k = ctx.hall_energy(Xk)
        own = net if net is not None else L.PivotNet(seed=0)
        if net is None:
            p = own.params
            p[:] = 0
            own.params = p  # zero network: every score equal, selection by spacing only
        ctx.set_pivot_net(own)
        ctx.layout(s=s, num_coord=2, strategy=L.PIVOTS_LEARNED, **kw)
        Xl, st_l = ctx.layout(s=s, num_coord=2, strategy=L.PIVOTS_LEARNED, **kw)
        e_l = ctx.hall_energy(Xl)
        ctx.set_pivot_net(None)
        if net is None:
            own.close()
    return {"n": g.n, "ms_kcenters": st_k.ms_total, "ms_learned": st_l.ms_total, "ms_select": st_l.ms_select,
            "E_kcenters": e_k, "E_learned": e_l, "gap": e_l / e_k - 1.0,
            "speedup": st_k.ms_total / st_l.ms_total}


def evaluate(weights: str, big: bool = True) -> dict:
    net = L.PivotNet(path=weights)
    res = {}
    held = [("rgg100k", Graph.rgg(100_000, 10.0, seed=5001)), ("holed300", Graph.holed_grid(300, 300, 12, 5002)),
            ("rmat15", Graph.rmat(15, 16, seed=5003)), ("rgg150k", Graph.rgg(150_000, 10.0, seed=5004))]
    if big:
        held.append(("rgg16m", Graph.rgg(1 << 24, 10.0, seed=5)))
    for name, g in held:
        r = {"trained": layout_gap(g, net), "no_network": layout_gap(g, None)}
        res[name] = r
        log(f"[eval] {name} n={g.n}: trained gap {100 * r['trained']['gap']:+.3f}% "
            f"(x{r['trained']['speedup']:.2f}), no network {100 * r['no_network']['gap']:+.3f}%")
    net.close()
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--out", default=str(DEFAULT_WEIGHTS))
    t.add_argument("--graphs", type=int, default=2000)
    t.add_argument("--epochs", type=int, default=24)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--report", default=None)
    e = sub.add_parser("eval")
    e.add_argument("--weights", default=str(DEFAULT_WEIGHTS))
    e.add_argument("--no-big", action="store_true")
    e.add_argument("--report", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "train":
        r = train(a.out, a.graphs, a.epochs, a.lr)
    else:
        r = evaluate(a.weights, big=not a.no_big)
    if a.report:
        Path(a.report).write_text(json.dumps(r, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
