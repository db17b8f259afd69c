"""ParHDE layout benchmark on one MI355X.

Workload (one "step" = one complete layout): random geometric graph of
4,194,304 points uniform in the unit square (seed 2), radius for average
degree 10, largest component, unweighted int32 CSR; ParHDE with s = 50
k-centers pivots (first pivot 0), c = 2 coordinates, FP64 orthogonalisation
and SpMM, the reference's power-method eigensolver, traversals run lazily
(only as many as the orthogonalisation consumes; outputs identical to the
reference's eager 3s traversals).  The graph is resident in HBM before the
timed region; each layout returns its coordinates to host memory.

Prints one JSON line (see the `record` dict below for the fields).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bfs_kernel_name(deep: bool, parts: int, boundary: int) -> str:
    """The traversal kernel the library launches: for a graph it found deep
    and cut into parts, the boundary-distance kernel (obfs_kernel, after a
    small init kernel and the source part's own search: the traversal's HIP
    events cover all three) or, without its per-part distance tables, the
    partitioned kernel; else the multi-hop kernel for a deep graph unless
    PHDE_BFS_HOPS=0, else the level-by-level one -- one launch per traversal
    -- or one launch per level with PHDE_BFS_PERSISTENT=0."""
    if os.environ.get("PHDE_BFS_PERSISTENT", "1") == "0":
        return "bfs_level_kernel"
    if deep and parts > 0 and boundary >= 0:
        return "obfs_kernel"
    if deep and parts > 0:
        return "pbfs_kernel"
    if deep and os.environ.get("PHDE_BFS_HOPS", "") not in ("0", "1"):
        return "bfs_hops_kernel"
    return "bfs_persistent_kernel"


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` on `workload`, as read by rocprofv3
    FETCH_SIZE + WRITE_SIZE counters (scaled by the plain-copy calibration)
    and recorded in profiles/traffic.json; None when not measured."""
    f = ROOT / "profiles" / "traffic.json"
    if not f.exists():
        return None
    rec = json.loads(f.read_text()).get(workload, {}).get(kernel)
    return None if rec is None else rec["bytes_per_launch"]


def copy_calibration(mib: int) -> None:
    """A plain device-to-device copy of `mib` MiB, three times, so that a
    counter run can relate FETCH_SIZE / WRITE_SIZE to known bytes.  Uses the
    HIP runtime's own copy kernel through hipMemcpy (no torch kernels)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    size = ctypes.c_size_t(mib << 20)
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(ctypes.byref(p), size) != 0:
            raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(src, 1, size) != 0:
        raise RuntimeError("hipMemset failed")
    for _ in range(3):
        if hip.hipMemcpy(dst, src, size, 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    log(f"[bench] copy calibration: 3 x {mib} MiB device copies")


def copy_bandwidth(mib: int = 1024, reps: int = 5) -> float:
    """Measured device-to-device copy rate in GB/s (bytes read + written per
    second, HIP events around hipMemcpy of `mib` MiB; best of `reps`): the
    practical HBM ceiling next to the 8 TB/s peak."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    size = ctypes.c_size_t(mib << 20)
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(ctypes.byref(p), size) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipEventCreate(ctypes.byref(a))
    hip.hipEventCreate(ctypes.byref(b))
    hip.hipMemset(src, 1, size)
    hip.hipMemcpy(dst, src, size, 3)  # hipMemcpyDeviceToDevice (warm-up)
    best = None
    for _ in range(reps):
        hip.hipEventRecord(a, None)
        hip.hipMemcpy(dst, src, size, 3)
        hip.hipEventRecord(b, None)
        hip.hipEventSynchronize(b)
        ms = ctypes.c_float()
        hip.hipEventElapsedTime(ctypes.byref(ms), a, b)
        best = ms.value if best is None else min(best, ms.value)
    hip.hipEventDestroy(a)
    hip.hipEventDestroy(b)
    hip.hipFree(src)
    hip.hipFree(dst)
    return 2.0 * (mib << 20) / (best * 1e-3) / 1e9


def bfs_bytes(n: int, m: int) -> int:
    # adjacency once, row pointers once, int32 level written once per vertex
    return 4 * m + 4 * (n + 1) + 4 * n


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=4194304)
    ap.add_argument("--s", type=int, default=50)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--eager", action="store_true")
    ap.add_argument("--extras", type=int, default=1, help="also time eager and fp32 once")
    ap.add_argument("--cpu-baseline", type=int, default=1,
                    help="time the FP64 oracle's full layouts on the host cores and report the GPU's FP32/FP64 "
                         "errors against them")
    ap.add_argument("--learned", type=int, default=1,
                    help="also time the learned-pivot layout against k-centers (needs the weights file)")
    ap.add_argument("--learned-n", type=int, default=1 << 24,
                    help="points of the second, held-out RGG for the learned comparison (0: skip)")
    ap.add_argument("--configs", type=int, default=1,
                    help="also time RMAT-22 (ParHDE) and the weighted 160^3 grid (WParHDE) in FP32 and FP64, "
                         "against the oracle unless --cpu-baseline 0")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the layouts, run a plain device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(every configuration of this workload fits one MI355X)", file=sys.stderr)
        sys.exit(2)

    from icpp_2020_parhde_amd import layout as L
    from icpp_2020_parhde_amd.graphs import Graph

    t0 = time.time()
    g = Graph.rgg(args.n, 10.0, seed=2)
    log(f"[bench] rgg n={g.n} nnz={g.m} generated in {time.time() - t0:.1f}s")
    t0 = time.time()
    ctx = L.Context(g)
    h2d_s = time.time() - t0
    kernel = bfs_kernel_name(ctx.vertex_order()[1], ctx.num_parts(), ctx.num_boundary())
    log(f"[bench] graph resident in HBM ({h2d_s * 1e3:.1f} ms incl. validation and H2D)")
    prec = L.PREC_FP32 if args.precision == "fp32" else L.PREC_FP64
    opts = dict(s=args.s, num_coord=2, precision=prec, eager=1 if args.eager else 0)

    for i in range(args.warmup):
        X, st = ctx.layout(**opts)
        log(f"[bench] warmup {i}: {st.ms_total:.1f} ms, {st.n_traversals} traversals")
    times, stats = [], []
    t_all = time.perf_counter()
    for i in range(args.steps):
        t = time.perf_counter()
        X, st = ctx.layout(**opts)
        times.append((time.perf_counter() - t) * 1e3)
        stats.append(st)
        log(f"[bench] step {i}: {times[-1]:.1f} ms (bfs {st.kernel_ms['bfs']:.1f} ms over "
            f"{st.n_traversals} traversals, orth {st.ms_orth:.1f}, mmult {st.ms_mmult:.1f}, "
            f"eig {st.ms_eig:.1f}, coords {st.ms_coords:.1f})")
    wall_ms = (time.perf_counter() - t_all) * 1e3
    ms_per_layout = wall_ms / args.steps
    st = stats[-1]
    nbfs = st.n_traversals
    bfs_ms = np.mean([s.kernel_ms["bfs"] for s in stats]) / nbfs
    gteps = (g.m / 2) / (bfs_ms * 1e-3) / 1e9
    achieved = bfs_bytes(g.n, g.m) / (bfs_ms * 1e-3) / 1e9

    extras = {}
    if args.extras:
        _, se = ctx.layout(s=args.s, num_coord=2, precision=prec, eager=1)
        extras["eager_ms_per_layout"] = round(se.ms_total, 2)
        extras["eager_traversals"] = se.n_traversals
        if prec == L.PREC_FP64:
            extras["fp32_ms_per_layout"] = round(timed_layouts(ctx, dict(opts, precision=L.PREC_FP32)), 2)
        extras["hbm_copy_gbs"] = round(copy_bandwidth(), 1)
        log(f"[bench] extras {extras}")
    accuracy = cpu = None
    if args.cpu_baseline:
        # the FP64 oracle's full layouts: the timed CPU baseline and the
        # reference for the FP32 / FP64 error report
        accuracy, cpu = accuracy_vs_oracle(g, ctx, L.ALGO_PARHDE, args.s, 2, eager_too=True)
        cpu["gpu_over_cpu_lazy"] = round(cpu["value"] / ms_per_layout, 2)
        if "eager_ms_per_layout" in extras:
            cpu["gpu_over_cpu_eager"] = round(cpu["eager_value"] / extras["eager_ms_per_layout"], 2)
    learned = None
    if args.learned:
        learned = {"rgg4m": learned_vs_kcenters(ctx, X, ms_per_layout, args.s, prec)}
        ctx.close()
        if args.learned_n:
            t0 = time.time()
            g5 = Graph.rgg(args.learned_n, 10.0, seed=5)
            log(f"[bench] rgg n={g5.n} nnz={g5.m} (seed 5) generated in {time.time() - t0:.1f}s")
            with L.Context(g5) as c5:
                t = time.perf_counter()
                X5, s5 = c5.layout(s=args.s, num_coord=2, precision=prec)
                ms5 = (time.perf_counter() - t) * 1e3
                log(f"[bench] rgg{g5.n} k-centers layout {ms5:.1f} ms")
                learned[f"rgg{args.learned_n >> 20}m"] = learned_vs_kcenters(c5, X5, ms5, args.s, prec)
            del g5
    else:
        ctx.close()

    other = other_configs(args) if args.configs else None

    workload = f"rgg{args.n >> 20}m_s{args.s}" if args.n >= 1 << 20 else f"rgg{args.n}_s{args.s}"
    record = {
        "metric": f"ParHDE ms/layout (RGG n={g.n}, s={args.s}, k-centers)",
        "value": round(ms_per_layout, 3),
        "unit": "ms/layout",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_layout, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "i32 levels / " + args.precision,
        "data": "synthetic (seeded random geometric graph)",
        "config": {"workload": workload, "n": g.n, "nnz": g.m, "s": args.s, "c": 2,
                   "pivots": "kcenters", "eig": "power", "traversals": "eager" if args.eager else "lazy",
                   "precision": args.precision},
        "bfs_gteps": round(gteps, 4),
        "bfs_ms_mean": round(bfs_ms, 4),
        "bfs_levels_mean": round(st.levels_total / nbfs, 1),
        "phases_ms": {"distance": round(st.ms_distance, 2), "orth": round(st.ms_orth, 2),
                      "mmult": round(st.ms_mmult, 2), "eig": round(st.ms_eig, 2),
                      "coords": round(st.ms_coords, 2)},
        "kernel_ms": {k: round(v, 3) for k, v in st.kernel_ms.items() if v},
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": measured_traffic(workload, kernel),
                     "algorithmic_bytes": bfs_bytes(g.n, g.m)},
        "h2d_ms": round(h2d_s * 1e3, 1),
        "ms_per_layout_incl_graph_upload": round(ms_per_layout + h2d_s * 1e3, 3),
        **extras,
        "accuracy_vs_fp64_oracle": accuracy,
        "learned_pivots": learned,
        "other_configs": other,
        "cpu_baseline": cpu,
    }
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    print(json.dumps(record), flush=True)


def timed_layouts(ctx, opts, reps=2):
    """Mean wall-clock ms of `reps` layouts (after one untimed)."""
    ctx.layout(**opts)
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        ctx.layout(**opts)
        ts.append((time.perf_counter() - t) * 1e3)
    return float(np.mean(ts))


def dorth_err_np(deg, S):
    """max_ij |B_ij / sqrt(B_ii B_jj) - delta_ij| of B = [1 S]^T D [1 S]."""
    U = np.column_stack([np.full(S.shape[0], 1 / np.sqrt(S.shape[0])), S])
    B = U.T @ (U * deg[:, None])
    d = np.sqrt(np.diag(B))
    return float(np.max(np.abs(B / np.outer(d, d) - np.eye(B.shape[0]))))


def accuracy_vs_oracle(g, ctx, algo, s, c, eager_too=False):
    """One full FP64 oracle layout (the reference's algorithm, OpenMP on the
    host cores; its HDE time is the CPU baseline), then the GPU's FP64 and
    FP32 layouts with diagnostics against it: coordinate error (relative to
    the largest entry, after sign alignment), eigenvalue error (relative to
    the largest eigenvalue) and the normalised D-orthogonality error of S.
    Returns (accuracy record, CPU timing record)."""
    from icpp_2020_parhde_amd import layout as L
    from oracle import pyoracle as O
    weighted = algo == L.ALGO_WPARHDE
    t0 = time.time()
    ref = O.wparhde(g, s, c, want_S=True) if weighted else O.parhde(g, s, c, lazy=True, want_S=True)
    wall = time.time() - t0
    if ref["rc"] != 0:
        raise RuntimeError(f"oracle layout failed: {ref['rc']}")
    if g.w is None:
        deg = np.diff(g.rowptr.astype(np.int64)).astype(np.float64)
    else:
        rows = np.repeat(np.arange(g.n), np.diff(g.rowptr.astype(np.int64)))
        deg = np.bincount(rows, weights=g.w.astype(np.float64), minlength=g.n)
    e_ref = dorth_err_np(deg, ref["S"])
    del ref["S"]
    acc = {"oracle_dorth_err": e_ref}
    lam = ref["eigval"]
    for name, prec in (("fp64", L.PREC_FP64), ("fp32", L.PREC_FP32)):
        X, st = ctx.layout(s=s, num_coord=c, algorithm=algo, precision=prec, diagnostics=1)
        ev = np.array(st.eigenvalues)
        acc[name] = {"coord_rel_err": aligned_rel_err(X, ref["coords"]),
                     "eig_rel_err": float(np.max(np.abs(ev - lam)) / np.max(np.abs(lam))),
                     "dorth_err": st.dorth_err, "pivots_equal": st.pivots == ref["pivots"]}
    cpu = {"value": round(ref["t_hde"] * 1e3, 1), "unit": "ms/layout", "cores": O.num_threads(), "kind": "port",
           "sample": (f"one complete {'WParHDE' if weighted else 'lazy ParHDE'} layout of the same graph "
                      f"({ref['n_bfs']} traversals, s={s}) by the oracle (g++ -O2 -fopenmp), timed over the "
                      f"reference's HDE span (traversals, orthogonalisation, L*S and S^T(LS), power method, "
                      f"coordinates); {wall:.1f}s wall"),
           "phases_ms": {"distance": round(ref["t_dist"] * 1e3, 1), "orth": round(ref["t_orth"] * 1e3, 1),
                         "mmult": round(ref["t_mmult"] * 1e3, 1)}}
    if eager_too:
        t0 = time.time()
        r = O.parhde(g, s, c, lazy=False)
        cpu["eager_value"] = round(r["t_hde"] * 1e3, 1)
        cpu["eager_traversals"] = r["n_bfs"]
        cpu["sample"] += f"; the eager layout ({r['n_bfs']} traversals, the reference as written) {time.time() - t0:.1f}s wall"
        acc["oracle_eager_equals_lazy"] = bool(np.array_equal(r["coords"], ref["coords"]))
    log(f"[bench] accuracy vs the FP64 oracle: {acc}; cpu {cpu}")
    return acc, cpu


def learned_vs_kcenters(ctx, X_kc, ms_kc, s, prec):
    """Learned-pivot layout (network scores + greedy suppression + one
    multi-source BFS) on the context's graph against its k-centers layout
    X_kc: time and Hall energy gap tr(Y'LY)/tr(Y'DY).  The same selection on
    constant scores (an all-zero network: picks by spacing alone) is timed
    beside it, so the network's own share of the result shows."""
    from icpp_2020_parhde_amd import layout as L
    from icpp_2020_parhde_amd.learned import DEFAULT_WEIGHTS
    if not DEFAULT_WEIGHTS.exists():
        log(f"[bench] no network weights at {DEFAULT_WEIGHTS}: learned comparison skipped")
        return None
    e_k = ctx.hall_energy(X_kc)
    opts = dict(s=s, num_coord=2, precision=prec, strategy=L.PIVOTS_LEARNED)
    rows = {}
    for name in ("trained", "no_network"):
        net = L.PivotNet(path=str(DEFAULT_WEIGHTS)) if name == "trained" else L.PivotNet(seed=0)
        if name == "no_network":
            p = net.params
            p[:] = 0
            net.params = p
        ctx.set_pivot_net(net)
        ctx.layout(**opts)
        times = []
        for _ in range(2):
            t = time.perf_counter()
            X, st = ctx.layout(**opts)
            times.append((time.perf_counter() - t) * 1e3)
        ms = float(np.mean(times))
        e_l = ctx.hall_energy(X)
        ctx.set_pivot_net(None)
        net.close()
        rows[name] = {"ms_per_layout": round(ms, 2), "speedup": round(ms_kc / ms, 3),
                      "ms_select": round(st.ms_select, 2), "ms_distance": round(st.ms_distance, 2),
                      "hall_energy": e_l, "gap": round(e_l / e_k - 1, 6), "within_2pct": bool(e_l <= 1.02 * e_k),
                      "n_candidates": st.n_candidates}
    r = {**rows["trained"], "ms_kcenters": round(ms_kc, 2), "hall_energy_kcenters": e_k,
         "no_network": rows["no_network"],
         "columns": "one multi-source search over the proposals (concurrent multi-hop searches on graphs "
                    "with a deep BFS, the 64-wide bit-parallel sweep otherwise); judge off"}
    log(f"[bench] learned pivots: {r}")
    return r


def aligned_rel_err(X, Y):
    err = 0.0
    for k in range(Y.shape[1]):
        x, y = X[:, k], Y[:, k]
        x = -x if np.dot(x, y) < 0 else x
        err = max(err, float(np.max(np.abs(x - y)) / np.max(np.abs(y))))
    return err


def other_configs(args):
    """RMAT scale 22 (a=.57 b=.19 c=.19, seed 3, LCC), ParHDE s=50, and the
    weighted 160^3 grid (weights 1..255, seed 4), WParHDE s=20 c=3: GPU ms
    per layout in FP32 and FP64, the FP32 / FP64 errors against the FP64
    oracle, and the oracle's own full layout timed on the host cores."""
    from icpp_2020_parhde_amd import layout as L
    from icpp_2020_parhde_amd.graphs import Graph
    out = {}
    for name, make, algo, s, c in [
            ("rmat22_s50", lambda: Graph.rmat(22, 16, seed=3), L.ALGO_PARHDE, 50, 2),
            ("grid160_wparhde_s20", lambda: Graph.grid3d(160, 160, 160, seed=4, wmax=255), L.ALGO_WPARHDE, 20, 3)]:
        t0 = time.time()
        g = make()
        log(f"[bench] {name}: n={g.n} nnz={g.m} ({time.time() - t0:.1f}s)")
        with L.Context(g) as ctx:
            opts = dict(s=s, num_coord=c, algorithm=algo)
            ms32 = timed_layouts(ctx, dict(opts, precision=L.PREC_FP32))
            ms64 = timed_layouts(ctx, dict(opts, precision=L.PREC_FP64))
            _, st = ctx.layout(precision=L.PREC_FP64, **opts)
            trav = st.kernel_ms["bfs" if algo == L.ALGO_PARHDE else "sssp"] / st.n_traversals
            out[name] = {"n": g.n, "nnz": g.m, "gpu_fp32_ms_per_layout": round(ms32, 2),
                         "gpu_fp64_ms_per_layout": round(ms64, 2),
                         "gpu_ms_per_traversal": round(trav, 3), "traversals": st.n_traversals,
                         "phases_ms_fp64": {"distance": round(st.ms_distance, 2), "orth": round(st.ms_orth, 2),
                                            "mmult": round(st.ms_mmult, 2), "eig": round(st.ms_eig, 2),
                                            "coords": round(st.ms_coords, 2)}}
            if args.cpu_baseline:
                acc, cpu = accuracy_vs_oracle(g, ctx, algo, s, c)
                out[name].update({"accuracy_vs_fp64_oracle": acc, "cpu_ms_per_layout": cpu["value"],
                                  "cpu_cores": cpu["cores"], "cpu_phases_ms": cpu["phases_ms"],
                                  "speedup_vs_cpu_fp64": round(cpu["value"] / ms64, 2),
                                  "speedup_vs_cpu_fp32": round(cpu["value"] / ms32, 2)})
        log(f"[bench] {name}: {out[name]}")
        del g
    return out



if __name__ == "__main__":
    main()
