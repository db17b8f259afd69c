"""ctypes view of oracle/liboracle.so and oracle/_ref/libref_pc.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() (as
the checker) and bench.py's CPU-baseline leg.  The product never imports it.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libref_pc.so"

c_u32p = ctypes.POINTER(ctypes.c_uint32)
c_i32p = ctypes.POINTER(ctypes.c_int32)
c_i64p = ctypes.POINTER(ctypes.c_int64)
c_f64p = ctypes.POINTER(ctypes.c_double)


class OrGraph(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("m", ctypes.c_int64), ("rowptr", c_u32p), ("adj", c_u32p),
                ("w", c_i32p)]


class OrResult(ctypes.Structure):
    _fields_ = [("coords", c_f64p), ("coords_raw", c_f64p), ("S", c_f64p), ("SLS", c_f64p),
                ("eigvec", c_f64p), ("eigval", c_f64p), ("pivots", c_i64p), ("accepted", c_i32p),
                ("normdist", c_f64p), ("n_bfs", ctypes.c_int32), ("n_accepted", ctypes.c_int32),
                ("eig_iters", ctypes.c_int32), ("t_hde", ctypes.c_double), ("t_dist", ctypes.c_double),
                ("t_orth", ctypes.c_double), ("t_mmult", ctypes.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} not built (make -C oracle)")
        _lib = ctypes.CDLL(str(LIB))
        _lib.or_seq_bfs.argtypes = [ctypes.POINTER(OrGraph), ctypes.c_int64, c_i32p]
        _lib.or_dobfs.argtypes = [ctypes.POINTER(OrGraph), ctypes.c_int64, c_f64p, ctypes.c_int, ctypes.c_int]
        _lib.or_deltastep.argtypes = [ctypes.POINTER(OrGraph), ctypes.c_int64, ctypes.c_int32, c_f64p]
        _lib.or_dijkstra.argtypes = [ctypes.POINTER(OrGraph), ctypes.c_int64, c_i64p]
        _lib.or_power_eig.argtypes = [c_f64p, c_f64p, ctypes.c_int64, c_i32p]
        _lib.or_power_eig.restype = ctypes.c_int
        for f in ("or_parhde", "or_pivotmds", "or_phde"):
            getattr(_lib, f).argtypes = [ctypes.POINTER(OrGraph), ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(OrResult)] if f != "or_parhde" else \
                [ctypes.POINTER(OrGraph), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(OrResult)]
            getattr(_lib, f).restype = ctypes.c_int
        _lib.or_wparhde.argtypes = [ctypes.POINTER(OrGraph), ctypes.c_int, ctypes.c_int, ctypes.c_int32,
                                    ctypes.POINTER(OrResult)]
        _lib.or_wparhde.restype = ctypes.c_int
        _lib.or_hall_energy.argtypes = [ctypes.POINTER(OrGraph), c_f64p, ctypes.c_int]
        _lib.or_hall_energy.restype = ctypes.c_double
        _lib.or_num_threads.restype = ctypes.c_int
    return _lib


def ref_lib():
    """The reference's own ParallelComponents (oracle/_ref), or None if not built."""
    if not REF_LIB.exists():
        return None
    r = ctypes.CDLL(str(REF_LIB))
    r.ref_power_eig.argtypes = [c_f64p, c_f64p, ctypes.c_long]
    r.ref_power_eig.restype = ctypes.c_int
    return r


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def graph(g) -> OrGraph:
    og = OrGraph()
    og.n, og.m = g.n, g.m
    og._keep = (np.ascontiguousarray(g.rowptr, np.uint32), np.ascontiguousarray(g.adj, np.uint32),
                None if g.w is None else np.ascontiguousarray(g.w, np.int32))
    og.rowptr = _p(og._keep[0], ctypes.c_uint32)
    og.adj = _p(og._keep[1], ctypes.c_uint32)
    if og._keep[2] is not None:
        og.w = _p(og._keep[2], ctypes.c_int32)
    return og


def seq_bfs(g, src):
    d = np.empty(g.n, np.int32)
    og = graph(g)
    lib().or_seq_bfs(ctypes.byref(og), src, _p(d, ctypes.c_int32))
    return d


def dobfs(g, src, alpha=15, beta=18):
    d = np.full(g.n, -1.0)
    og = graph(g)
    lib().or_dobfs(ctypes.byref(og), src, _p(d, ctypes.c_double), alpha, beta)
    return d


def deltastep(g, src, delta=10000):
    d = np.empty(g.n)
    og = graph(g)
    lib().or_deltastep(ctypes.byref(og), src, delta, _p(d, ctypes.c_double))
    return d


def dijkstra(g, src):
    d = np.empty(g.n, np.int64)
    og = graph(g)
    lib().or_dijkstra(ctypes.byref(og), src, _p(d, ctypes.c_int64))
    return d


def power_eig(K):
    """K: flat column-major buffer (dim*dim). Returns (K_out flat, evals, iters, rc)."""
    Kf = np.array(K, dtype=np.float64).copy()
    dim = int(round(len(Kf) ** 0.5))
    ev = np.empty(dim)
    it = ctypes.c_int32()
    rc = lib().or_power_eig(_p(Kf, ctypes.c_double), _p(ev, ctypes.c_double), dim, ctypes.byref(it))
    return Kf, ev, it.value, rc


def ref_power_eig(K):
    r = ref_lib()
    if r is None:
        return None
    Kf = np.array(K, dtype=np.float64).copy()
    dim = int(round(len(Kf) ** 0.5))
    ev = np.empty(dim)
    rc = r.ref_power_eig(_p(Kf, ctypes.c_double), _p(ev, ctypes.c_double), dim)
    return Kf, ev, rc


def _layout(fn, g, s, c, *extra, want_S=False, lim=None):
    n = g.n
    lim = 3 * s if lim is None else lim
    out = {
        "coords": np.zeros(n * c), "coords_raw": np.zeros(n * c), "SLS": np.zeros(s * s),
        "eigvec": np.zeros(s * c), "eigval": np.zeros(s), "pivots": np.full(lim, -1, np.int64),
        "accepted": np.full(lim, -1, np.int32), "normdist": np.zeros(lim),
    }
    if want_S:
        out["S"] = np.zeros(n * s)
    r = OrResult()
    for k, v in out.items():
        t = {np.float64: ctypes.c_double, np.int64: ctypes.c_int64, np.int32: ctypes.c_int32}[v.dtype.type]
        setattr(r, k, _p(v, t))
    og = graph(g)
    rc = fn(ctypes.byref(og), s, c, *extra, ctypes.byref(r))
    res = {
        "rc": rc, "coords": out["coords"].reshape(c, n).T, "coords_raw": out["coords_raw"].reshape(c, n).T,
        "SLS": out["SLS"], "eigvec": out["eigvec"].reshape(c, s).T, "eigval": out["eigval"],
        "pivots": out["pivots"][: r.n_bfs].tolist(), "accepted": out["accepted"], "normdist": out["normdist"],
        "n_bfs": r.n_bfs, "n_accepted": r.n_accepted, "eig_iters": r.eig_iters,
        "t_hde": r.t_hde, "t_dist": r.t_dist, "t_orth": r.t_orth, "t_mmult": r.t_mmult,
    }
    if want_S:
        res["S"] = out["S"].reshape(s, n).T
    return res


def parhde(g, s, c=2, lazy=True, want_S=False):
    return _layout(lib().or_parhde, g, s, c, 1 if lazy else 0, want_S=want_S)


def parhde_sources(g, s, sources, c=2, want_S=False):
    """ParHDE with the candidate sources in the given order (random / learned
    pivots): CGS until s columns are accepted."""
    src = np.ascontiguousarray(sources, dtype=np.int64)
    L = lib()
    L.or_parhde_sources.argtypes = [ctypes.POINTER(OrGraph), ctypes.c_int, ctypes.c_int, c_i64p, ctypes.c_int,
                                    ctypes.POINTER(OrResult)]
    L.or_parhde_sources.restype = ctypes.c_int
    return _layout(L.or_parhde_sources, g, s, c, _p(src, ctypes.c_int64), len(src), want_S=want_S,
                   lim=len(src))


def wparhde(g, s, c=2, delta=10000, want_S=False):
    return _layout(lib().or_wparhde, g, s, c, delta, want_S=want_S)


def pivotmds(g, s, c=2, want_S=False):
    return _layout(lib().or_pivotmds, g, s, c, want_S=want_S)


def phde(g, s, c=2, want_S=False):
    return _layout(lib().or_phde, g, s, c, want_S=want_S)


def hall_energy(g, Y):
    Y = np.asarray(Y, np.float64)
    if Y.ndim == 1:
        Y = Y[:, None]
    Yc = np.ascontiguousarray(Y.T.reshape(-1))
    og = graph(g)
    return lib().or_hall_energy(ctypes.byref(og), _p(Yc, ctypes.c_double), Y.shape[1])


def num_threads():
    return lib().or_num_threads()


def set_threads(k: int):
    os.environ["OMP_NUM_THREADS"] = str(k)


class OrSample(ctypes.Structure):
    _fields_ = [("t_bfs_each", ctypes.c_double), ("t_orth_total", ctypes.c_double),
                ("t_mmult", ctypes.c_double), ("t_eig", ctypes.c_double), ("t_coords", ctypes.c_double),
                ("t_spent", ctypes.c_double), ("n_bfs_timed", ctypes.c_int32)]


def parhde_sample(g, s, nbfs=6, c=2, discards=0):
    """CPU baseline of one ParHDE layout on graph g from a bounded sample.

    Returns the cpu_baseline record: value = estimated ms per layout with
    the reference's eager 3s traversals; also the lazy (s + discards)
    estimate and the measured phase times."""
    L = lib()
    L.or_parhde_sample.argtypes = [ctypes.POINTER(OrGraph), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(OrSample)]
    L.or_parhde_sample.restype = ctypes.c_int
    og = graph(g)
    out = OrSample()
    rc = L.or_parhde_sample(ctypes.byref(og), s, c, nbfs, ctypes.byref(out))
    if rc != 0:
        raise RuntimeError(f"or_parhde_sample failed: {rc}")
    rest = out.t_orth_total + out.t_mmult + out.t_eig + out.t_coords
    eager = 3 * s * out.t_bfs_each + rest
    lazy = (s + discards) * out.t_bfs_each + rest
    return {
        "value": round(eager * 1e3, 1),
        "unit": "ms/layout",
        "cores": num_threads(),
        "kind": "port",
        "sample": (f"full graph; {out.n_bfs_timed} of the {3 * s} k-centers BFS timed (x{3 * s}), "
                   f"CGS timed at 3 column counts and summed over s={s}, L*S + S^T(LS), power method "
                   f"and coordinates timed once at full size; {out.t_spent:.1f}s of CPU work"),
        "lazy_value": round(lazy * 1e3, 1),
        "phases_ms": {"bfs_each": round(out.t_bfs_each * 1e3, 2), "orth": round(out.t_orth_total * 1e3, 1),
                      "mmult": round(out.t_mmult * 1e3, 1), "eig": round(out.t_eig * 1e3, 2),
                      "coords": round(out.t_coords * 1e3, 1)},
    }
