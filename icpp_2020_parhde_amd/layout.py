"""Python view of the MI355X ParHDE library.

``Context`` owns one graph resident in GPU memory (``phde_create``) and
exposes the traversals and the layout.  ``DrawingBase`` / ``ParHDE`` /
``WParHDE`` / ``PivotMDS`` / ``PHDE`` mirror the reference's C++ classes
(variants/DrawingBase.h:13-48 and subclasses): same constructor arguments,
``run(tries, alpha, beta, delta)`` and ``getCoordinates()``.  Every call goes
through libparhde.so; there is no host fallback.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _native
from ._native import PhdeOptions, PhdeStats, as_ptr
from .graphs import Graph

# include/parhde.h enums
ALGO_PIVOTMDS, ALGO_PHDE, ALGO_WPARHDE, ALGO_PARHDE = 0, 1, 2, 3
PIVOTS_RANDCOARSE, PIVOTS_RANDFINE, PIVOTS_KCENTERS, PIVOTS_LEARNED = 0, 1, 2, 3
PREC_FP64, PREC_FP32 = 0, 1
EIG_POWER, EIG_JACOBI, EIG_GENERALIZED = 0, 1, 2

STATUS = {
    0: "ok", -1: "bad argument", -2: "device error", -3: "out of device memory",
    -4: "too few independent pivots", -5: "eigensolver failed", -6: "weights mismatch",
    -7: "graph not connected", -8: "unsupported",
}

KERNEL_CLASSES = ["bfs", "kcenter", "orth", "spmm", "gram", "eig", "coords", "sssp", "msbfs", "other"]


class PhdeError(RuntimeError):
    def __init__(self, code: int, what: str):
        msg = _native.parhde_lib().phde_last_error().decode(errors="replace")
        super().__init__(f"{what}: {STATUS.get(code, code)} ({code}) {msg}")
        self.code = code


def _check(rc: int, what: str):
    if rc != 0:
        raise PhdeError(rc, what)


def device_count() -> int:
    c = ctypes.c_int(0)
    _native.parhde_lib().phde_device_count(ctypes.byref(c))
    return c.value


@dataclass
class LayoutStats:
    ms_total: float = 0.0
    ms_distance: float = 0.0
    ms_orth: float = 0.0
    ms_mmult: float = 0.0
    ms_eig: float = 0.0
    ms_coords: float = 0.0
    n_traversals: int = 0
    n_accepted: int = 0
    n_discarded: int = 0
    eig_iters: int = 0
    levels_total: int = 0
    pivots: list = field(default_factory=list)
    kernel_ms: dict = field(default_factory=dict)
    kernel_launches: dict = field(default_factory=dict)
    ms_select: float = 0.0
    hall_energy: float = 0.0
    hall_energy_kcenters: float = 0.0
    fallback: int = 0
    n_candidates: int = 0
    normdist: list = field(default_factory=list)     # ||a|| of each examined candidate after projection
    eigenvalues: list = field(default_factory=list)  # the s eigenvalue estimates, solver order
    dorth_err: float = 0.0                           # with diagnostics=1: D-orthogonality error of [1 S]
    n_gram: int = 0                                  # candidates taken through the Gram-space CGS

    @classmethod
    def from_c(cls, s: PhdeStats, dim: int = 0) -> "LayoutStats":
        nt = int(s.n_traversals)
        ne = min(int(s.n_examined), 192)
        return cls(s.ms_total, s.ms_distance, s.ms_orth, s.ms_mmult, s.ms_eig, s.ms_coords, nt,
                   int(s.n_accepted), int(s.n_discarded), int(s.eig_iters), int(s.levels_total),
                   [int(s.pivots[i]) for i in range(min(nt, 192))],
                   {k: float(s.kernel_ms[i]) for i, k in enumerate(KERNEL_CLASSES)},
                   {k: int(s.kernel_launches[i]) for i, k in enumerate(KERNEL_CLASSES)},
                   float(s.ms_select), float(s.hall_energy), float(s.hall_energy_kcenters),
                   int(s.fallback), int(s.n_candidates), [float(s.normdist[i]) for i in range(ne)],
                   [float(s.eigenvalues[i]) for i in range(min(dim, 64))], float(s.dorth_err),
                   int(s.n_gram))


def default_options(**kw) -> PhdeOptions:
    o = PhdeOptions()
    _native.parhde_lib().phde_default_options(ctypes.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


class Context:
    """One graph in HBM plus a HIP stream (phde_create / phde_destroy)."""

    def __init__(self, g: Graph):
        self.lib = _native.parhde_lib()
        self.graph = g
        self.n = g.n
        h = ctypes.c_void_p()
        rp = np.ascontiguousarray(g.rowptr, dtype=np.uint32)
        ad = np.ascontiguousarray(g.adj, dtype=np.uint32) if g.m > 0 else np.zeros(1, np.uint32)
        wp = None
        if g.w is not None:
            w = np.ascontiguousarray(g.w, dtype=np.int32)
            wp = as_ptr(w, ctypes.c_int32)
        _check(self.lib.phde_create(as_ptr(rp, ctypes.c_uint32), as_ptr(ad, ctypes.c_uint32), wp,
                                    g.n, g.m, ctypes.byref(h)), "phde_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.phde_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ----------------------------------------------------------- traversals
    def bfs(self, source: int, alpha: int = 15, beta: int = 18) -> np.ndarray:
        """Levels from `source` (DOBFS, gapbs/bfs.cc:144); -1 when unreached."""
        d = np.empty(self.n, dtype=np.int32)
        _check(self.lib.phde_bfs(self.h, source, alpha, beta, as_ptr(d, ctypes.c_int32)), "phde_bfs")
        return d

    def vertex_order(self):
        """(caller's ids in the context's internal order, renumbered?)."""
        order = np.empty(self.n, dtype=np.uint32)
        r = ctypes.c_int()
        _check(self.lib.phde_vertex_order(self.h, as_ptr(order, ctypes.c_uint32), ctypes.byref(r)),
               "phde_vertex_order")
        return order, bool(r.value)

    def num_parts(self) -> int:
        """Parts of the partitioned traversal (0: not in use)."""
        k = ctypes.c_int64()
        _check(self.lib.phde_num_parts(self.h, ctypes.byref(k)), "phde_num_parts")
        return k.value

    def num_boundary(self) -> int:
        """Boundary vertices of the boundary-distance traversal (-1: not in use)."""
        k = ctypes.c_int64()
        _check(self.lib.phde_num_boundary(self.h, ctypes.byref(k)), "phde_num_boundary")
        return k.value

    def sssp(self, source: int, delta: int = 0) -> np.ndarray:
        """Exact distances (DeltaStep, gapbs/sssp.cc:53); 2^30-1 when unreached."""
        d = np.empty(self.n, dtype=np.int32)
        _check(self.lib.phde_sssp(self.h, source, delta, as_ptr(d, ctypes.c_int32)), "phde_sssp")
        return d

    def msbfs(self, sources: Sequence[int]) -> np.ndarray:
        """Up to 64 BFS in one bit-parallel sweep; n x k column-major levels."""
        src = np.ascontiguousarray(sources, dtype=np.int64)
        k = len(src)
        d = np.empty(self.n * k, dtype=np.int32)
        _check(self.lib.phde_msbfs(self.h, as_ptr(src, ctypes.c_int64), k, as_ptr(d, ctypes.c_int32)),
               "phde_msbfs")
        return d.reshape(k, self.n).T

    # --------------------------------------------------------------- layout
    def layout(self, options: Optional[PhdeOptions] = None, **kw):
        """One layout; returns (coords n x c float64, LayoutStats).  The
        coordinates are the library's column-major buffer viewed as n x c
        (Fortran order, no copy)."""
        o = options if options is not None else default_options(**kw)
        if options is not None:
            for k, v in kw.items():
                setattr(o, k, v)
        X = np.empty(self.n * o.num_coord, dtype=np.float64)
        st = PhdeStats()
        _check(self.lib.phde_layout(self.h, ctypes.byref(o), as_ptr(X, ctypes.c_double), ctypes.byref(st)),
               "phde_layout")
        return X.reshape(o.num_coord, self.n).T, LayoutStats.from_c(st, o.s)

    def power_eig(self, K: np.ndarray):
        """Reference power method on one s x s matrix (column-major semantics)."""
        dim = K.shape[0]
        Kf = np.ascontiguousarray(K.T.reshape(-1), dtype=np.float64).copy()  # column-major buffer
        ev = np.empty(dim, dtype=np.float64)
        it = ctypes.c_int32()
        _check(self.lib.phde_power_eig(self.h, as_ptr(Kf, ctypes.c_double), as_ptr(ev, ctypes.c_double),
                                       dim, ctypes.byref(it)), "phde_power_eig")
        return Kf, ev, it.value

    def sym_eig(self, A: np.ndarray, B: Optional[np.ndarray] = None):
        dim = A.shape[0]
        Af = np.ascontiguousarray(A.T, dtype=np.float64)
        Bp = None
        if B is not None:
            Bf = np.ascontiguousarray(B.T, dtype=np.float64)
            Bp = as_ptr(Bf, ctypes.c_double)
        ev = np.empty(dim)
        V = np.empty(dim * dim)
        _check(self.lib.phde_sym_eig(self.h, as_ptr(Af, ctypes.c_double), Bp, dim,
                                     as_ptr(ev, ctypes.c_double), as_ptr(V, ctypes.c_double)), "phde_sym_eig")
        return ev, V.reshape(dim, dim).T.copy()

    # -------------------------------------------------------- learned pivots
    def kcenters(self, s: int, first: int = 0):
        """Plain k-centers: (sources[s], final distance-to-nearest-source field)."""
        src = np.empty(s, dtype=np.int64)
        md = np.empty(self.n, dtype=np.int32)
        _check(self.lib.phde_kcenters(self.h, s, first, as_ptr(src, ctypes.c_int64), as_ptr(md, ctypes.c_int32)),
               "phde_kcenters")
        return src, md

    def pivot_features(self, first: int = 0):
        """(n x 4 float32 features, the two traversal sources)."""
        f = np.empty(self.n * 4, dtype=np.float32)
        src = np.empty(2, dtype=np.int64)
        _check(self.lib.phde_pivot_features(self.h, first, as_ptr(f, ctypes.c_float), as_ptr(src, ctypes.c_int64)),
               "phde_pivot_features")
        return f.reshape(self.n, 4), [int(x) for x in src]

    def set_training_data(self, feat: np.ndarray, target: np.ndarray) -> None:
        f = np.ascontiguousarray(feat, dtype=np.float32).reshape(-1)
        t = np.ascontiguousarray(target, dtype=np.float32).reshape(-1)
        if f.size != 4 * self.n or t.size != self.n:
            raise ValueError("features must be n x 4 and the target n")
        self._train = (f, t)
        _check(self.lib.phde_set_training_data(self.h, as_ptr(f, ctypes.c_float), as_ptr(t, ctypes.c_float)),
               "phde_set_training_data")

    def set_pivot_net(self, net: Optional["PivotNet"]) -> None:
        self._net = net
        _check(self.lib.phde_set_pivot_net(self.h, net.h if net is not None else None), "phde_set_pivot_net")

    def select_pivots(self, score: np.ndarray, k: int, radius: int, first: int = 0) -> np.ndarray:
        sc = np.ascontiguousarray(score, dtype=np.float32)
        out = np.empty(k, dtype=np.int64)
        _check(self.lib.phde_select_pivots(self.h, as_ptr(sc, ctypes.c_float), first, k, radius,
                                           as_ptr(out, ctypes.c_int64)), "phde_select_pivots")
        return out

    def hall_energy(self, Y: np.ndarray) -> float:
        Y = np.asarray(Y, dtype=np.float64)
        if Y.ndim == 1:
            Y = Y[:, None]
        Yc = np.ascontiguousarray(Y.T.reshape(-1))
        e = ctypes.c_double()
        _check(self.lib.phde_hall_energy(self.h, as_ptr(Yc, ctypes.c_double), Y.shape[1], ctypes.byref(e)),
               "phde_hall_energy")
        return e.value


class PivotNet:
    """The learned pivot-proposal network (phde_net_*): 3 message-passing
    layers of width 64 and a linear read-out, trained on the GPU."""

    def __init__(self, seed: int = 1, device: int = 0, path: Optional[str] = None):
        self.lib = _native.parhde_lib()
        h = ctypes.c_void_p()
        _check(self.lib.phde_net_create(device, seed, ctypes.byref(h)), "phde_net_create")
        self.h = h
        if path is not None:
            self.load(path)

    @staticmethod
    def num_params() -> int:
        return int(_native.parhde_lib().phde_net_num_params())

    def close(self):
        if getattr(self, "h", None):
            self.lib.phde_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def params(self) -> np.ndarray:
        p = np.empty(self.num_params(), dtype=np.float32)
        _check(self.lib.phde_net_get_params(self.h, as_ptr(p, ctypes.c_float)), "phde_net_get_params")
        return p

    @params.setter
    def params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.num_params():
            raise ValueError("wrong parameter count")
        _check(self.lib.phde_net_set_params(self.h, as_ptr(p, ctypes.c_float)), "phde_net_set_params")

    def save(self, path: str) -> None:
        _check(self.lib.phde_net_save(self.h, str(path).encode()), "phde_net_save")

    def load(self, path: str) -> None:
        _check(self.lib.phde_net_load(self.h, str(path).encode()), "phde_net_load")

    def scores(self, ctx: Context) -> np.ndarray:
        y = np.empty(ctx.n, dtype=np.float32)
        _check(self.lib.phde_net_scores(self.h, ctx.h, as_ptr(y, ctypes.c_float)), "phde_net_scores")
        return y

    def gradient(self, ctx: Context):
        """(mean squared error, gradient) on the context's training data."""
        g = np.empty(self.num_params(), dtype=np.float32)
        loss = ctypes.c_double()
        _check(self.lib.phde_net_gradient(self.h, ctx.h, ctypes.byref(loss), as_ptr(g, ctypes.c_float)),
               "phde_net_gradient")
        return loss.value, g

    def train_step(self, ctx: Context, lr: float) -> float:
        loss = ctypes.c_double()
        _check(self.lib.phde_net_train_step(self.h, ctx.h, lr, ctypes.byref(loss)), "phde_net_train_step")
        return loss.value


class DrawingBase:
    """Mirror of the reference's DrawingBase (variants/DrawingBase.h:13-48)."""

    algorithm = ALGO_PARHDE

    def __init__(self, rowPtrs, adj, weights, n: int, m: int, numSources: int, doRand: int = 2,
                 numCoord: int = 2):
        self.graph = Graph(int(n), int(m), np.asarray(rowPtrs, np.uint32), np.asarray(adj, np.uint32),
                           None if weights is None else np.asarray(weights, np.int32))
        self.n, self.m = int(n), int(m)
        self.numSources, self.numCoord, self.doRand = numSources, numCoord, doRand
        self.ctx = Context(self.graph)
        strategy = {2: PIVOTS_KCENTERS, 1: PIVOTS_RANDFINE, 0: PIVOTS_RANDCOARSE}.get(doRand, PIVOTS_KCENTERS)
        self.options = default_options(s=numSources, num_coord=numCoord, strategy=strategy,
                                       algorithm=self.algorithm)
        self.coordinates = np.zeros((self.n, numCoord))
        self.stats: Optional[LayoutStats] = None

    def run(self, tries: int = 1, alpha: int = 15, beta: int = 18, delta: int = 1) -> int:
        self.options.alpha = alpha
        if delta > 1:
            self.options.delta = delta
        for _ in range(max(tries, 1)):
            X, self.stats = self.ctx.layout(self.options)
        self.coordinates = X
        return 0

    def getCoordinates(self) -> np.ndarray:
        # DrawingBase.cpp:196-205: columns normalised in place on every call
        self.coordinates = self.coordinates / np.linalg.norm(self.coordinates, axis=0)
        return self.coordinates


class ParHDE(DrawingBase):
    algorithm = ALGO_PARHDE

    def __init__(self, rowPtrs, adj, weights, n, m, s, r=2, numCoord=2):
        if weights is not None:
            raise PhdeError(-6, "ParHDE code doesn't requires the weighted edges")
        super().__init__(rowPtrs, adj, None, n, m, s, r, numCoord)


class WParHDE(DrawingBase):
    algorithm = ALGO_WPARHDE

    def __init__(self, rowPtrs, adj, weights, n, m, s, numCoord=2):
        if weights is None:
            raise PhdeError(-6, "Weighted ParHDE code requires the weighted edges")
        super().__init__(rowPtrs, adj, weights, n, m, s, 2, numCoord)


class PivotMDS(DrawingBase):
    algorithm = ALGO_PIVOTMDS

    def __init__(self, rowPtrs, adj, weights, n, m, s, r=2, numCoord=2):
        super().__init__(rowPtrs, adj, weights, n, m, s, r, numCoord)


class PHDE(DrawingBase):
    algorithm = ALGO_PHDE

    def __init__(self, rowPtrs, adj, weights, n, m, s, r=2, numCoord=2):
        super().__init__(rowPtrs, adj, weights, n, m, s, r, numCoord)
