"""Graphs in the reference's binary CSR layout, and seeded generators.

``Graph`` holds ``rowptr`` (uint32, n+1), ``adj`` (uint32, m) and optional
``w`` (int32, m) as numpy arrays: exactly the arrays the reference's
``embed`` reads (spectralDrawing.cpp:136-164) and ``mtx2csr`` writes
(mtx2csr.cpp:281-296).  Generators and file I/O run in libphde_graph.so.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _native
from ._native import PgCsr, as_ptr

# Synthetic workload parameters (seeds fixed so vertex numbering, and with it
# every k-centers tie, is reproducible).
SEED_RGG = 2
SEED_RMAT = 3
SEED_GRID3D = 4
SEED_RGG_LARGE = 5


@dataclass
class Graph:
    n: int
    m: int
    rowptr: np.ndarray
    adj: np.ndarray
    w: Optional[np.ndarray] = None
    name: str = ""

    # -------------------------------------------------------------- plumbing
    @staticmethod
    def _take(csr: PgCsr, name: str = "") -> "Graph":
        lib = _native.graph_lib()
        n, m = int(csr.n), int(csr.m)
        rowptr = np.ctypeslib.as_array(csr.rowptr, shape=(n + 1,)).copy()
        adj = np.ctypeslib.as_array(csr.adj, shape=(max(m, 1),))[:m].copy()
        w = None
        if csr.w:
            w = np.ctypeslib.as_array(csr.w, shape=(max(m, 1),))[:m].copy()
        lib.pg_free(ctypes.byref(csr))
        return Graph(n, m, rowptr, adj, w, name)

    def _csr(self) -> PgCsr:
        self.rowptr = np.ascontiguousarray(self.rowptr, dtype=np.uint32)
        self.adj = np.ascontiguousarray(self.adj, dtype=np.uint32)
        c = PgCsr()
        c.n, c.m = self.n, self.m
        c.rowptr = as_ptr(self.rowptr, ctypes.c_uint32)
        c.adj = as_ptr(self.adj, ctypes.c_uint32)
        if self.w is not None:
            self.w = np.ascontiguousarray(self.w, dtype=np.int32)
            c.w = as_ptr(self.w, ctypes.c_int32)
        return c

    @staticmethod
    def _check(rc: int, what: str):
        if rc != 0:
            raise RuntimeError(f"{what} failed with status {rc}")

    @property
    def degrees(self) -> np.ndarray:
        return np.diff(self.rowptr.astype(np.int64))

    @property
    def weighted(self) -> bool:
        return self.w is not None

    def unweighted(self) -> "Graph":
        return Graph(self.n, self.m, self.rowptr, self.adj, None, self.name)

    # ------------------------------------------------------------ generators
    @classmethod
    def mesh(cls, rows: int, cols: int) -> "Graph":
        """Triangulated mesh: edges right, down and down-right (i+1, j+1)."""
        c = PgCsr()
        cls._check(_native.graph_lib().pg_mesh(rows, cols, ctypes.byref(c)), "pg_mesh")
        return cls._take(c, f"mesh{rows}x{cols}")

    @classmethod
    def grid3d(cls, nx: int, ny: int, nz: int, seed: int = SEED_GRID3D, wmax: int = 255) -> "Graph":
        """6-neighbour 3D grid with int32 weights uniform in [1, wmax] (0: unweighted)."""
        c = PgCsr()
        cls._check(_native.graph_lib().pg_grid3d(nx, ny, nz, seed, wmax, ctypes.byref(c)), "pg_grid3d")
        return cls._take(c, f"grid{nx}x{ny}x{nz}")

    @classmethod
    def rgg(cls, n: int, avg_degree: float = 10.0, seed: int = SEED_RGG, lcc: bool = True,
            with_points: bool = False):
        """Random geometric graph in the unit square (largest component)."""
        c = PgCsr()
        xy = _native.c_f64p()
        rc = _native.graph_lib().pg_rgg(n, avg_degree, seed, 1 if lcc else 0, ctypes.byref(c),
                                        ctypes.byref(xy) if with_points else None)
        cls._check(rc, "pg_rgg")
        g = cls._take(c, f"rgg{n}")
        if not with_points:
            return g
        pts = np.ctypeslib.as_array(xy, shape=(2 * g.n,)).copy().reshape(g.n, 2)
        ctypes.CDLL(None).free(xy)
        return g, pts

    @classmethod
    def rmat(cls, scale: int, edgefactor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             seed: int = SEED_RMAT, lcc: bool = True) -> "Graph":
        """R-MAT graph, symmetrised, self loops and duplicates removed."""
        cs = PgCsr()
        rc = _native.graph_lib().pg_rmat(scale, edgefactor, a, b, c, seed, 1 if lcc else 0, ctypes.byref(cs))
        cls._check(rc, "pg_rmat")
        return cls._take(cs, f"rmat{scale}")

    @classmethod
    def holed_grid(cls, rows: int, cols: int, holes: int, seed: int) -> "Graph":
        c = PgCsr()
        cls._check(_native.graph_lib().pg_holed_grid(rows, cols, holes, seed, ctypes.byref(c)),
                   "pg_holed_grid")
        return cls._take(c, f"holedgrid{rows}x{cols}")

    @classmethod
    def from_edges(cls, n: int, u, v, w=None) -> "Graph":
        u = np.ascontiguousarray(u, dtype=np.uint32)
        v = np.ascontiguousarray(v, dtype=np.uint32)
        wp = None
        if w is not None:
            w = np.ascontiguousarray(w, dtype=np.int32)
            wp = as_ptr(w, ctypes.c_int32)
        c = PgCsr()
        rc = _native.graph_lib().pg_from_edges(n, len(u), as_ptr(u, ctypes.c_uint32),
                                               as_ptr(v, ctypes.c_uint32), wp, ctypes.byref(c))
        cls._check(rc, "pg_from_edges")
        return cls._take(c)

    @classmethod
    def path(cls, n: int) -> "Graph":
        i = np.arange(n - 1)
        return cls.from_edges(n, i, i + 1)

    @classmethod
    def cycle(cls, n: int) -> "Graph":
        i = np.arange(n)
        return cls.from_edges(n, i, (i + 1) % n)

    @classmethod
    def star(cls, n: int) -> "Graph":
        i = np.arange(1, n)
        return cls.from_edges(n, np.zeros(n - 1), i)

    def lcc(self) -> "Graph":
        c = PgCsr()
        ncomp = ctypes.c_int64()
        src = self._csr()
        self._check(_native.graph_lib().pg_lcc(ctypes.byref(src), ctypes.byref(c), ctypes.byref(ncomp)),
                    "pg_lcc")
        return self._take(c, self.name)

    # ------------------------------------------------------------------ I/O
    @classmethod
    def read_csr(cls, path: str, weighted: bool = False) -> "Graph":
        c = PgCsr()
        rc = _native.graph_lib().pg_read_csr(str(path).encode(), 1 if weighted else 0, ctypes.byref(c))
        cls._check(rc, "pg_read_csr")
        return cls._take(c, str(path))

    @classmethod
    def read_mtx(cls, path: str, lcc: bool = True) -> "Graph":
        """Matrix Market coordinate file, read as the reference's mtx2csr.cpp
        reads it (weights kept for integer files), then its largest
        component unless lcc=False."""
        c = PgCsr()
        kind = ctypes.c_int()
        cls._check(_native.graph_lib().pg_read_mtx(str(path).encode(), ctypes.byref(c), ctypes.byref(kind)),
                   "pg_read_mtx")
        g = cls._take(c, str(path))
        return g.lcc() if lcc else g

    @staticmethod
    def mtx2csr(mtx_path: str, csr_path: str):
        """The reference converter's whole job (mtx2csr.cpp): Matrix Market in,
        binary CSR of the largest component out; returns (n, m) written."""
        n, m = ctypes.c_int64(), ctypes.c_int64()
        Graph._check(_native.graph_lib().pg_mtx2csr(str(mtx_path).encode(), str(csr_path).encode(),
                                                     ctypes.byref(n), ctypes.byref(m)), "pg_mtx2csr")
        return n.value, m.value

    def draw_png(self, coords, path: str, first: int = 0, second: int = 1):
        """PNG of the edges at `coords` (n x c), x / y from columns first /
        second, drawn as the reference's draw_graph.c draws them; returns
        (width, height)."""
        X = np.ascontiguousarray(np.asarray(coords, dtype=np.float64).T)  # column-major n x c
        w, h = ctypes.c_uint(), ctypes.c_uint()
        c = self._csr()
        self._check(_native.graph_lib().pg_draw_png(ctypes.byref(c), X.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                    X.shape[0], first, second, str(path).encode(),
                                                    ctypes.byref(w), ctypes.byref(h)), "pg_draw_png")
        return w.value, h.value

    def write_csr(self, path: str) -> None:
        c = self._csr()
        self._check(_native.graph_lib().pg_write_csr(str(path).encode(), ctypes.byref(c)), "pg_write_csr")

    def validate(self) -> bool:
        c = self._csr()
        return _native.graph_lib().pg_validate(ctypes.byref(c)) == 0
