"""Loading of the in-tree native libraries and their C-ABI signatures.

libparhde.so (HIP, gfx950) and libphde_graph.so (host) are built in-tree by
``csrc/Makefile`` (driven by ``__graft_entry__.build()``).  Nothing here
falls back to another implementation: a missing library raises.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

LIBDIR = Path(__file__).resolve().parent / "lib"
BINDIR = Path(__file__).resolve().parent / "bin"

c_u32p = ctypes.POINTER(ctypes.c_uint32)
c_i32p = ctypes.POINTER(ctypes.c_int32)
c_i64p = ctypes.POINTER(ctypes.c_int64)
c_f64p = ctypes.POINTER(ctypes.c_double)
c_f32p = ctypes.POINTER(ctypes.c_float)


class PgCsr(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64),
        ("m", ctypes.c_int64),
        ("rowptr", c_u32p),
        ("adj", c_u32p),
        ("w", c_i32p),
    ]


class PhdeOptions(ctypes.Structure):
    _fields_ = [
        ("algorithm", ctypes.c_int32),
        ("s", ctypes.c_int32),
        ("num_coord", ctypes.c_int32),
        ("strategy", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("eig_mode", ctypes.c_int32),
        ("eager", ctypes.c_int32),
        ("delta", ctypes.c_int32),
        ("alpha", ctypes.c_int32),
        ("beta", ctypes.c_int32),
        ("verbose", ctypes.c_int32),
        ("first_pivot", ctypes.c_int64),
        ("seed", ctypes.c_uint64),
        ("judge_tol", ctypes.c_double),
        ("nms_rho", ctypes.c_double),
        ("judge", ctypes.c_int32),
        ("diagnostics", ctypes.c_int32),
    ]


class PhdeStats(ctypes.Structure):
    _fields_ = [
        ("ms_total", ctypes.c_double),
        ("ms_distance", ctypes.c_double),
        ("ms_orth", ctypes.c_double),
        ("ms_mmult", ctypes.c_double),
        ("ms_eig", ctypes.c_double),
        ("ms_coords", ctypes.c_double),
        ("n_traversals", ctypes.c_int32),
        ("n_accepted", ctypes.c_int32),
        ("n_discarded", ctypes.c_int32),
        ("eig_iters", ctypes.c_int32),
        ("levels_total", ctypes.c_int64),
        ("pivots", ctypes.c_int64 * 192),
        ("kernel_ms", ctypes.c_double * 16),
        ("kernel_launches", ctypes.c_int64 * 16),
        ("ms_select", ctypes.c_double),
        ("hall_energy", ctypes.c_double),
        ("hall_energy_kcenters", ctypes.c_double),
        ("fallback", ctypes.c_int32),
        ("n_candidates", ctypes.c_int32),
        ("n_examined", ctypes.c_int32),
        ("n_gram", ctypes.c_int32),
        ("normdist", ctypes.c_double * 192),
        ("eigenvalues", ctypes.c_double * 64),
        ("dorth_err", ctypes.c_double),
    ]


# Symbols declared in include/parhde.h and include/phde_graph.h, with their
# ctypes signatures (restype, argtypes).
PARHDE_SYMBOLS = {
    "phde_default_options": (None, [ctypes.POINTER(PhdeOptions)]),
    "phde_version": (ctypes.c_char_p, []),
    "phde_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "phde_create": (ctypes.c_int, [c_u32p, c_u32p, c_i32p, ctypes.c_int64, ctypes.c_int64,
                                   ctypes.POINTER(ctypes.c_void_p)]),
    "phde_create_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_int64, ctypes.c_int64,
                                          ctypes.POINTER(ctypes.c_void_p)]),
    "phde_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "phde_num_vertices": (ctypes.c_int64, [ctypes.c_void_p]),
    "phde_vertex_order": (ctypes.c_int, [ctypes.c_void_p, c_u32p, ctypes.POINTER(ctypes.c_int)]),
    "phde_bfs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, c_i32p]),
    "phde_bfs_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_void_p]),
    "phde_sssp": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, c_i32p]),
    "phde_msbfs": (ctypes.c_int, [ctypes.c_void_p, c_i64p, ctypes.c_int, c_i32p]),
    "phde_layout": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PhdeOptions), c_f64p,
                                   ctypes.POINTER(PhdeStats)]),
    "phde_power_eig": (ctypes.c_int, [ctypes.c_void_p, c_f64p, c_f64p, ctypes.c_int, c_i32p]),
    "phde_num_parts": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "phde_num_boundary": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "phde_power_eig_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                              ctypes.c_void_p]),
    "phde_sym_eig": (ctypes.c_int, [ctypes.c_void_p, c_f64p, c_f64p, ctypes.c_int, c_f64p, c_f64p]),
    "phde_hall_energy": (ctypes.c_int, [ctypes.c_void_p, c_f64p, ctypes.c_int, c_f64p]),
    "phde_last_error": (ctypes.c_char_p, []),
    "phde_kcenters": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, c_i64p, c_i32p]),
    # learned pivot proposal
    "phde_net_num_params": (ctypes.c_int, []),
    "phde_net_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]),
    "phde_net_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "phde_net_get_params": (ctypes.c_int, [ctypes.c_void_p, c_f32p]),
    "phde_net_set_params": (ctypes.c_int, [ctypes.c_void_p, c_f32p]),
    "phde_net_save": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "phde_net_load": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "phde_set_pivot_net": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "phde_pivot_features": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, c_f32p, c_i64p]),
    "phde_set_training_data": (ctypes.c_int, [ctypes.c_void_p, c_f32p, c_f32p]),
    "phde_net_scores": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, c_f32p]),
    "phde_net_gradient": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, c_f64p, c_f32p]),
    "phde_net_adam": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float]),
    "phde_net_train_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, c_f64p]),
    "phde_select_pivots": (ctypes.c_int, [ctypes.c_void_p, c_f32p, ctypes.c_int64, ctypes.c_int,
                                          ctypes.c_int, c_i64p]),
}

GRAPH_SYMBOLS = {
    "pg_free": (None, [ctypes.POINTER(PgCsr)]),
    "pg_mesh": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(PgCsr)]),
    "pg_grid3d": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_uint64,
                                 ctypes.c_int32, ctypes.POINTER(PgCsr)]),
    "pg_rgg": (ctypes.c_int, [ctypes.c_int64, ctypes.c_double, ctypes.c_uint64, ctypes.c_int,
                              ctypes.POINTER(PgCsr), ctypes.POINTER(c_f64p)]),
    "pg_rmat": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                               ctypes.c_double, ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(PgCsr)]),
    "pg_holed_grid": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_uint64,
                                     ctypes.POINTER(PgCsr)]),
    "pg_from_edges": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, c_u32p, c_u32p, c_i32p,
                                     ctypes.POINTER(PgCsr)]),
    "pg_lcc": (ctypes.c_int, [ctypes.POINTER(PgCsr), ctypes.POINTER(PgCsr), c_i64p]),
    "pg_write_csr": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(PgCsr)]),
    "pg_read_csr": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(PgCsr)]),
    "pg_validate": (ctypes.c_int, [ctypes.POINTER(PgCsr)]),
    "pg_read_mtx": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(PgCsr), ctypes.POINTER(ctypes.c_int)]),
    "pg_mtx2csr": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, c_i64p, c_i64p]),
    "pg_draw_png": (ctypes.c_int, [ctypes.POINTER(PgCsr), c_f64p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint), ctypes.POINTER(ctypes.c_uint)]),
}

_cache: dict[str, ctypes.CDLL] = {}


def _load(name: str, symbols: dict) -> ctypes.CDLL:
    if name in _cache:
        return _cache[name]
    path = LIBDIR / name
    if not path.exists():
        raise RuntimeError(
            f"{path} is not built; run __graft_entry__.build() (make -C icpp_2020_parhde_amd/csrc)")
    lib = ctypes.CDLL(str(path), mode=os.RTLD_NOW | os.RTLD_LOCAL)
    for sym, (res, args) in symbols.items():
        f = getattr(lib, sym)
        f.restype = res
        f.argtypes = args
    _cache[name] = lib
    return lib


def parhde_lib() -> ctypes.CDLL:
    """libparhde.so: the HIP layout library (raises when not built)."""
    return _load("libparhde.so", PARHDE_SYMBOLS)


def graph_lib() -> ctypes.CDLL:
    """libphde_graph.so: host graph generators and CSR I/O."""
    return _load("libphde_graph.so", GRAPH_SYMBOLS)


def as_ptr(arr, ctype):
    return arr.ctypes.data_as(ctypes.POINTER(ctype))
