"""phde_power_eig_noalloc, the entry point that allocates nothing (its
scratch is a static device array of the library), against phde_power_eig
on the same matrices: bit-identical eigenvectors, eigenvalues and
iteration counts, on two calls in a row, the second with a smaller matrix
(the scratch keeps the first call's contents).  Device buffers are made
here with the HIP runtime; the entry point itself allocates none."""
import ctypes
from pathlib import Path

import numpy as np
import pytest

from icpp_2020_parhde_amd import layout as L
from icpp_2020_parhde_amd.graphs import Graph

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


class DevBuf:
    """hipMalloc'd buffer owned by the test."""

    hip = None

    def __init__(self, nbytes):
        if DevBuf.hip is None:
            DevBuf.hip = ctypes.CDLL("libamdhip64.so")
        self.p = ctypes.c_void_p()
        assert DevBuf.hip.hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(nbytes)) == 0
        self.nbytes = nbytes

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert DevBuf.hip.hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1) == 0

    def download(self, a):
        assert DevBuf.hip.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), self.p, ctypes.c_size_t(a.nbytes), 2) == 0
        return a

    def free(self):
        DevBuf.hip.hipFree(self.p)


def column_major(name):
    f = np.load(GOLDEN / f"power_eig_{name}.npz")
    dim = int(round(len(f["K_in"]) ** 0.5))
    return f["K_in"].reshape(dim, dim).T, dim


def test_power_eig_noalloc_matches_allocating_entry_point(gpu):
    lib = gpu
    with L.Context(Graph.path(4)) as ctx:
        for name in ("mesh128_s50", "spd20"):  # second call smaller than the first
            K, dim = column_major(name)
            K_ref, ev_ref, it_ref = ctx.power_eig(K)
            Kd, evd, std = DevBuf(8 * dim * dim), DevBuf(8 * dim), DevBuf(8)
            try:
                Kd.upload(np.ascontiguousarray(K.T.reshape(-1)))  # column-major buffer
                rc = lib.phde_power_eig_noalloc(ctx.h, Kd.p, evd.p, dim, std.p)
                assert rc == 0, L._native.parhde_lib().phde_last_error()
                Kout = Kd.download(np.empty(dim * dim))
                ev = evd.download(np.empty(dim))
                status = std.download(np.empty(2, np.int32))
            finally:
                for b in (Kd, evd, std):
                    b.free()
            assert status[0] == 0 and status[1] == it_ref, (name, status, it_ref)
            assert np.array_equal(Kout, K_ref), name
            assert np.array_equal(ev, ev_ref), name


def test_power_eig_noalloc_rejects_oversize(gpu):
    lib = gpu
    with L.Context(Graph.path(4)) as ctx:
        b = DevBuf(8 * 65 * 65)
        try:
            assert lib.phde_power_eig_noalloc(ctx.h, b.p, b.p, 65, b.p) == -1
            assert b"outside 1..64" in lib.phde_last_error()
        finally:
            b.free()
