"""The built libraries load and export every symbol their headers declare.
CPU only: no compute call needs a GPU here."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

from icpp_2020_parhde_amd import _native

ROOT = Path(__file__).resolve().parents[1]


def declared(header):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:phde|pg)_[a-z0-9_]+)\s*\(", text)))


def test_parhde_exports_every_declared_symbol():
    lib = _native.parhde_lib()
    names = declared("parhde.h")
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
        assert n in _native.PARHDE_SYMBOLS, f"{n} has no ctypes signature"


def test_graph_lib_exports_every_declared_symbol():
    lib = _native.graph_lib()
    for n in declared("phde_graph.h"):
        assert hasattr(lib, n), n
        assert n in _native.GRAPH_SYMBOLS, n


def test_cxx_drop_in_symbols_exported():
    """DOBFS / DeltaStep keep the reference's C++ signatures (mangled names of
    variants/DrawingBase.h:62-67) and the classes their vtables."""
    out = subprocess.run(["nm", "-D", "--defined-only", str(_native.LIBDIR / "libparhde.so")],
                         capture_output=True, text=True, check=True).stdout
    for sym in ["_Z5DOBFSPjS_lllPdii", "_Z9DeltaStepPKjPjPilllPdi", "_ZTV6ParHDE", "_ZTV7WParHDE",
                "_ZTV8PivotMDS", "_ZTV4PHDE", "_ZN11DrawingBase14getCoordinatesEv",
                "_ZN18ParallelComponents21eigenValuePowerMethodEPdS0_li"]:
        assert sym in out, sym


def test_version_and_device_count_without_gpu():
    lib = _native.parhde_lib()
    assert b"gfx950" in lib.phde_version()
    c = ctypes.c_int(-1)
    assert lib.phde_device_count(ctypes.byref(c)) == 0
    assert c.value >= 0


def test_bad_arguments_rejected_on_host():
    """Malformed CSR is refused before anything reaches a device."""
    lib = _native.parhde_lib()
    import numpy as np
    h = ctypes.c_void_p()
    rp = np.array([0, 2, 1], np.uint32)  # decreasing row pointer
    adj = np.array([1, 0], np.uint32)
    rc = lib.phde_create(_native.as_ptr(rp, ctypes.c_uint32), _native.as_ptr(adj, ctypes.c_uint32), None,
                         2, 2, ctypes.byref(h))
    assert rc == -1
    rp = np.array([0, 1, 2], np.uint32)
    adj = np.array([1, 5], np.uint32)  # id out of range
    rc = lib.phde_create(_native.as_ptr(rp, ctypes.c_uint32), _native.as_ptr(adj, ctypes.c_uint32), None,
                         2, 2, ctypes.byref(h))
    assert rc == -1
    assert lib.phde_layout(None, None, None, None) == -1


def test_embed_binary_built():
    exe = _native.BINDIR / "embed"
    assert exe.exists()
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 1 and "Usage" in out.stderr
