"""The Matrix Market converter (csrc/mtx.cpp, bin/mtx2csr) against the
reference's own mtx2csr.cpp, compiled from its source into oracle/_ref by
oracle/Makefile (byte-identical output files), and against hand-worked
graphs.  CPU only."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from icpp_2020_parhde_amd.graphs import Graph

ROOT = Path(__file__).resolve().parents[1]
OURS = ROOT / "icpp_2020_parhde_amd" / "bin" / "mtx2csr"
REF = ROOT / "oracle" / "_ref" / "mtx2csr"

FIXTURES = {
    # self loop, duplicate, two components of 3: the first found is kept
    "sym_pattern": """%%MatrixMarket matrix coordinate pattern symmetric
% a comment line
6 6 7
2 1
3 2
3 3
3 1
3 1
5 4
6 5
""",
    # general square: the graph of A + A^T, real weights dropped
    "general_real": """%%MatrixMarket matrix coordinate real general
4 4 4
1 2 0.5
2 1 3.0
3 4 1
2 3 2e-1
""",
    # rectangular: bipartite, columns are vertices 0..2, rows 3..4
    "bipartite": """%%MatrixMarket matrix coordinate pattern general
2 3 4
1 1
1 3
2 2
2 3
""",
    # integer weights: duplicates keep the smallest (signed) weight
    "integer": """%%MatrixMarket matrix coordinate integer symmetric
4 4 6
2 1 5
2 1 3
3 2 -2
4 3 7
3 2 4
4 1 1
""",
    # the second component is the larger one: ids relabelled in old order
    "lcc_second": """%%MatrixMarket matrix coordinate pattern symmetric
7 7 5
2 1
4 3
5 4
6 5
7 3
""",
}


def read_bin(path):
    b = Path(path).read_bytes()
    hdr = np.frombuffer(b[:48], dtype=np.int64)
    n, m = int(hdr[0]), int(hdr[1])
    rp = np.frombuffer(b[48:48 + 4 * (n + 1)], dtype=np.uint32)
    adj = np.frombuffer(b[48 + 4 * (n + 1):48 + 4 * (n + 1) + 4 * m], dtype=np.uint32)
    rest = b[48 + 4 * (n + 1) + 4 * m:]
    w = np.frombuffer(rest, dtype=np.int32) if rest else None
    return hdr, rp, adj, w


def convert(exe, mtx, out):
    return subprocess.run([str(exe), str(mtx), str(out)], capture_output=True, text=True)


@pytest.fixture(scope="module")
def built():
    if not OURS.exists():
        pytest.fail("bin/mtx2csr is not built: run __graft_entry__.build()")
    return OURS


def test_hand_worked(built, tmp_path):
    def run(name):
        f = tmp_path / f"{name}.mtx"
        f.write_text(FIXTURES[name])
        assert convert(built, f, tmp_path / f"{name}.csr").returncode == 0
        return read_bin(tmp_path / f"{name}.csr")

    hdr, rp, adj, w = run("sym_pattern")
    assert hdr.tolist() == [3, 6, 1, 0, 1, 0] and w is None
    assert rp.tolist() == [0, 2, 4, 6] and adj.tolist() == [1, 2, 0, 2, 0, 1]
    hdr, rp, adj, w = run("general_real")
    assert hdr[0] == 4 and w is None
    assert rp.tolist() == [0, 1, 3, 5, 6] and adj.tolist() == [1, 0, 2, 1, 3, 2]
    hdr, rp, adj, w = run("bipartite")
    assert hdr[0] == 5
    assert rp.tolist() == [0, 1, 2, 4, 6, 8] and adj.tolist() == [3, 4, 3, 4, 0, 2, 1, 2]
    hdr, rp, adj, w = run("integer")
    assert rp.tolist() == [0, 2, 4, 6, 8] and adj.tolist() == [1, 3, 0, 2, 1, 3, 0, 2]
    assert w.tolist() == [3, 1, 3, -2, -2, 7, 1, 7]
    hdr, rp, adj, w = run("lcc_second")
    # component {2,3,4,5,6} (old ids) -> 0..4
    assert hdr[0] == 5 and rp.tolist() == [0, 2, 4, 6, 7, 8]
    assert adj.tolist() == [1, 4, 0, 2, 1, 3, 2, 0]


def test_graph_read_mtx_matches_file(built, tmp_path):
    f = tmp_path / "integer.mtx"
    f.write_text(FIXTURES["integer"])
    g = Graph.read_mtx(f)
    n, m = Graph.mtx2csr(f, tmp_path / "x.csr")
    h = Graph.read_csr(tmp_path / "x.csr", weighted=True)
    assert (g.n, g.m) == (n, m) == (h.n, h.m)
    assert np.array_equal(g.rowptr, h.rowptr) and np.array_equal(g.adj, h.adj) and np.array_equal(g.w, h.w)
    assert g.validate()


def test_rejects_malformed(built, tmp_path):
    bad = {"array": "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",
           "nonsquare_sym": "%%MatrixMarket matrix coordinate pattern symmetric\n2 3 1\n1 2\n",
           "zero_index": "%%MatrixMarket matrix coordinate pattern symmetric\n3 3 1\n0 2\n",
           "out_of_range": "%%MatrixMarket matrix coordinate pattern symmetric\n3 3 1\n4 2\n"}
    for name, text in bad.items():
        f = tmp_path / f"{name}.mtx"
        f.write_text(text)
        r = convert(built, f, tmp_path / f"{name}.csr")
        assert r.returncode != 0, name
        with pytest.raises(Exception):
            Graph.read_mtx(f)


def _random_mtx(path, rng, rows, cols, nnz, kind, sym):
    lines = [f"%%MatrixMarket matrix coordinate {kind} {'symmetric' if sym else 'general'}",
             "% generated", f"{rows} {cols} {nnz}"]
    u = rng.integers(1, rows + 1, nnz)
    v = rng.integers(1, cols + 1, nnz)
    if sym:
        u, v = np.maximum(u, v), np.minimum(u, v)
    for a, b in zip(u, v):
        if kind == "pattern":
            lines.append(f"{a} {b}")
        elif kind == "integer":
            lines.append(f"{a} {b} {int(rng.integers(-50, 1000))}")
        else:
            lines.append(f"{a} {b} {rng.standard_normal():.6g}")
    path.write_text("\n".join(lines) + "\n")


@pytest.mark.parametrize("case", [("pattern", True, 300, 300, 900), ("integer", True, 400, 400, 1500),
                                  ("real", False, 250, 250, 700), ("integer", False, 200, 200, 800),
                                  ("pattern", False, 120, 90, 400), ("integer", False, 60, 140, 500)])
def test_matches_reference_binary(built, tmp_path, case):
    if not REF.exists():
        pytest.skip("oracle/_ref/mtx2csr not built (needs the reference tree)")
    kind, sym, rows, cols, nnz = case
    rng = np.random.default_rng(rows * 7 + nnz)
    f = tmp_path / "r.mtx"
    _random_mtx(f, rng, rows, cols, nnz, kind, sym)
    for name, text in FIXTURES.items():
        g = tmp_path / f"{name}.mtx"
        g.write_text(text)
    for src in [f] + [tmp_path / f"{name}.mtx" for name in FIXTURES]:
        a, b = tmp_path / "ours.csr", tmp_path / "ref.csr"
        assert convert(built, src, a).returncode == 0
        assert convert(REF, src, b).returncode == 0
        assert a.read_bytes() == b.read_bytes(), src.name
