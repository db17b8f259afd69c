"""The PNG renderer (csrc/draw_png.cpp, bin/draw) against the reference's
draw_graph.c + lodepng.c, compiled from their sources into oracle/_ref by
oracle/Makefile: identical decoded RGBA pixels on the same CSR and
coordinate files.  CPU only."""
import subprocess
import zlib
from pathlib import Path

import numpy as np
import pytest

from icpp_2020_parhde_amd.graphs import Graph

ROOT = Path(__file__).resolve().parents[1]
OURS = ROOT / "icpp_2020_parhde_amd" / "bin" / "draw"
REF = ROOT / "oracle" / "_ref" / "draw_graph"


def decode_png(path):
    """RGBA pixels of a non-interlaced PNG (grey, RGB, palette, grey+alpha or
    RGBA; 1-8 bits per sample)."""
    b = Path(path).read_bytes()
    assert b[:8] == bytes([137, 80, 78, 71, 13, 10, 26, 10])
    pos, idat, plte, trns = 8, b"", None, None
    while pos < len(b):
        ln = int.from_bytes(b[pos:pos + 4], "big")
        typ, data = b[pos + 4:pos + 8], b[pos + 8:pos + 8 + ln]
        assert zlib.crc32(typ + data) == int.from_bytes(b[pos + 8 + ln:pos + 12 + ln], "big")
        pos += 12 + ln
        if typ == b"IHDR":
            w, h = int.from_bytes(data[:4], "big"), int.from_bytes(data[4:8], "big")
            depth, ctype, interlace = data[8], data[9], data[12]
            assert interlace == 0 and depth <= 8
        elif typ == b"PLTE":
            plte = np.frombuffer(data, np.uint8).reshape(-1, 3)
        elif typ == b"tRNS":
            trns = np.frombuffer(data, np.uint8)
        elif typ == b"IDAT":
            idat += data
    spp = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    bpp = max(1, spp * depth // 8)
    stride = (w * spp * depth + 7) // 8
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, stride + 1)
    out = np.zeros((h, stride), np.int64)
    prev = np.zeros(stride, np.int64)
    for y in range(h):
        f, line = raw[y, 0], raw[y, 1:].astype(np.int64)
        cur = np.zeros(stride, np.int64)
        for x in range(stride):
            a = cur[x - bpp] if x >= bpp else 0
            c = prev[x - bpp] if x >= bpp else 0
            bb = prev[x]
            if f == 0:
                pr = 0
            elif f == 1:
                pr = a
            elif f == 2:
                pr = bb
            elif f == 3:
                pr = (a + bb) // 2
            else:
                p = a + bb - c
                pa, pb, pc = abs(p - a), abs(p - bb), abs(p - c)
                pr = a if pa <= pb and pa <= pc else (bb if pb <= pc else c)
            cur[x] = (line[x] + pr) & 255
        out[y] = cur
        prev = cur
    bits = np.unpackbits(out.astype(np.uint8), axis=1) if depth < 8 else None
    if depth < 8:
        samples = np.array([[int("".join(map(str, bits[y, i * depth:(i + 1) * depth])), 2)
                             for i in range(w * spp)] for y in range(h)])
    else:
        samples = out[:, :w * spp]
    s = samples.reshape(h, w, spp)
    rgba = np.zeros((h, w, 4), np.int64)
    rgba[..., 3] = 255
    if ctype == 3:
        rgba[..., :3] = plte[s[..., 0]]
        if trns is not None:
            alpha = np.full(len(plte), 255)
            alpha[:len(trns)] = trns
            rgba[..., 3] = alpha[s[..., 0]]
    elif ctype in (0, 4):
        scale = 255 // (2 ** depth - 1)
        rgba[..., :3] = s[..., :1] * scale
        if ctype == 4:
            rgba[..., 3] = s[..., 1]
    else:
        rgba[..., :spp] = s
    return rgba


def _files(tmp_path, g, X):
    csr, xyz = tmp_path / "g.csr", tmp_path / "g.nxyz"
    g.write_csr(csr)
    xyz.write_text("\n".join(",".join(f"{v:.17g}" for v in row) for row in X) + "\n")
    return csr, xyz


@pytest.mark.parametrize("name", ["mesh", "rgg", "rmat"])
def test_draw_matches_reference(tmp_path, name):
    if not OURS.exists():
        pytest.fail("bin/draw is not built: run __graft_entry__.build()")
    if not REF.exists():
        pytest.skip("oracle/_ref/draw_graph not built (needs the reference tree)")
    g = {"mesh": lambda: Graph.mesh(24, 30), "rgg": lambda: Graph.rgg(1500, 8.0, seed=3),
         "rmat": lambda: Graph.rmat(9, 8, seed=4)}[name]()
    rng = np.random.default_rng(7)
    # three coordinate columns in (-1, 1), drawn with columns 1 and 3
    X = rng.uniform(-0.9, 0.9, size=(g.n, 3))
    csr, xyz = _files(tmp_path, g, X)
    for exe, out in ((OURS, "ours.png"), (REF, "ref.png")):
        r = subprocess.run([str(exe), str(csr), str(xyz), str(tmp_path / out), "3", "1", "3"],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
    a, b = decode_png(tmp_path / "ours.png"), decode_png(tmp_path / "ref.png")
    assert a.shape == b.shape and np.array_equal(a, b)
    # not a blank picture: the background and several edge colours
    assert len(np.unique(a.reshape(-1, 4), axis=0)) >= 3


def test_graph_draw_png_matches_command(tmp_path):
    g = Graph.mesh(12, 15)
    X = np.random.default_rng(1).uniform(-0.5, 0.5, size=(g.n, 2))
    csr, xyz = _files(tmp_path, g, X)
    w, h = g.draw_png(X, tmp_path / "py.png")
    assert subprocess.run([str(OURS), str(csr), str(xyz), str(tmp_path / "cli.png"), "2", "1", "2"],
                          capture_output=True).returncode == 0
    a, b = decode_png(tmp_path / "py.png"), decode_png(tmp_path / "cli.png")
    assert a.shape == (h, w, 4) and np.array_equal(a, b)
