"""The batched classical Gram-Schmidt paths against the one-candidate-at-a-
time path, on graphs with and without discarded candidates:

- the blocked CGS (several candidates share the passes over the columns
  accepted before them, PHDE_CGS_GRAM=0): the same accepted columns and
  bit-identical coordinates;
- the Gram-space CGS (the default for FP64 ParHDE: exact integer Gram
  matrices of the distance columns, the projections on coefficient
  vectors): the same pivots and discards, coordinates within the FP64
  tolerance and every candidate's norm within 1e-9 relative;
- after the Gram-space CGS, S^T L S from the exact C^T L C and the
  coordinates as C (M V) (the default) against forming S and taking the
  SpMM path (PHDE_GRAM_TAIL=0): coordinates within the FP64 tolerance.

The switches are read once per process, so every variant runs in a child
process."""
import json
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from icpp_2020_parhde_amd import layout as L
from icpp_2020_parhde_amd.graphs import Graph

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]

CASES = {
    "rgg": ("rgg", 30000, 12, 2, False),
    "mesh_eager": ("mesh", 90, 20, 3, True),
    "path_discards": ("path", 400, 12, 2, False),
    "rgg_s50": ("rgg", 60000, 50, 2, False),
}


def _graph(kind, size):
    if kind == "rgg":
        return Graph.rgg(size, 10.0, seed=4)
    if kind == "mesh":
        return Graph.mesh(size, size + 7)
    return Graph.path(size)


def _layout(kind, size, s, c, eager):
    g = _graph(kind, size)
    with L.Context(g) as ctx:
        X, st = ctx.layout(s=s, num_coord=c, eager=1 if eager else 0)
    return X, st


CHILD = r"""
import json, sys
import numpy as np
sys.path.insert(0, {root!r})
sys.path.insert(0, {tests!r})
from test_gpu_cgs_block import _layout
X, st = _layout({kind!r}, {size}, {s}, {c}, {eager})
np.save({out!r}, X)
print(json.dumps({{"pivots": list(st.pivots), "ndisc": int(st.n_discarded),
                   "normdist": list(st.normdist), "ngram": int(st.n_gram)}}))
"""


def _run(case, tmp_path, tag, **env_over):
    kind, size, s, c, eager = CASES[case]
    out = str(tmp_path / f"{tag}.npy")
    env = dict(os.environ, **env_over)
    code = CHILD.format(root=str(ROOT), tests=str(ROOT / "tests"), kind=kind, size=size, s=s, c=c,
                        eager=eager, out=out)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    res["X"] = np.load(out)
    return res


@pytest.mark.parametrize("case", list(CASES))
def test_batched_cgs_against_one_at_a_time(gpu, case, tmp_path):
    one = _run(case, tmp_path, "one", PHDE_CGS_BLOCK="0")
    blk = _run(case, tmp_path, "blk", PHDE_CGS_GRAM="0")
    gram = _run(case, tmp_path, "gram")
    gram_s = _run(case, tmp_path, "gram_s", PHDE_GRAM_TAIL="0")
    assert one["ngram"] == 0 and blk["ngram"] == 0
    # the Gram path took the first min(s, candidates) candidates
    assert gram["ngram"] == min(CASES[case][2], len(gram["normdist"]))
    assert blk["pivots"] == one["pivots"] and blk["ndisc"] == one["ndisc"]
    assert np.array_equal(blk["X"], one["X"])
    assert gram["pivots"] == one["pivots"] and gram["ndisc"] == one["ndisc"]
    a, b = np.array(gram["normdist"]), np.array(one["normdist"])
    assert len(a) == len(b)
    kept = b >= 0.001
    assert np.all(a[~kept] < 0.001)
    assert np.max(np.abs(a[kept] / b[kept] - 1)) < 1e-9
    assert gram_s["pivots"] == gram["pivots"] and gram_s["normdist"] == gram["normdist"]
    for X, Y in ((gram["X"], one["X"]), (gram_s["X"], one["X"]), (gram["X"], gram_s["X"])):
        for k in range(Y.shape[1]):
            x = X[:, k] if np.dot(X[:, k], Y[:, k]) >= 0 else -X[:, k]
            assert np.max(np.abs(x - Y[:, k])) / np.max(np.abs(Y[:, k])) < 1e-8, k


def test_gram_cgs_declines_sums_beyond_2_53(gpu):
    """A path of 250,000 vertices: max(sum deg, n) * max(d)^2 passes 2^53, so
    the Gram matrices of its distance columns could not be summed exactly in
    FP64; the Gram-space CGS takes no candidate and the vector CGS gives the
    oracle's layout (with the oracle's discard of the far end's column,
    which equals the ones vector minus the first column)."""
    from oracle import pyoracle as O
    from test_gpu_core import FP64_TOL, aligned_rel_err

    g = Graph.path(250000)
    ref = O.parhde(g, 3, 2, lazy=True)
    with L.Context(g) as ctx:
        X, st = ctx.layout(s=3, num_coord=2)
    assert st.n_gram == 0
    assert st.n_discarded == 1
    assert list(st.pivots) == [int(p) for p in ref["pivots"][: len(st.pivots)]]
    assert aligned_rel_err(X, ref["coords"]) < FP64_TOL
