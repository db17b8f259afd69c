"""Blocked classical Gram-Schmidt (several candidates share the passes over
the columns accepted before them) against the one-candidate-at-a-time
path: the same accepted columns and bit-identical coordinates, on graphs
with and without discarded candidates.  PHDE_CGS_BLOCK is read once per
process, so the one-at-a-time reference runs in a child process."""
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
print(json.dumps({{"pivots": list(st.pivots), "ndisc": int(st.n_discarded)}}))
"""


@pytest.mark.parametrize("case", list(CASES))
def test_blocked_cgs_equals_one_at_a_time(gpu, case, tmp_path):
    kind, size, s, c, eager = CASES[case]
    X, st = _layout(kind, size, s, c, eager)
    out = str(tmp_path / "x.npy")
    env = dict(os.environ, PHDE_CGS_BLOCK="0")
    code = CHILD.format(root=str(ROOT), tests=str(ROOT / "tests"), kind=kind, size=size, s=s, c=c,
                        eager=eager, out=out)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    ref = json.loads(r.stdout.strip().splitlines()[-1])
    assert list(st.pivots) == ref["pivots"]
    assert int(st.n_discarded) == ref["ndisc"]
    assert np.array_equal(X, np.load(out))
