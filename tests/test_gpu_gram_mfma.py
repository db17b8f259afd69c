"""The Gram products on the FP64 matrix cores against the LDS-tiled kernel
(PHDE_GRAM_MFMA=0, read once per process, so it runs in a child process):
the layouts that use S^T(LS) (ParHDE), S^T D S (WParHDE, generalised
eigenproblem) and C^T C (PivotMDS), in FP64 and FP32 columns, agree to the
coordinate tolerance of the oracle parity tests (1e-8 relative in FP64,
1e-4 in FP32): the two kernels sum in different orders."""
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
    "parhde": ("rgg", dict(s=12, num_coord=2)),
    "wparhde_generalized": ("wgrid", dict(s=10, num_coord=3, algorithm=L.ALGO_WPARHDE, eig_mode=L.EIG_GENERALIZED)),
    # the power method stops at |<e_i, t_i>| >= 0.999, so PivotMDS coordinates
    # follow the last bits of C^T C; compared through the Jacobi eigensolver
    "pivotmds": ("rgg", dict(s=16, num_coord=2, algorithm=L.ALGO_PIVOTMDS, eig_mode=L.EIG_JACOBI)),
    "parhde_fp32": ("rgg", dict(s=12, num_coord=2, precision=L.PREC_FP32)),
}


def _graph(kind):
    if kind == "rgg":
        return Graph.rgg(20000, 10.0, seed=6)
    return Graph.grid3d(20, 18, 16, seed=4, wmax=255)


def _layout(case):
    kind, opts = CASES[case]
    with L.Context(_graph(kind)) as ctx:
        X, _ = ctx.layout(**opts)
    return np.ascontiguousarray(X)


CHILD = r"""
import sys
import numpy as np
sys.path.insert(0, {root!r})
sys.path.insert(0, {tests!r})
from test_gpu_gram_mfma import _layout
np.save({out!r}, _layout({case!r}))
"""


@pytest.mark.parametrize("case", list(CASES))
def test_mfma_gram_matches_tiled(gpu, case, tmp_path):
    X = _layout(case)
    out = str(tmp_path / "x.npy")
    code = CHILD.format(root=str(ROOT), tests=str(ROOT / "tests"), case=case, out=out)
    r = subprocess.run([sys.executable, "-c", code], env=dict(os.environ, PHDE_GRAM_MFMA="0"),
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    Y = np.load(out)
    tol = 1e-4 if CASES[case][1].get("precision") == L.PREC_FP32 else 1e-8
    for k in range(Y.shape[1]):
        x, y = X[:, k], Y[:, k]
        x = -x if np.dot(x, y) < 0 else x
        err = np.max(np.abs(x - y)) / np.max(np.abs(y))
        print(f"{case} column {k}: relative difference {err:.3e}")
        assert err <= tol, (case, k, err)
