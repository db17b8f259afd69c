"""Writes the golden fixtures under tests/golden/ (run where the reference
tree exists; elsewhere the tests only read the fixtures).

power_eig_*.npz: inputs and outputs of the reference's own
eigenValuePowerMethod (variants/ParallelComponents.cpp:48-158), compiled from
the reference source into oracle/_ref/libref_pc.so by oracle/Makefile and
called with one OpenMP thread.  Inputs are the S^T(LS) matrices of the
128x128 mesh at s = 10 and s = 50 (as computed by the oracle) and seeded
symmetric test matrices.
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from icpp_2020_parhde_amd.graphs import Graph  # noqa: E402
from oracle import pyoracle as O  # noqa: E402


def main():
    out = Path(__file__).resolve().parent
    if O.ref_lib() is None:
        raise SystemExit("oracle/_ref/libref_pc.so not built (make -C oracle ref)")
    cases = {}
    g = Graph.mesh(128, 128)
    for s in (10, 50):
        r = O.parhde(g, s, 2, lazy=True)
        cases[f"mesh128_s{s}"] = r["SLS"]
    rng = np.random.default_rng(7)
    for dim in (3, 8, 20, 64):
        A = rng.standard_normal((dim, dim))
        A = A @ A.T + dim * np.eye(dim)
        cases[f"spd{dim}"] = np.ascontiguousarray(A.T.reshape(-1))
    for name, K in cases.items():
        Kout, ev, rc = O.ref_power_eig(K)
        np.savez(out / f"power_eig_{name}.npz", K_in=K, K_out=Kout, evals=ev, rc=rc)
        print(name, "rc", rc)


if __name__ == "__main__":
    main()
