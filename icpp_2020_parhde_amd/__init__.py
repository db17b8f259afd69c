"""MI355X-native ParHDE: spectral graph layout on one gfx950 GPU.

The layout path of the reference (k-centers BFS/SSSP pivots, degree-weighted
Gram-Schmidt, Laplacian SpMM, s x s eigensolve, coordinates) runs in
hand-written HIP kernels in ``lib/libparhde.so``, reached through the C ABI of
``include/parhde.h``.  This package is the Python view of that library.
"""
from .graphs import Graph  # noqa: F401
from .layout import (  # noqa: F401
    ALGO_PARHDE, ALGO_PHDE, ALGO_PIVOTMDS, ALGO_WPARHDE, EIG_GENERALIZED, EIG_JACOBI, EIG_POWER,
    PIVOTS_KCENTERS, PIVOTS_LEARNED, PIVOTS_RANDCOARSE, PIVOTS_RANDFINE, PREC_FP32, PREC_FP64,
    Context, DrawingBase, LayoutStats, ParHDE, PHDE, PhdeError, PivotMDS, WParHDE, default_options,
    device_count,
)

__version__ = "0.1.0"
