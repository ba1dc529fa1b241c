"""bmmpy_amd: the bmmpy row-XOR weight-minimization API on MI355X.

Same public names, keyword arguments, defaults and exceptions as the
reference's ``bmmpy`` package (bmmpy/__init__.py:34-105) for the search,
apply and runtime surface; every search and replay runs as HIP kernels
(libbmmpy_hip.so) driven by the C++ host layer in ``_bmmpy``.
"""
from __future__ import annotations

from . import _runtime  # noqa: F401  (must precede the native imports)

from ._bmmpy import (  # noqa: F401
    ApplyResult,
    BitMatrix,
    Candidate,
    MatrixErr,
    MatrixError,
    RowWindow,
    RuntimeFeatures,
    add,
    fixed_weight_masks_u32,
    fixed_weight_masks_u64,
    fwht_i16,
    fwht_i32,
    get_runtime_features,
    get_version,
)
from .apply import GlobalGreedyApplier, GreedyApplier
from .matrix import matrix_from_rows
from .search import (
    BruteforceSearch,
    CudaBruteforceSearch,
    CudaMitmFwhtSearch,
    FwhtSearch,
    MitmFwhtSearch,
    SASelectionResult,
    SASelector,
)
from .workflows import search_apply
from . import solver  # noqa: E402  (bmm.solver.ExactBasisSolver, as in the reference)
from . import ga  # noqa: E402
from .ga import (  # noqa: E402
    GeneticAlgorithm,
    IslandModel,
    IslandModelSnapshot,
    IslandSnapshot,
    IslandSpec,
    MigrationPolicy,
    OverflowPolicy,
    RunStats,
    StopCriteria,
)

__all__ = [
    "ApplyResult", "BitMatrix", "BruteforceSearch", "Candidate", "CudaBruteforceSearch",
    "CudaMitmFwhtSearch", "FwhtSearch", "GlobalGreedyApplier", "GreedyApplier", "MatrixErr", "MatrixError",
    "MitmFwhtSearch", "RowWindow", "SASelectionResult", "SASelector", "RuntimeFeatures", "add", "fwht_i16", "fwht_i32",
    "get_runtime_features", "get_version", "matrix_from_rows", "search_apply", "solver",
    "fixed_weight_masks_u32", "fixed_weight_masks_u64", "ga", "GeneticAlgorithm", "IslandModel",
    "IslandModelSnapshot", "IslandSnapshot", "IslandSpec", "MigrationPolicy", "OverflowPolicy",
    "RunStats", "StopCriteria",
]

__version__ = get_version()
