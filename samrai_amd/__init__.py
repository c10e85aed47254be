"""MI355X-native FAC Poisson V-cycle (the SAMRAI CellPoissonFACSolver path).

The compute path is libsamrai_fac.so (hand-written gfx950 HIP kernels and the
C++ CellPoissonFACOps / FACPreconditioner / CellPoissonFACSolver), reached
through the C ABI in include/samrai_fac.h.  This package is the Python-side
mirror of that interface plus hierarchy generators for the benchmark decks.
"""
from .fac import (CellPoissonFACSolver, PatchData, PatchHierarchy, SolverError,  # noqa: F401
                  lib)
