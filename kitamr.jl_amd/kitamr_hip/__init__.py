"""kitamr_hip: MI355X (gfx950) backend of KitAMR.jl's per-step phase-space update.

The per-step work (slope!, flux!, iterate!) runs in hand-written HIP kernels behind the
C ABI of include/kitamr_hip.h; this package is the host side: forest and face-list
flattening, velocity grids, and the solver API mirror.
"""
from .solver import (Configure, Domain, Gas, PCoordFn, Solver, Uniform, build_case, flux,
                     initialize, iterate, limit_dt, slope, solve, step)

__all__ = ["Configure", "Domain", "Gas", "PCoordFn", "Solver", "Uniform", "build_case", "flux",
           "initialize", "iterate", "limit_dt", "slope", "solve", "step"]
