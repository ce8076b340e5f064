"""MI355X-native finite-volume time step of the Barton-2009 hyperelastic solver.

Module names mirror the reference's Julia modules (EquationsOfState, Strains,
HyperelasticityMPh, Hyperelasticity, NumFluxes); the time loop of main.jl is
`solver.Solver`.  All numerics run in libhes.so (gfx950 HIP kernels).
"""
from ._lib import DomainError, HesError, HesLibraryMissing  # noqa: F401

__all__ = ["DomainError", "HesError", "HesLibraryMissing"]
