"""MI355X (gfx950) split patch update for the perfect-gas Euler equations.

A drop-in for FiniteVolumeSolver's FluxMethod / HyperbolicPatchIntegrator /
CompleteFromCons path (MUSCL-Hancock + HLLEM or exact Godunov, CFL reduction),
with a learned initial guess for the exact solver's Newton iteration, computed by the
HIP library libfvs_hip.so behind the C ABI of include/fvs.h.
"""
from . import _lib
from .patch import (CompleteFromCons, FluxMethod, HyperbolicPatchIntegrator, PatchView,
                    PerfectGas, SinglePatchLevel, TransmissiveBoundary, godunov_sweep,
                    make_desc, patch_stable_dt, split_sweep)
from ._lib import FvsError, NewtonNotConverged, TimeStepTooLarge
from .newton_guess import LearnedNewtonGuess, exact_middle_state

__all__ = [
    "CompleteFromCons", "FluxMethod", "HyperbolicPatchIntegrator", "PatchView", "PerfectGas",
    "SinglePatchLevel", "TransmissiveBoundary", "godunov_sweep", "make_desc", "patch_stable_dt",
    "split_sweep",
    "FvsError", "NewtonNotConverged", "TimeStepTooLarge", "LearnedNewtonGuess", "exact_middle_state",
]
