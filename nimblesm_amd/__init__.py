"""nimblesm_amd -- MI355X-native explicit internal-force path of NimbleSM.

HEX8 deformation gradients at 8 Gauss points, Neo-Hookean Cauchy stress,
B^T.sigma nodal forces with a deterministic node gather, and the
central-difference update, as hand-written HIP kernels for gfx950 behind the
C ABI of include/nimblesm_amd.h.  This package is the host side.
"""
from ._capi import NimbleError, lib  # noqa: F401
from .model import (  # noqa: F401
    Block,
    ExplicitTimeIntegrator,
    HexElement,
    ModelData,
    NeohookeanMaterial,
)
from .ensemble import Ensemble, EnsembleProblem, config2_draws, draw_parameters  # noqa: F401
from .problems import CubeProblem, jittered_cube, workload  # noqa: F401

__all__ = [
    "Block",
    "CubeProblem",
    "Ensemble",
    "EnsembleProblem",
    "ExplicitTimeIntegrator",
    "HexElement",
    "ModelData",
    "NeohookeanMaterial",
    "NimbleError",
    "config2_draws",
    "draw_parameters",
    "jittered_cube",
    "lib",
    "workload",
]
