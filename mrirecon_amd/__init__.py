"""mrirecon_amd -- MI355X-native NUFFT-driven NLCG reconstruction (graspcg hot path).

Public names follow the reference package graspcg:
``NUFFT, NUFFTConfig, AxisSpec`` (graspcg/nufft/api.py), ``TVND, TVParams,
RegManager`` (graspcg/regularization), ``CGSolver, CGConfig``
(graspcg/solvers/cg.py), ``Roles`` (graspcg/core/roles.py).  The compute runs
in hand-written gfx950 kernels behind the C ABI of include/mrirecon_hip.h.
"""
from .roles import Roles, IMAGE_ROLES
from .nufft.layout import AxisSpec
from .nufft.api import NUFFT, NUFFTConfig
from .regularization import TVND, TVParams, RegManager, GraphLaplacian, GraphLapParams
from .solver import CGSolver, CGConfig, NLCGEngine
from .mapping import TemporalBasisOp, MappedTV

__all__ = ["TemporalBasisOp", "MappedTV", "NUFFT", "NUFFTConfig", "AxisSpec", "TVND", "TVParams", "RegManager", "GraphLaplacian", "GraphLapParams", "CGSolver",
           "CGConfig", "NLCGEngine", "Roles", "IMAGE_ROLES"]
