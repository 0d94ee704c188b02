"""Huber total variation with the reference's parameters and manager surface.

``TVParams`` / ``TVND`` mirror graspcg/regularization/tv_nd.py:19-88 and
``RegManager.energy_and_grad`` mirrors graspcg/regularization/manager.py:38-142;
the arithmetic runs in the fused HIP kernel ``mrh_tv_energy_grad``.

Axes are resolved against the solver image layout (F, 1, S0, S1) with roles
(unlike=1, like=1, nufft=2): 'temporal' is the frame axis, 'spatial' the two
image axes, and axis 1 is the size-1 coil slot whose differences vanish.  (The
reference's NUFFT exposes Roles(1, #like, 2) without the coil slot, which makes
'spatial' resolve to (1, 2) on a (F, 1, S0, S1) tensor; the solver here passes
the roles explicitly instead.)
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .roles import IMAGE_ROLES, AxesSpec, Roles

_AXIS_BIT = {0: _lib.TV_AXIS_T, 2: _lib.TV_AXIS_X, 3: _lib.TV_AXIS_Y}
_AXIS_SLOT = {0: 0, 2: 1, 3: 2}


@dataclass(frozen=True)
class TVParams:
    """weight = lambda, eps = Huber knee, axes = tokens or indices, isotropic = l2 across axes,
    axis_weights = per-axis scalars in the order of ``axes`` (tv_nd.py:19-33)."""
    weight: float = 0.0
    eps: float = 1e-3
    axes: AxesSpec = "spatial"
    isotropic: bool = True
    axis_weights: Optional[Sequence[float]] = None


class TVND:
    """N-D Huber TV term (tv_nd.py:35-88)."""
    Params = TVParams

    def __init__(self, name: str, params: TVParams):
        self.name = name
        self.params = params

    def term(self, roles: Roles = IMAGE_ROLES) -> Optional[_lib.TVTerm]:
        axes = roles.resolve_axes(self.params.axes)
        if self.params.axis_weights is not None and len(self.params.axis_weights) != len(axes):
            raise ValueError("axis_weights must match len(axes).")
        ws = self.params.axis_weights or [1.0] * len(axes)
        mask, w3 = 0, [1.0, 1.0, 1.0]
        for ax, w in zip(axes, ws):
            if ax == 1:          # size-1 coil slot: zero differences
                continue
            if ax not in _AXIS_BIT:
                raise ValueError(f"axis {ax} is not an axis of the (F, 1, S0, S1) image")
            mask |= _AXIS_BIT[ax]
            w3[_AXIS_SLOT[ax]] = float(w)
        if mask == 0 or self.params.weight == 0.0:
            return None
        t = _lib.TVTerm()
        t.weight, t.eps, t.axes, t.isotropic = float(self.params.weight), float(self.params.eps), mask, int(
            bool(self.params.isotropic))
        for i in range(3):
            t.axis_w[i] = w3[i]
        return t


@dataclass(frozen=True)
class GraphLapParams:
    """weight = lambda; normalize "none" (L = Deg - W), "rw" (I - Deg^-1 W) or "sym"
    (I - Deg^-1/2 W Deg^-1/2) (graspcg/regularization/graph_laplacian.py:11-20)."""
    weight: float = 0.0
    normalize: str = "none"
    mm_chunk_cols: int = 65536


class GraphLaplacian:
    """Graph-Laplacian term along the frame axis (graph_laplacian.py:22-321):
    E = (lambda / 2) sum_p Re<x[:, p], (L x)[:, p]>, gradient lambda L x.  W (F, F) dense or sparse
    COO; it is symmetrised and its diagonal zeroed as the reference does (:40-78).  The term runs in
    the NLCG engine's kernels (``k_gl_grad``, ``k_gl_ls``), at most one per manager."""
    Params = GraphLapParams

    def __init__(self, name: str, W, params: GraphLapParams):
        self.name = name
        self.params = params
        if isinstance(W, torch.Tensor):
            W = W.detach().to("cpu")
            W = W.to_dense() if W.is_sparse else W
            W = W.to(torch.float64).numpy()
        W = np.asarray(W, dtype=np.float64)
        if W.ndim != 2 or W.shape[0] != W.shape[1]:
            raise ValueError("W must be a square (F, F) matrix")
        W = 0.5 * (W + W.T)
        np.fill_diagonal(W, 0.0)
        self._W = W
        self._deg = np.maximum(W.sum(axis=1), 0.0)
        if (params.normalize or "none").lower() not in ("none", "rw", "sym"):
            raise ValueError("normalize must be 'none', 'rw' or 'sym'")

    @property
    def frames(self) -> int:
        return int(self._W.shape[0])

    def laplacian(self) -> np.ndarray:
        """L in float64 (graph_laplacian.py:140-154): the degree is clamped at 1e-20 where divided."""
        W, deg, n = self._W, self._deg, self._W.shape[0]
        norm = (self.params.normalize or "none").lower()
        if norm == "none":
            return np.diag(deg) - W
        if norm == "rw":
            return np.eye(n) - (1.0 / np.maximum(deg, 1e-20))[:, None] * W
        dmh = 1.0 / np.sqrt(np.maximum(deg, 1e-20))
        return np.eye(n) - dmh[:, None] * W * dmh[None, :]

    def profile(self) -> np.ndarray:
        """Diagonal majoriser per frame (graph_laplacian.py:283-321): lambda deg or lambda [deg > 0]."""
        v = self._deg if (self.params.normalize or "none").lower() == "none" else (self._deg > 0).astype(np.float64)
        return self.params.weight * v


class RegManager:
    """Orchestrates the TV terms (one fused kernel pass) and at most one graph-Laplacian term."""

    def __init__(self, regs: Sequence[object], *, compile_kernels: bool = True):
        self._regs: List[object] = list(regs)
        if len(self.terms()) > 3:
            raise NotImplementedError("at most 3 TV terms are fused per pass")
        if sum(isinstance(r, GraphLaplacian) for r in self._regs) > 1:
            raise NotImplementedError("at most one graph-Laplacian term")

    @property
    def regs(self) -> List[object]:
        return list(self._regs)

    def terms(self, roles: Roles = IMAGE_ROLES) -> List[_lib.TVTerm]:
        return [t for t in (r.term(roles) for r in self._regs if isinstance(r, TVND)) if t is not None]

    def graph(self) -> Optional[GraphLaplacian]:
        """The graph-Laplacian term with a non-zero weight, if any."""
        for r in self._regs:
            if isinstance(r, GraphLaplacian) and r.params.weight != 0.0:
                return r
        return None

    def tv_desc(self, P: int, F: int, S0: int, S1: int, roles: Roles = IMAGE_ROLES) -> _lib.TVDesc:
        d = _lib.TVDesc()
        d.P, d.F, d.S0, d.S1 = int(P), int(F), int(S0), int(S1)
        ts = self.terms(roles)
        d.n_terms = len(ts)
        for i, t in enumerate(ts):
            d.term[i] = t
        return d

    @torch.no_grad()
    def energy_and_grad_tensors(self, x: torch.Tensor, g: Optional[torch.Tensor] = None, *,
                                problems: int = 1) -> torch.Tensor:
        """TV energy per problem (float64, device) of x (P*F, [1,] S0, S1); g += gradient if given."""
        if self.graph() is not None:
            raise NotImplementedError("the graph-Laplacian term runs inside the NLCG engine only")
        if x.dtype != torch.complex64 or not x.is_contiguous():
            raise ValueError("x must be a contiguous complex64 tensor")
        if g is not None and (g.dtype != torch.complex64 or not g.is_contiguous() or g.shape != x.shape):
            raise ValueError("g must be a contiguous complex64 tensor shaped like x")
        if not (x.dim() == 3 or (x.dim() == 4 and int(x.shape[1]) == 1)):
            raise ValueError(f"x must be (P*F, S0, S1) or (P*F, 1, S0, S1), got shape {tuple(x.shape)}")
        S0, S1 = int(x.shape[-2]), int(x.shape[-1])
        B = int(x.shape[0])
        if B * S0 * S1 != x.numel():
            raise ValueError(f"x of shape {tuple(x.shape)} is not a stack of S0 x S1 frames")
        if B % problems:
            raise ValueError("frames not divisible by the number of problems")
        desc = self.tv_desc(problems, B // problems, S0, S1)
        lib = _lib.load()
        work = torch.empty(int(lib.mrh_reduce_workspace_len()), dtype=torch.float64, device=x.device)
        energy = torch.zeros(problems, dtype=torch.float64, device=x.device)
        _lib.check(lib.mrh_tv_energy_grad(ctypes.byref(desc), x.data_ptr(), _lib.ptr(g), work.data_ptr(),
                                          energy.data_ptr(), _lib.stream_of(x)))
        return energy

    @torch.no_grad()
    def energy_and_grad(self, ws) -> torch.Tensor:
        """Reference signature: reads ``var`` and accumulates into ``grad`` of a workspace."""
        return self.energy_and_grad_tensors(ws.get("var"), ws.get("grad")).sum()
