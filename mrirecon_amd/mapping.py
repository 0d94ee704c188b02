"""Low-rank temporal basis: the parameter-space map u = U V and TV applied through it.

Mirrors graspcg/regularization/mapping.py:112-231 (``TemporalBasisOp``) and the
``Transformed(base=TVND, op=TemporalBasisOp)`` wrapper (graspcg/regularization/wrappers.py:9-105,
``MappedRegularizer`` at mapping.py:293-356) on the HIP kernels of ``csrc/tbasis.hip``:

* ``U`` (T, K): frame t's weights on K basis images; ``V`` (K, ...): the coefficient images.
* ``forward_apply(V)``     u = U V                       (mapping.py:129-147)
* ``adjoint_apply(g_u, g_V)``  g_V += U^H g_u; the field gets no gradient (mapping.py:149-164)
* ``diag_push(k_u, diag_V)``   diag_V[k] += sum_t |U_tk|^2 k_u[t]   (mapping.py:170-199)
* ``diag_push_profile(0, v, diag_V)``  diag_V[k] += sum_t |U_tk|^2 v[t]  (mapping.py:200-231)

In the reference's CG the op is terminal in parameter space: its regulariser adds TV(U V)
to the energy and nothing to the image gradient, so the NLCG engine (which optimises the
frames) does not take it; ``MappedTV`` evaluates the energy, the coefficient gradient and the
coefficient-space diagonal for callers that optimise V.
"""
from __future__ import annotations

from typing import Optional, Sequence

import torch

from . import _lib
from .regularization import RegManager, TVND
from .roles import IMAGE_ROLES, Roles

MAX_K = 16     # MRH_TBASIS_MAX_K


def _cplx(t: torch.Tensor, what: str) -> torch.Tensor:
    if t.dtype != torch.complex64 or not t.is_contiguous() or t.device.type != "cuda":
        raise ValueError(f"{what} must be a contiguous complex64 CUDA tensor")
    return t


class TemporalBasisOp:
    """u = U V with U (T, K) (graspcg/regularization/mapping.py:112-231).

    ``param_key`` / ``grad_key`` / ``diag_key`` name V, g_V and diag_V in a workspace, as in
    the reference; the tensor methods below take them directly."""

    def __init__(self, U, param_key: str = "V", grad_key: str = "g_V", diag_key: str = "diag_V",
                 device: Optional[torch.device] = None):
        U = torch.as_tensor(U)
        if U.dim() != 2:
            raise ValueError(f"U must be (T, K), got shape {tuple(U.shape)}")
        T, K = int(U.shape[0]), int(U.shape[1])
        if T < 1 or not 1 <= K <= MAX_K:
            raise ValueError(f"U is ({T}, {K}); need T >= 1 and 1 <= K <= {MAX_K}")
        dev = torch.device(device) if device is not None else (U.device if U.is_cuda else torch.device("cuda"))
        self.U = U.to(device=dev, dtype=torch.complex64).contiguous()
        self.T, self.K = T, K
        self.param_key, self.grad_key, self.diag_key = param_key, grad_key, diag_key
        self._lib = _lib.load()

    def _n(self, t: torch.Tensor, lead: int, what: str) -> int:
        if t.dim() < 1 or int(t.shape[0]) != lead:
            raise ValueError(f"{what} must have {lead} entries on axis 0, got shape {tuple(t.shape)}")
        return t.numel() // lead

    @torch.no_grad()
    def forward_apply(self, V: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        """u = U V: (K, ...) -> (T, ...)."""
        _cplx(V, "V")
        N = self._n(V, self.K, "V")
        u = torch.empty((self.T,) + tuple(V.shape[1:]), dtype=torch.complex64, device=V.device) if out is None \
            else _cplx(out, "out")
        if u.numel() != self.T * N:
            raise ValueError("out must hold T x (elements per coefficient image)")
        _lib.check(self._lib.mrh_tbasis_forward(self.U.data_ptr(), self.T, self.K, V.data_ptr(), N, u.data_ptr(),
                                                _lib.stream_of(V)))
        return u

    @torch.no_grad()
    def adjoint_apply(self, grad_u: torch.Tensor, g_V: torch.Tensor, accumulate: bool = True) -> torch.Tensor:
        """g_V += U^H grad_u (``accumulate`` False: g_V = U^H grad_u); returns g_V."""
        _cplx(grad_u, "grad_u")
        _cplx(g_V, "g_V")
        N = self._n(grad_u, self.T, "grad_u")
        if g_V.numel() != self.K * N:
            raise ValueError("g_V must hold K x (elements per frame)")
        _lib.check(self._lib.mrh_tbasis_adjoint(self.U.data_ptr(), self.T, self.K, grad_u.data_ptr(), N,
                                                g_V.data_ptr(), int(bool(accumulate)), _lib.stream_of(grad_u)))
        return g_V

    @torch.no_grad()
    def diag_push(self, k_u: torch.Tensor, diag_V: torch.Tensor) -> None:
        """diag_V[k] += sum_t |U_tk|^2 k_u[t] for a (T, ...) float32 field; terminal (returns None)."""
        N = self._n(k_u, self.T, "k_u")
        self._diag(k_u, 0, N, diag_V)

    @torch.no_grad()
    def diag_push_profile(self, axis: int, v1d: torch.Tensor, diag_V: torch.Tensor) -> None:
        """diag_V[k, ...] += sum_t |U_tk|^2 v[t] for a frame profile (axis 0); other axes are
        dropped, as in the reference.  Terminal (returns None)."""
        if axis != 0:
            return None
        v = torch.as_tensor(v1d, dtype=torch.float32, device=self.U.device).reshape(-1).contiguous()
        if v.numel() != self.T:
            raise ValueError(f"the frame profile needs {self.T} values, got {v.numel()}")
        self._diag(v, 1, diag_V.numel() // self.K, diag_V)
        return None

    def _diag(self, ku: torch.Tensor, profile: int, N: int, diag_V: torch.Tensor) -> None:
        if ku.dtype != torch.float32 or not ku.is_contiguous() or ku.device.type != "cuda":
            raise ValueError("diagonal weights must be a contiguous float32 CUDA tensor")
        if diag_V.dtype != torch.float32 or not diag_V.is_contiguous() or diag_V.numel() != self.K * N:
            raise ValueError("diag_V must be a contiguous float32 tensor of K x (elements per frame)")
        _lib.check(self._lib.mrh_tbasis_diag(self.U.data_ptr(), self.T, self.K, ku.data_ptr(), profile, N,
                                             diag_V.data_ptr(), _lib.stream_of(diag_V)))

    def roles_for_output(self, roles: Roles) -> Roles:
        return roles          # u keeps the frame / spatial meaning of the axes (mapping.py:166-169)


class MappedTV:
    """TV terms on u = U V with their gradient pushed to V: the reference's
    ``Transformed(base=TVND, op=TemporalBasisOp)`` (wrappers.py:9-105, mapping.py:293-356).
    V is (K, S0, S1) or (K, 1, S0, S1); u has the T frames of U on axis 0."""

    def __init__(self, terms: Sequence[TVND], op: TemporalBasisOp, roles: Roles = IMAGE_ROLES):
        self.terms = list(terms)
        if not all(isinstance(t, TVND) for t in self.terms):
            raise TypeError("MappedTV takes TVND terms")
        self.op, self.roles = op, roles
        self.regm = RegManager(self.terms)

    @torch.no_grad()
    def energy_grad(self, V: torch.Tensor, g_V: torch.Tensor) -> torch.Tensor:
        """Energy of the terms at u = U V (FP64 device scalar); g_V += U^H grad_u."""
        if not (V.dim() == 3 or (V.dim() == 4 and int(V.shape[1]) == 1)):
            raise ValueError(f"V must be (K, S0, S1) or (K, 1, S0, S1) for TV on its frames, got {tuple(V.shape)}")
        u = self.op.forward_apply(V)
        g_u = torch.zeros_like(u)
        E = self.regm.energy_and_grad_tensors(u, g_u, problems=1)
        self.op.adjoint_apply(g_u, g_V)
        return E[0]

    @torch.no_grad()
    def add_diag(self, diag_V: torch.Tensor, image_diag: bool = False) -> None:
        """Coefficient-space diagonal majoriser, as Transformed.add_diag (wrappers.py:39-105):
        with ``image_diag`` (the caller's context carries an image diagonal) a term over frames
        first pushes its frame profile 2 lambda w_0^2 x degree; that push is terminal for this
        op, so every term then also pushes its constant 2 lambda sum_a w_a^2."""
        T = self.op.T
        for term in self.terms:
            p = term.params
            axes = tuple(self.roles.resolve_axes(p.axes))
            if p.weight == 0.0 or not axes:
                continue
            if p.axis_weights is not None and len(p.axis_weights) != len(axes):
                continue    # the reference pushes no profile and k = 0 then (tv_nd.py:176-179,197-199)
            ws = list(p.axis_weights) if p.axis_weights is not None else [1.0] * len(axes)
            if image_diag and 0 in axes:
                w0 = ws[axes.index(0)]
                deg = torch.ones(T, dtype=torch.float32)
                if T > 2:
                    deg[1:-1] = 2.0
                self.op.diag_push_profile(0, deg * float(2.0 * p.weight * w0 * w0), diag_V)
            k = 2.0 * float(p.weight) * sum(float(w) ** 2 for w in ws)
            self.op.diag_push_profile(0, torch.full((T,), k, dtype=torch.float32), diag_V)
