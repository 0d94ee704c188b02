"""NUFFT front-end with the reference's public surface (graspcg/nufft/api.py:28-376).

``NUFFT(maps, traj, dcf, axis=AxisSpec, config=NUFFTConfig)`` exposes
``A(x, out=None)`` / ``AH(y, out=None)`` in the user's axis order, writes
results out-first into ``out`` views, ``k_per_frame()``, ``image_shape(y)``,
``domain_shape()`` and ``roles_image`` / ``roles_kspace``.  The only backend
built here is ``'hip'`` (the MI355X kernels); Like axes are folded into the
batch like the reference's torchkb path (api.py:195-201).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, Literal, Optional, Sequence, Tuple

import torch

from ..roles import Roles
from .backends.hip import HipNUFFTAdapter
from .layout import (AxisSpec, plan_image_layout, plan_image_layout_from_sizes, plan_kspace_layout,
                     plan_kspace_layout_from_sizes)


@dataclass(frozen=True)
class NUFFTConfig:
    """Front-end configuration (graspcg/nufft/api.py:30-47) plus the KB knobs of the 'hip' backend."""
    ndim: Literal[2, 3]
    backend: Literal["hip"] = "hip"
    traj_units: Literal["norm", "rad"] = "norm"
    dcf_mode: Literal["standard", "balanced", "none"] = "balanced"
    apply_dcf_in_fwd: bool = False
    apply_dcf_in_adj: bool = True
    vectorize_coils: bool = True
    eps: float = 1e-6
    isign: int = -1
    kb_width: int = 6
    osf: float = 2.0
    strip_cols: int = 0


class NUFFT:
    """Forward ``A``: image (B, Like..., 1, FFT...) -> k-space (B, Like..., C, K);
    adjoint ``AH`` the reverse, coil-combined with conj(maps) inside the operator."""

    def __init__(self, maps: torch.Tensor, traj: torch.Tensor, dcf: Optional[torch.Tensor], *,
                 axis: AxisSpec, config: NUFFTConfig, dtype: torch.dtype = torch.complex64,
                 device: Optional[torch.device] = None):
        if maps.ndim not in (3, 4):
            raise ValueError("maps must be (C,X,Y) or (C,X,Y,Z)")
        ndim = 2 if maps.ndim == 3 else 3
        if config.ndim != ndim:
            raise ValueError(f"config.ndim={config.ndim} must match maps dimensionality {ndim}")
        if config.backend != "hip":
            raise ValueError("backend must be 'hip' (the only backend built in this package)")
        if config.isign != -1:
            raise ValueError("the 'hip' backend implements isign=-1 (exp(-i w.r) forward)")
        self.maps, self.traj0, self.dcf0 = maps, traj, dcf
        self.axis, self.cfg, self.dtype = axis, config, dtype
        self.device = torch.device(device) if device is not None else maps.device
        self._traj_BndK = self._to_BndK(traj, ndim)
        be = HipNUFFTAdapter(ndim=ndim, traj_units=config.traj_units, dcf_mode=config.dcf_mode,
                             apply_dcf_in_fwd=config.apply_dcf_in_fwd,
                             apply_dcf_in_adj=config.apply_dcf_in_adj, kb_width=config.kb_width,
                             osf=config.osf, strip_cols=config.strip_cols)
        d = None
        if dcf is not None:
            d = dcf.view(1, -1) if dcf.ndim == 1 else dcf
            if d.ndim != 2:
                raise ValueError("dcf must be (B,K) or (K)")
        be.prepare((int(maps.shape[0]),) + tuple(int(s) for s in maps.shape[1:]), self._traj_BndK, d,
                   maps.to(self.device, dtype=torch.complex64), dtype=dtype, device=self.device,
                   like_prod=int(maps.shape[0]))
        self._backend = be
        like_img = axis.like_image()
        like_ks = axis.like_kspace()
        self.roles_image = Roles(unlike=1, like=len(like_img), nufft=len(axis.image_fft))
        self.roles_kspace = Roles(unlike=1, like=len(like_ks), nufft=1)

    @property
    def backend(self) -> HipNUFFTAdapter:
        return self._backend

    # -------------------------------------------------------------- operators
    @torch.no_grad()
    def A(self, x_user: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        ip = plan_image_layout(x_user, self.axis)
        C, K = int(self.maps.shape[0]), self.k_per_frame()
        shape = self._user_kspace_shape(ip.B, ip.like_sizes, C, K)
        if out is None:
            out = torch.empty(shape, device=self.device, dtype=x_user.dtype)
        elif tuple(out.shape) != tuple(shape):
            raise ValueError(f"out has shape {tuple(out.shape)} but expected {tuple(shape)}")
        x_view = self._image_view(x_user, ip).reshape(-1, 1, *ip.spatial)
        self._write(out, self._kspace_perm(None, ip.like_labels),
                    lambda o: self._backend.A(x_view, out=o), (x_view.shape[0], C, K))
        return out

    @torch.no_grad()
    def AH(self, y_user: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        kp = plan_kspace_layout(y_user, self.axis)
        spatial = tuple(int(s) for s in self.maps.shape[1:])
        C = int(self.maps.shape[0])
        if kp.C != C:
            raise ValueError(f"y coil dim {kp.C} must equal maps C {C}")
        shape = self._user_image_shape(kp.B, kp.like_sizes, spatial)
        if out is None:
            out = torch.empty(shape, device=self.device, dtype=y_user.dtype)
        elif tuple(out.shape) != tuple(shape):
            raise ValueError(f"out has shape {tuple(out.shape)} but expected {tuple(shape)}")
        y_view = self._kspace_perm(y_user, kp.like_labels).reshape(-1, C, kp.K)
        ip = plan_image_layout_from_sizes(kp.B, kp.like_sizes, spatial, self.axis)
        self._write(out, self._image_perm(out, ip.like_labels),
                    lambda o: self._backend.AH(y_view, out=o), (y_view.shape[0], 1, *spatial))
        return out

    @staticmethod
    def _write(out: torch.Tensor, perm, compute, canon_shape) -> None:
        """Run ``compute`` straight into ``out`` when its canonical permutation is contiguous
        (out-first, no result allocation); otherwise into a temporary copied back."""
        view = out.permute(*perm)
        if view.is_contiguous() and out.dtype == torch.complex64:
            compute(view.view(canon_shape))
        else:
            view.copy_(compute(None).reshape(view.shape))

    # -------------------------------------------------------------- shapes
    def k_per_frame(self) -> int:
        return int(self._backend.k_per_frame())

    def image_shape(self, y_user: Optional[torch.Tensor] = None) -> Tuple[int, ...]:
        return (int(self._traj_BndK.shape[0]), 1, *(int(s) for s in self.maps.shape[1:]))

    def domain_shape(self) -> Tuple[int, ...]:
        return (1, 1, *(int(s) for s in self.maps.shape[1:]))

    # -------------------------------------------------------------- views
    def _user_kspace_shape(self, B, L_sizes, C, K) -> Tuple[int, ...]:
        kp = plan_kspace_layout_from_sizes(B, L_sizes, C, K, self.axis)
        size: Dict[str, int] = {"B": kp.B, self.axis.coil: kp.C, self.axis.kspace_fft[0]: kp.K}
        size.update(zip(kp.like_labels, kp.like_sizes))
        return tuple(int(size[l]) for l in self.axis.kspace)

    def _user_image_shape(self, B, L_sizes, spatial) -> Tuple[int, ...]:
        ip = plan_image_layout_from_sizes(B, L_sizes, spatial, self.axis)
        size: Dict[str, int] = {"B": ip.B, self.axis.coil: 1}
        size.update(zip(ip.like_labels, ip.like_sizes))
        size.update(zip(self.axis.image_fft, spatial))
        return tuple(int(size[l]) for l in self.axis.image)

    def _kspace_perm(self, y: torch.Tensor, like_labels) -> torch.Tensor:
        """Permutation of a user k-space tensor to (B, Like..., C, K) (api.py:304-323)."""
        target = ("B",) + tuple(like_labels) + (self.axis.coil, self.axis.kspace_fft[0])
        pos = {l: i for i, l in enumerate(self.axis.kspace)}
        perm = [pos[l] for l in target]
        return perm if y is None else y.permute(*perm)

    def _image_perm(self, x: torch.Tensor, like_labels):
        """Permutation of a user image tensor to (B, Like..., [coil], FFT...) (api.py:325-349)."""
        pos = {l: i for i, l in enumerate(self.axis.image)}
        target = ["B"] + list(like_labels) + ([self.axis.coil] if self.axis.coil in pos else []) \
            + list(self.axis.image_fft)
        return [pos[l] for l in target]

    def _image_view(self, x: torch.Tensor, ip) -> torch.Tensor:
        return x.permute(*self._image_perm(x, ip.like_labels))

    @staticmethod
    def _to_BndK(traj: torch.Tensor, ndim: int) -> torch.Tensor:
        if traj.ndim == 2:
            t = traj
            if int(t.shape[0]) != ndim and int(t.shape[1]) == ndim:
                t = t.transpose(0, 1)
            if int(t.shape[0]) != ndim:
                raise ValueError(f"traj has wrong nd; expected {ndim}")
            return t.unsqueeze(0).contiguous()
        if traj.ndim == 3:
            if int(traj.shape[1]) == ndim:
                return traj
            if int(traj.shape[2]) == ndim:
                return traj.transpose(1, 2).contiguous()
            raise ValueError("traj must be (B,nd,K) or (B,K,nd)")
        raise ValueError("traj must have 2 or 3 dims")


NSNUFFT = NUFFT

__all__ = ["NUFFT", "NSNUFFT", "NUFFTConfig", "AxisSpec"]
