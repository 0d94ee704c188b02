"""Axis grammar for user layouts (same contract as graspcg/nufft/layout.py:11-194).

Image tensors are (B, Like..., [coil=1], FFT...) in any label order; k-space
tensors are (B, Like..., coil, K).  Like axes are every label that is neither
'B', the coil label nor an FFT label.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, Optional, Sequence, Tuple

import torch


class AxisSpec:
    """User axis labels; accepts the new (``image_labels=...``) or legacy (``image=...``) names
    (graspcg/nufft/layout.py:28-60)."""

    def __init__(self, *, image_labels=None, kspace_labels=None, image_fft_labels=None,
                 kspace_fft_labels=None, coil_label=None, image=None, kspace=None, image_fft=None,
                 kspace_fft=None, coil=None):
        image_labels = image_labels if image_labels is not None else image
        kspace_labels = kspace_labels if kspace_labels is not None else kspace
        image_fft_labels = image_fft_labels if image_fft_labels is not None else image_fft
        kspace_fft_labels = kspace_fft_labels if kspace_fft_labels is not None else kspace_fft
        coil_label = coil_label if coil_label is not None else (coil if coil is not None else "coil")
        if image_labels is None or kspace_labels is None or image_fft_labels is None:
            raise TypeError("AxisSpec requires image(_labels)=..., kspace(_labels)=..., image_fft(_labels)=...")
        self.image_labels = tuple(image_labels)
        self.kspace_labels = tuple(kspace_labels)
        self.image_fft_labels = tuple(image_fft_labels)
        self.kspace_fft_labels = tuple(kspace_fft_labels) if kspace_fft_labels is not None else ("K",)
        self.coil_label = str(coil_label)

    image = property(lambda self: self.image_labels)
    kspace = property(lambda self: self.kspace_labels)
    image_fft = property(lambda self: self.image_fft_labels)
    kspace_fft = property(lambda self: self.kspace_fft_labels)
    coil = property(lambda self: self.coil_label)

    def like_image(self) -> Tuple[str, ...]:
        return tuple(l for l in self.image if l not in self.image_fft and l not in ("B", self.coil))

    def like_kspace(self) -> Tuple[str, ...]:
        return tuple(l for l in self.kspace if l not in self.kspace_fft and l not in ("B", self.coil))

    def __repr__(self) -> str:
        return (f"AxisSpec(image={self.image}, kspace={self.kspace}, image_fft={self.image_fft}, "
                f"coil={self.coil!r})")


@dataclass(frozen=True)
class ImageLayoutPlan:
    axis: AxisSpec
    B: int
    like_labels: Tuple[str, ...]
    like_sizes: Tuple[int, ...]
    L_other: int
    spatial: Tuple[int, ...]


@dataclass(frozen=True)
class KspaceLayoutPlan:
    axis: AxisSpec
    B: int
    like_labels: Tuple[str, ...]
    like_sizes: Tuple[int, ...]
    L_other: int
    C: int
    K: int


def _sizes(t: torch.Tensor, labels: Sequence[str]) -> Dict[str, int]:
    if t.ndim != len(labels):
        raise ValueError(f"tensor has {t.ndim} dims but the axis spec names {len(labels)}: {labels}")
    return {lbl: int(t.shape[i]) for i, lbl in enumerate(labels)}


def plan_image_layout(x: torch.Tensor, axis: AxisSpec) -> ImageLayoutPlan:
    s = _sizes(x, axis.image)
    if axis.coil in s and s[axis.coil] != 1:
        raise ValueError("Image is expected coil-combined; axis.coil dim must be size 1 if present.")
    like = axis.like_image()
    ls = tuple(s[l] for l in like)
    return ImageLayoutPlan(axis, s.get("B", 1), like, ls, int(math.prod(ls)) if ls else 1,
                           tuple(s[l] for l in axis.image_fft))


def plan_kspace_layout(y: torch.Tensor, axis: AxisSpec) -> KspaceLayoutPlan:
    s = _sizes(y, axis.kspace)
    like = axis.like_kspace()
    ls = tuple(s[l] for l in like)
    return KspaceLayoutPlan(axis, s.get("B", 1), like, ls, int(math.prod(ls)) if ls else 1,
                            s[axis.coil], s[axis.kspace_fft[0]])


def _like_sizes(labels: Tuple[str, ...], L_sizes: Sequence[int]) -> Tuple[Tuple[int, ...], int]:
    if len(labels) == 0:
        return tuple(), 1
    if len(L_sizes) == len(labels):
        ls = tuple(int(v) for v in L_sizes)
        return ls, int(math.prod(ls))
    if len(L_sizes) == 1:   # product assigned to the first like label (layout.py:164-168)
        ls = [1] * len(labels)
        ls[0] = int(L_sizes[0])
        return tuple(ls), int(L_sizes[0])
    raise ValueError("L_sizes length must be 0, 1, or match number of like labels.")


def plan_image_layout_from_sizes(B: int, L_sizes: Sequence[int], spatial: Sequence[int],
                                 axis: AxisSpec) -> ImageLayoutPlan:
    like = axis.like_image()
    ls, L = _like_sizes(like, L_sizes)
    return ImageLayoutPlan(axis, int(B), like, ls, L, tuple(int(v) for v in spatial))


def plan_kspace_layout_from_sizes(B: int, L_sizes: Sequence[int], C: int, K: int,
                                  axis: AxisSpec) -> KspaceLayoutPlan:
    like = axis.like_kspace()
    ls, L = _like_sizes(like, L_sizes)
    return KspaceLayoutPlan(axis, int(B), like, ls, L, int(C), int(K))
