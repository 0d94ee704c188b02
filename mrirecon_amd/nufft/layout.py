"""Axis labels of user tensors and the layout plans derived from them.

Interface of graspcg/nufft/layout.py:11-194.  An image tensor carries the
labels of ``AxisSpec.image`` (batch 'B', any number of Like labels, an optional
size-1 coil axis and the FFT axes, in any order); a k-space tensor carries
``AxisSpec.kspace`` (batch, Like labels, the coil axis and the sample axis).
Every label that is none of batch, coil or an FFT label is a Like label, and
the Like axes are folded into the batch by the operator.  Field names, both
keyword styles of ``AxisSpec`` (``image_labels=`` or ``image=``), the default
labels ('coil', ('K',)) and the exception types are the reference's; the code
is this package's.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, Optional, Sequence, Tuple

import torch

_BATCH = "B"


class AxisSpec:
    """Labels of image and k-space tensors.  Each field may be given under its long keyword
    (``image_labels``, ``kspace_labels``, ``image_fft_labels``, ``kspace_fft_labels``,
    ``coil_label``) or its short one (``image``, ``kspace``, ``image_fft``, ``kspace_fft``,
    ``coil``); the long one wins when both are given."""

    _FIELDS = (("image_labels", "image"), ("kspace_labels", "kspace"),
               ("image_fft_labels", "image_fft"), ("kspace_fft_labels", "kspace_fft"),
               ("coil_label", "coil"))
    _DEFAULTS = {"kspace_fft_labels": ("K",), "coil_label": "coil"}

    def __init__(self, **kw):
        known = {n for pair in self._FIELDS for n in pair}
        extra = sorted(set(kw) - known)
        if extra:
            raise TypeError(f"AxisSpec got unexpected keyword(s) {extra}")
        given: Dict[str, object] = {}
        for long_name, short_name in self._FIELDS:
            v = kw.get(long_name)
            if v is None:
                v = kw.get(short_name)
            if v is None:
                v = self._DEFAULTS.get(long_name)
            given[long_name] = v
        missing = [f"{ln} (or {sn})" for ln, sn in self._FIELDS if given[ln] is None]
        if missing:
            raise TypeError("AxisSpec needs labels for " + ", ".join(missing))
        self.image_labels: Tuple[str, ...] = tuple(given["image_labels"])
        self.kspace_labels: Tuple[str, ...] = tuple(given["kspace_labels"])
        self.image_fft_labels: Tuple[str, ...] = tuple(given["image_fft_labels"])
        self.kspace_fft_labels: Tuple[str, ...] = tuple(given["kspace_fft_labels"])
        self.coil_label: str = str(given["coil_label"])

    # short names (read-only)
    image = property(lambda self: self.image_labels)
    kspace = property(lambda self: self.kspace_labels)
    image_fft = property(lambda self: self.image_fft_labels)
    kspace_fft = property(lambda self: self.kspace_fft_labels)
    coil = property(lambda self: self.coil_label)

    def _like(self, labels: Sequence[str], fft: Sequence[str]) -> Tuple[str, ...]:
        structural = set(fft) | {_BATCH, self.coil_label}
        return tuple(lbl for lbl in labels if lbl not in structural)

    def like_image(self) -> Tuple[str, ...]:
        """Like labels of image tensors, in label order."""
        return self._like(self.image_labels, self.image_fft_labels)

    def like_kspace(self) -> Tuple[str, ...]:
        """Like labels of k-space tensors, in label order."""
        return self._like(self.kspace_labels, self.kspace_fft_labels)

    def __repr__(self) -> str:
        return (f"AxisSpec(image={self.image_labels}, kspace={self.kspace_labels}, "
                f"image_fft={self.image_fft_labels}, kspace_fft={self.kspace_fft_labels}, "
                f"coil={self.coil_label!r})")


@dataclass(frozen=True)
class ImageLayoutPlan:
    axis: AxisSpec
    B: int
    like_labels: Tuple[str, ...]
    like_sizes: Tuple[int, ...]
    L_other: int
    spatial: Tuple[int, ...]          # sizes of the FFT axes, in image_fft order


@dataclass(frozen=True)
class KspaceLayoutPlan:
    axis: AxisSpec
    B: int
    like_labels: Tuple[str, ...]
    like_sizes: Tuple[int, ...]
    L_other: int
    C: int
    K: int                            # size of the first k-space FFT label


def _label_sizes(t: torch.Tensor, labels: Sequence[str], what: str) -> Dict[str, int]:
    if t.dim() != len(labels):
        raise ValueError(f"{what} tensor of shape {tuple(t.shape)} does not match its {len(labels)} labels {tuple(labels)}")
    return dict(zip(labels, (int(n) for n in t.shape)))


def _fold_like(like_labels: Tuple[str, ...], L_sizes: Sequence[int]) -> Tuple[Tuple[int, ...], int]:
    """Sizes of the Like labels from a size list: one size per label, or a single total that the
    first label takes (the others 1), or nothing when there are no Like labels."""
    n = len(like_labels)
    if n == 0:
        return (), 1
    sizes = [int(v) for v in L_sizes]
    if len(sizes) == 1 and n > 1:
        sizes = sizes + [1] * (n - 1)
    if len(sizes) != n:
        raise ValueError(f"got {len(L_sizes)} Like sizes for the Like labels {like_labels}; "
                         f"give one per label or a single total")
    return tuple(sizes), math.prod(sizes)


def plan_image_layout(x_user: torch.Tensor, axis: AxisSpec) -> ImageLayoutPlan:
    sz = _label_sizes(x_user, axis.image_labels, "image")
    ncoil = sz.get(axis.coil_label, 1)
    if ncoil != 1:
        raise ValueError(f"images are coil-combined: the '{axis.coil_label}' axis must have size 1, not {ncoil}")
    like = axis.like_image()
    ls = tuple(sz[lbl] for lbl in like)
    return ImageLayoutPlan(axis=axis, B=sz.get(_BATCH, 1), like_labels=like, like_sizes=ls,
                           L_other=math.prod(ls), spatial=tuple(sz[lbl] for lbl in axis.image_fft_labels))


def plan_kspace_layout(y_user: torch.Tensor, axis: AxisSpec) -> KspaceLayoutPlan:
    sz = _label_sizes(y_user, axis.kspace_labels, "k-space")
    like = axis.like_kspace()
    ls = tuple(sz[lbl] for lbl in like)
    return KspaceLayoutPlan(axis=axis, B=sz.get(_BATCH, 1), like_labels=like, like_sizes=ls,
                            L_other=math.prod(ls), C=sz[axis.coil_label], K=sz[axis.kspace_fft_labels[0]])


def plan_image_layout_from_sizes(B: int, L_sizes: Sequence[int], spatial: Sequence[int],
                                 axis: AxisSpec) -> ImageLayoutPlan:
    like = axis.like_image()
    ls, L = _fold_like(like, L_sizes)
    return ImageLayoutPlan(axis=axis, B=int(B), like_labels=like, like_sizes=ls, L_other=int(L),
                           spatial=tuple(int(s) for s in spatial))


def plan_kspace_layout_from_sizes(B: int, L_sizes: Sequence[int], C: int, K: int,
                                  axis: AxisSpec) -> KspaceLayoutPlan:
    like = axis.like_kspace()
    ls, L = _fold_like(like, L_sizes)
    return KspaceLayoutPlan(axis=axis, B=int(B), like_labels=like, like_sizes=ls, L_other=int(L),
                            C=int(C), K=int(K))


__all__ = ["AxisSpec", "ImageLayoutPlan", "KspaceLayoutPlan", "plan_image_layout", "plan_kspace_layout",
           "plan_image_layout_from_sizes", "plan_kspace_layout_from_sizes"]
