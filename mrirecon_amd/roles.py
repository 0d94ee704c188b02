"""Axis roles of image-like tensors.

Interface of graspcg/core/roles.py:7-81: an image-like tensor is laid out as
(unlike..., like..., nufft...) and a regulariser names the axes it acts on
either by index or by a token that stands for one or more of those three
groups.  The tokens, the de-duplicated order of the result and the exception
types (KeyError for an unknown token, TypeError for a bad spec, IndexError for
an axis outside the tensor) are the reference's; the code is this package's.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, Iterable, Iterator, Sequence, Tuple, Union

AxesSpec = Union[str, Sequence[Union[int, str]]]

# token -> the axis groups it selects, in layout order
_GROUPS_OF_TOKEN: Dict[str, Tuple[str, ...]] = {}
for _names, _groups in ((("temporal", "time", "unlike"), ("unlike",)),
                        (("like",), ("like",)),
                        (("spatial", "image", "nufft"), ("nufft",)),
                        (("all",), ("unlike", "like", "nufft")),
                        (("none",), ())):
    for _n in _names:
        _GROUPS_OF_TOKEN[_n] = _groups
del _names, _groups, _n


@dataclass(frozen=True)
class Roles:
    """Counts of the three axis groups: x.shape == (unlike..., like..., nufft...)."""
    unlike: int
    like: int
    nufft: int

    def total(self) -> int:
        return int(self.unlike) + int(self.like) + int(self.nufft)

    @property
    def ndim(self) -> int:
        return self.total()

    def _group(self, name: str) -> Tuple[int, ...]:
        first = {"unlike": 0, "like": int(self.unlike), "nufft": int(self.unlike) + int(self.like)}[name]
        return tuple(range(first, first + int(getattr(self, name))))

    def _token(self, token: str) -> Tuple[int, ...]:
        groups = _GROUPS_OF_TOKEN.get(token.lower())
        if groups is None:
            raise KeyError(f"axes token {token!r} is not one of {sorted(_GROUPS_OF_TOKEN)}")
        return tuple(a for g in groups for a in self._group(g))

    def token_map(self) -> Dict[str, Tuple[int, ...]]:
        """Every token with the axes it resolves to."""
        return {t: self._token(t) for t in _GROUPS_OF_TOKEN}

    def _index(self, axis: int) -> int:
        n = self.total()
        a = axis + n if axis < 0 else axis
        if a < 0 or a >= n:
            raise IndexError(f"axis {axis} does not exist in a tensor with {n} axes")
        return a

    def _expand(self, items: Iterable) -> Iterator[int]:
        for it in items:
            if isinstance(it, str):
                yield from self._token(it)
            elif isinstance(it, int) and not isinstance(it, bool):
                yield self._index(it)
            else:
                raise TypeError(f"an axes entry must be an int or a token string, not {type(it).__name__}")

    def resolve_axes(self, spec: AxesSpec) -> Tuple[int, ...]:
        """Axis indices named by a token, or by a list/tuple of tokens and ints, first occurrence kept."""
        if isinstance(spec, str):
            return self._token(spec)
        if not isinstance(spec, (list, tuple)):
            raise TypeError(f"axes must be a token string or a list/tuple, not {type(spec).__name__}")
        seen: Dict[int, None] = {}
        for a in self._expand(spec):
            seen.setdefault(a, None)
        return tuple(seen)

    @property
    def temporal(self) -> Tuple[int, ...]:
        return self._group("unlike")

    @property
    def like_axes(self) -> Tuple[int, ...]:
        return self._group("like")

    @property
    def spatial(self) -> Tuple[int, ...]:
        return self._group("nufft")


# The solver's image layout (F, 1, S0, S1): frames, the size-1 coil slot, two spatial axes.
IMAGE_ROLES = Roles(unlike=1, like=1, nufft=2)
