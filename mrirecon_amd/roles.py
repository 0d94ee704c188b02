"""Axis roles of image-like tensors (same grammar as graspcg/core/roles.py:7-81)."""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, List, Sequence, Tuple, Union

AxesSpec = Union[str, Sequence[Union[int, str]]]


@dataclass(frozen=True)
class Roles:
    """x.shape == (unlike..., like..., nufft...)."""
    unlike: int
    like: int
    nufft: int

    def total(self) -> int:
        return int(self.unlike) + int(self.like) + int(self.nufft)

    def token_map(self) -> Dict[str, Tuple[int, ...]]:
        u, l, n = int(self.unlike), int(self.like), int(self.nufft)
        un, lk, sp = tuple(range(u)), tuple(range(u, u + l)), tuple(range(u + l, u + l + n))
        return {"temporal": un, "time": un, "unlike": un, "like": lk, "spatial": sp, "image": sp,
                "nufft": sp, "all": un + lk + sp, "none": tuple()}

    def resolve_axes(self, spec: AxesSpec) -> Tuple[int, ...]:
        tm, t = self.token_map(), self.total()
        if isinstance(spec, str):
            if spec.lower() not in tm:
                raise KeyError(f"Unknown axes token '{spec}'.")
            return tm[spec.lower()]
        if not isinstance(spec, (list, tuple)):
            raise TypeError(f"Bad axes spec type: {type(spec)}")
        out: List[int] = []
        for s in spec:
            if isinstance(s, str):
                if s.lower() not in tm:
                    raise KeyError(f"Unknown axes token '{s}'.")
                out.extend(tm[s.lower()])
            elif isinstance(s, int):
                ax = s if s >= 0 else t + s
                if not 0 <= ax < t:
                    raise IndexError(f"Axis {s} out of range for total dims {t}")
                out.append(ax)
            else:
                raise TypeError(f"Bad axes element type: {type(s)}")
        return tuple(dict.fromkeys(out))

    @property
    def temporal(self) -> Tuple[int, ...]:
        return self.token_map()["temporal"]

    @property
    def spatial(self) -> Tuple[int, ...]:
        return self.token_map()["spatial"]

    @property
    def ndim(self) -> int:
        return self.total()


# The solver's image layout (F, 1, S0, S1): frames, the size-1 coil slot, two spatial axes.
IMAGE_ROLES = Roles(unlike=1, like=1, nufft=2)
