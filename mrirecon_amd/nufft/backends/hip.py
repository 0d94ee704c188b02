"""'hip' NUFFT backend: the MI355X kernels behind the reference adapter surface.

Drop-in for the duck-typed adapters selected in graspcg/nufft/api.py:112-143
(TorchKbNUFFTAdapter, graspcg/nufft/backends/torchkb.py:9-323):

* ``prepare(im_shape, traj, dcf, maps, dtype, device, like_prod)`` -- torchkb.py:98-144
* ``k_per_frame()``                                              -- torchkb.py:146-147
* ``A(x, out=None, *, scratch=None, arena=None)``                -- torchkb.py:203-270
* ``AH(y, out=None, *, scratch=None, arena=None)``               -- torchkb.py:273-323
* ``diag_AHA_profile()`` / ``diag_AHA_scalar()``                 -- torchkb.py:329-364

Shapes follow the torchkb surface: A maps (B_eff, Cin in {1, C}, X, Y) or the
Like variant (B, L, Cin, X, Y) to (B_eff, C, K) / (B, L, C, K); AH maps
(B_eff, C, K) to (B_eff, 1, X, Y).  B_eff must be a multiple of the trajectory
batch B and batch b_eff uses frame b_eff // (B_eff / B), as torchkb's
``repeat_interleave`` does (torchkb.py:150-166).  DCF policy as
``_weights`` (torchkb.py:185-199).  Ragged trajectories (per-frame sample
counts) are accepted through ``k_offsets``; k-space is then a flat vector of
per-frame (C, K_b) blocks.

Errors: ValueError for shape / coil / batch mismatch (torchkb.py:225-234,
289-297), RuntimeError for a non-GPU device.  All work is enqueued on the
stream that is current at call time.
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass, field
from typing import Dict, List, Literal, Optional, Sequence, Tuple

import numpy as np
import torch

from ... import _lib

_DCF_CODES = {"none": _lib.DCF_NONE, "balanced": _lib.DCF_BALANCED, "standard": _lib.DCF_STANDARD}


class NufftPlan:
    """Owns one C-ABI plan (device footprints, twiddles, deapodisation, workspace)."""

    def __init__(self, *, S: Tuple[int, int], G: Tuple[int, int], W: int, C: int,
                 k_offsets: np.ndarray, traj_flat: np.ndarray, dcf_flat: Optional[np.ndarray],
                 dcf_mode: str, dcf_in_fwd: bool, dcf_in_adj: bool, maps: torch.Tensor,
                 like: int = 1, strip_cols: int = 0, beta: float = 0.0):
        lib = _lib.load()
        self._lib = lib
        self.k_offsets = np.ascontiguousarray(k_offsets, dtype=np.int64)
        self.traj_flat = np.ascontiguousarray(traj_flat, dtype=np.float32)
        self.dcf_flat = None if dcf_flat is None else np.ascontiguousarray(dcf_flat, dtype=np.float32)
        self.B = len(self.k_offsets) - 1
        self.C, self.S, self.G, self.W, self.like = int(C), tuple(S), tuple(G), int(W), int(like)
        self.maps = maps
        d = _lib.NufftDesc()
        d.ndim = 2
        d.S0, d.S1 = self.S
        d.G0, d.G1 = self.G
        d.W = self.W
        d.beta = float(beta)
        d.C, d.B, d.like = self.C, self.B, self.like
        d.k_offsets = self.k_offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        d.traj = self.traj_flat.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        d.dcf = (self.dcf_flat.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
                 if self.dcf_flat is not None else None)
        d.dcf_mode = _DCF_CODES[dcf_mode]
        d.dcf_in_fwd, d.dcf_in_adj = int(dcf_in_fwd), int(dcf_in_adj)
        d.maps = maps.data_ptr()
        d.strip_cols = int(strip_cols)
        h = ctypes.c_void_p()
        t0 = time.perf_counter()
        with torch.cuda.device(maps.device):
            _lib.check(lib.mrh_nufft_plan_create(ctypes.byref(d), ctypes.byref(h)))
        self.create_ms = (time.perf_counter() - t0) * 1e3     # host binning + uploads
        self.handle = h
        self.tuning = None

    def info(self) -> Dict[str, float]:
        inf = _lib.NufftInfo()
        _lib.check(self._lib.mrh_nufft_plan_info(self.handle, ctypes.byref(inf)))
        return {n: getattr(inf, n) for n, _ in inf._fields_}

    @property
    def n_k(self) -> int:
        return int(self.k_offsets[-1]) * self.C

    def forward(self, x: torch.Tensor, y: torch.Tensor, like: int = 1, x_has_coils: bool = False) -> None:
        _lib.check(self._lib.mrh_nufft_forward(self.handle, x.data_ptr(), y.data_ptr(), int(like),
                                               int(bool(x_has_coils)), _lib.stream_of(x)))

    def adjoint(self, y: torch.Tensor, x: torch.Tensor, like: int = 1) -> None:
        _lib.check(self._lib.mrh_nufft_adjoint(self.handle, y.data_ptr(), x.data_ptr(), int(like),
                                               _lib.stream_of(y)))

    def forward_masked(self, x: torch.Tensor, y: torch.Tensor, problem_active: int, frames_per_problem: int,
                       like: int = 1) -> None:
        """Forward over the frames of active problems only (``problem_active``: device int* address)."""
        _lib.check(self._lib.mrh_nufft_forward_masked(self.handle, x.data_ptr(), y.data_ptr(), int(like), 0,
                                                      int(problem_active), int(frames_per_problem),
                                                      _lib.stream_of(x)))

    def adjoint_masked(self, y: torch.Tensor, x: torch.Tensor, problem_active: int, frames_per_problem: int,
                       like: int = 1) -> None:
        """Adjoint over the frames of active problems only."""
        _lib.check(self._lib.mrh_nufft_adjoint_masked(self.handle, y.data_ptr(), x.data_ptr(), int(like),
                                                      int(problem_active), int(frames_per_problem),
                                                      _lib.stream_of(y)))

    def forward_stage(self, stage: int, x: torch.Tensor, y: torch.Tensor, like: int = 1) -> None:
        _lib.check(self._lib.mrh_nufft_forward_stage(self.handle, int(stage), x.data_ptr(), y.data_ptr(),
                                                     int(like), 0, _lib.stream_of(x)))

    def adjoint_stage(self, stage: int, y: torch.Tensor, x: torch.Tensor, like: int = 1) -> None:
        _lib.check(self._lib.mrh_nufft_adjoint_stage(self.handle, int(stage), y.data_ptr(), x.data_ptr(),
                                                     int(like), _lib.stream_of(y)))

    @property
    def coil_chunk(self) -> int:
        return int(self.info()["coil_chunk"])

    def set_coil_chunk(self, coils_per_block: int) -> None:
        _lib.check(self._lib.mrh_nufft_set_coil_chunk(self.handle, int(coils_per_block)))

    @torch.no_grad()
    def autotune(self, reps: int = 5, like: int = 1, candidates: Optional[Sequence[int]] = None) -> Dict[int, float]:
        """Pick the strip kernels' coils per workgroup by timing forward + adjoint pairs.

        Candidates are ceil(C / k) for k = 1..C (one workgroup per (frame, strip) down to one
        per coil), timed as ``tune_coil_chunks`` describes; the fastest is kept and
        ``self.tuning`` records every time (ms per pair).  Results do not depend on the choice."""
        t_start = time.perf_counter()
        times = tune_coil_chunks([self], like=like, reps=reps, candidates=candidates)[0]
        self.tuning = {"coil_chunk": self.coil_chunk, "ms_per_pair": times,
                       "autotune_ms": (time.perf_counter() - t_start) * 1e3}
        return times

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            self._lib.mrh_nufft_plan_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Nufft3Plan:
    """Owns one 3-D C-ABI plan (samples, deapodisation, twiddles, maps, grid workspace)."""

    def __init__(self, *, S: Tuple[int, int, int], G: Tuple[int, int, int], W: int, C: int,
                 k_offsets: np.ndarray, traj_flat: np.ndarray, dcf_flat: Optional[np.ndarray],
                 dcf_mode: str, dcf_in_fwd: bool, dcf_in_adj: bool, maps: torch.Tensor, beta: float = 0.0):
        lib = _lib.load()
        self._lib = lib
        self.k_offsets = np.ascontiguousarray(k_offsets, dtype=np.int64)
        self.traj_flat = np.ascontiguousarray(traj_flat, dtype=np.float32)
        self.dcf_flat = None if dcf_flat is None else np.ascontiguousarray(dcf_flat, dtype=np.float32)
        self.B = len(self.k_offsets) - 1
        self.C, self.S, self.G, self.W = int(C), tuple(S), tuple(G), int(W)
        self.maps = maps
        d = _lib.Nufft3Desc()
        d.S0, d.S1, d.S2 = self.S
        d.G0, d.G1, d.G2 = self.G
        d.W, d.beta, d.C, d.B = self.W, float(beta), self.C, self.B
        d.k_offsets = self.k_offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        d.traj = self.traj_flat.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        d.dcf = (self.dcf_flat.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
                 if self.dcf_flat is not None else None)
        d.dcf_mode = _DCF_CODES[dcf_mode]
        d.dcf_in_fwd, d.dcf_in_adj = int(dcf_in_fwd), int(dcf_in_adj)
        d.maps = maps.data_ptr()
        h = ctypes.c_void_p()
        with torch.cuda.device(maps.device):
            _lib.check(lib.mrh_nufft3_plan_create(ctypes.byref(d), ctypes.byref(h)))
        self.handle = h

    @property
    def n_k(self) -> int:
        return int(self.k_offsets[-1]) * self.C

    def forward(self, x: torch.Tensor, y: torch.Tensor) -> None:
        _lib.check(self._lib.mrh_nufft3_forward(self.handle, x.data_ptr(), y.data_ptr(), _lib.stream_of(x)))

    def adjoint(self, y: torch.Tensor, x: torch.Tensor) -> None:
        _lib.check(self._lib.mrh_nufft3_adjoint(self.handle, y.data_ptr(), x.data_ptr(), _lib.stream_of(y)))

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            self._lib.mrh_nufft3_plan_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@torch.no_grad()
def tune_coil_chunks(plans: Sequence["NufftPlan"], like: int = 1, reps: int = 5,
                     candidates: Optional[Sequence[int]] = None) -> List[Dict[int, float]]:
    """Time every (plan, coils per workgroup) candidate of ``plans`` (plans of one problem that
    differ in strip width) and leave each plan at its fastest chunk; returns each plan's
    {chunk: ms per pair}.

    Every candidate runs one warm-up pair, then ``reps`` rounds time one pair of each candidate
    in turn (HIP events on the current stream) and a candidate's time is its median.  Timed one
    candidate after another instead, the GPU's clock ramp during the sweep favoured whichever
    candidates ran last.  The inputs are fixed seeded Gaussian image and k-space data, each
    written to its own output: fed back into each other, A^H A would grow them about tenfold per
    pair and, over many candidates, to inf, which the adjoint's strips do not spread."""
    p0 = plans[0]
    S0, S1 = p0.S
    dev = p0.maps.device
    gen = torch.Generator(device=dev)
    gen.manual_seed(1234)
    x_in = torch.randn((p0.B * like, S0, S1), dtype=torch.complex64, device=dev, generator=gen)
    y_in = torch.randn(p0.n_k * like, dtype=torch.complex64, device=dev, generator=gen)
    x = torch.empty_like(x_in)
    y = torch.empty_like(y_in)
    cand = []
    for i, p in enumerate(plans):
        cs = (sorted({-(-p.C // k) for k in range(1, p.C + 1)}, reverse=True) if candidates is None
              else sorted({int(c) for c in candidates}, reverse=True))
        cand += [(i, c) for c in cs]

    def pair(i: int, c: int) -> None:
        plans[i].set_coil_chunk(c)
        plans[i].forward(x_in, y, like)
        plans[i].adjoint(y_in, x, like)

    for i, c in cand:
        pair(i, c)
    ts: Dict[Tuple[int, int], List[float]] = {ic: [] for ic in cand}
    for _ in range(max(1, reps)):
        for i, c in cand:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            pair(i, c)
            e1.record()
            e1.synchronize()
            ts[(i, c)].append(e0.elapsed_time(e1))
    out: List[Dict[int, float]] = [dict() for _ in plans]
    for (i, c), v in ts.items():
        out[i][c] = float(sorted(v)[len(v) // 2])
    for i, p in enumerate(plans):
        p.set_coil_chunk(min(out[i], key=out[i].get))
    return out


def _grid_size(S: int, osf: float) -> int:
    return int(math.ceil(osf * S))


def _kb_beta(W: int, osf: float) -> float:
    """Kaiser-Bessel beta from the configured oversampling (Beatty et al. 2005), as
    graspcg/nufft/backends/ptnufft.py:125 computes it: with a grid of ceil(osf S) points the
    ratio G / S can differ from osf, and the window follows osf."""
    return math.pi * math.sqrt(((W / float(osf)) * (float(osf) - 0.5)) ** 2 - 0.8)


@dataclass
class HipNUFFTAdapter:
    """MI355X KB-gridding backend with the torchkb adapter surface."""

    ndim: int
    backend_name: str = "hip"
    traj_units: Literal["rad", "norm"] = "rad"
    dcf_mode: Literal["balanced", "standard", "none"] = "balanced"
    apply_dcf_in_fwd: bool = False
    apply_dcf_in_adj: bool = True
    kb_width: int = 6          # torchkb numpoints=6 (torchkb.py:139)
    osf: float = 2.0           # torchkb default 2x grid
    strip_cols: int = 0        # 0 = chosen by the autotuner (or 8 with autotune=False)
    autotune: bool = True      # time strip widths and coil chunks per plan (NufftPlan.autotune)

    _im_shape: Optional[Tuple[int, ...]] = None
    _maps: Optional[torch.Tensor] = None
    _dev: Optional[torch.device] = None
    _dtype: torch.dtype = torch.complex64
    _B: int = 0
    _K: int = 0
    _k_offsets: Optional[np.ndarray] = None
    _traj_flat: Optional[np.ndarray] = None
    _dcf_flat: Optional[np.ndarray] = None
    _plans: Dict[int, NufftPlan] = field(default_factory=dict)
    _alpha_scalar: Optional[float] = None

    # ------------------------------------------------------------ prepare
    def prepare(self, im_shape: Tuple[int, ...], traj: torch.Tensor, dcf: Optional[torch.Tensor],
                maps: torch.Tensor, dtype: torch.dtype = torch.complex64,
                device: Optional[torch.device] = None, like_prod: int = 1, *,
                k_offsets: Optional[Sequence[int]] = None) -> None:
        """Build the device plan.  ``traj`` is (B, nd, K) | (B, K, nd) | (nd, K), or with
        ``k_offsets`` a flat (nd, sum K_b) array of per-frame blocks."""
        if self.ndim not in (2, 3):
            raise ValueError("the 'hip' backend supports ndim in {2, 3}")
        device = torch.device(device) if device is not None else maps.device
        if device.type != "cuda":
            raise RuntimeError(f"'hip' backend needs a GPU device, got {device}")
        if dtype not in (torch.complex64, torch.complex128):
            raise ValueError(f"unsupported dtype {dtype}")
        self._im_shape = tuple(int(s) for s in im_shape)
        nd = self.ndim
        if len(self._im_shape) != nd + 1:
            raise ValueError(f"im_shape must be (C, *S) with {nd} spatial axes, got {self._im_shape}")
        C = self._im_shape[0]
        self._dev, self._dtype = device, dtype
        self._maps = maps.to(device=device, dtype=torch.complex64).contiguous()
        if tuple(self._maps.shape) != self._im_shape:
            raise ValueError(f"maps shape {tuple(self._maps.shape)} != {self._im_shape}")
        t = traj.detach().to("cpu", torch.float32)
        scale = 2.0 * math.pi if self.traj_units == "norm" else 1.0
        if k_offsets is not None:
            off = np.asarray(k_offsets, dtype=np.int64)
            if t.ndim != 2 or int(t.shape[0]) != nd or int(t.shape[1]) != int(off[-1]):
                raise ValueError(f"ragged traj must be ({nd}, sum K_b) matching k_offsets")
            B = len(off) - 1
            tn = t.numpy() * scale                       # (nd, sum K) -> per-frame (nd, K_b) blocks
            flat = np.concatenate([tn[:, off[b]:off[b + 1]].reshape(-1) for b in range(B)])
        else:
            if t.ndim == 2:
                if int(t.shape[0]) != nd and int(t.shape[1]) == nd:
                    t = t.t()
                if int(t.shape[0]) != nd:
                    raise ValueError(f"traj has wrong nd; expected {nd}")
                t = t.unsqueeze(0)
            elif t.ndim == 3:
                if int(t.shape[1]) != nd:
                    if int(t.shape[2]) == nd:
                        t = t.transpose(1, 2)
                    else:
                        raise ValueError("traj must be (B,nd,K) or (B,K,nd)")
            else:
                raise ValueError("traj must have 2 or 3 dims")
            B, _, K = t.shape
            off = np.arange(B + 1, dtype=np.int64) * int(K)
            flat = t.contiguous().numpy() * scale           # (B, nd, K) == per-frame (nd, K) blocks
        self._B, self._k_offsets = int(B), off
        ks = np.diff(off)
        self._K = int(ks[0]) if np.all(ks == ks[0]) else -1
        self._traj_flat = np.ascontiguousarray(flat.reshape(-1), dtype=np.float32)
        if dcf is None:
            self._dcf_flat = None
        else:
            d = dcf.detach().to("cpu", torch.float32)
            if k_offsets is None:
                if d.ndim == 1:
                    d = d.view(1, -1)
                if d.ndim != 2:
                    raise ValueError("dcf must be (B,K) or (K)")
                if int(d.shape[-1]) != self._K:
                    raise ValueError("dcf must have K samples")
                if int(d.shape[0]) == 1 and self._B > 1:
                    d = d.expand(self._B, -1)
                if int(d.shape[0]) != self._B:
                    raise ValueError(f"dcf batch {int(d.shape[0])} incompatible with traj batch {self._B}")
            self._dcf_flat = np.ascontiguousarray(d.reshape(-1).numpy(), dtype=np.float32)
            if self._dcf_flat.size != int(off[-1]):
                raise ValueError("dcf length does not match the trajectory")
        for p in self._plans.values():
            p.close()
        self._plans = {}
        self._alpha_scalar = None
        self._plan(1)

    def _new_plan(self, like: int, strip_cols: int):
        if self.ndim == 3:
            if like != 1:
                raise NotImplementedError("3-D plans take one like image per frame")
            C, S0, S1, S2 = self._im_shape
            return Nufft3Plan(S=(S0, S1, S2), G=tuple(_grid_size(v, self.osf) for v in (S0, S1, S2)),
                              W=self.kb_width, C=C, k_offsets=self._k_offsets, traj_flat=self._traj_flat,
                              dcf_flat=self._dcf_flat, dcf_mode=self.dcf_mode, dcf_in_fwd=self.apply_dcf_in_fwd,
                              dcf_in_adj=self.apply_dcf_in_adj, maps=self._maps,
                              beta=_kb_beta(self.kb_width, self.osf))
        C, S0, S1 = self._im_shape
        return NufftPlan(S=(S0, S1), G=(_grid_size(S0, self.osf), _grid_size(S1, self.osf)),
                         W=self.kb_width, C=C, k_offsets=self._k_offsets, traj_flat=self._traj_flat,
                         dcf_flat=self._dcf_flat, dcf_mode=self.dcf_mode,
                         dcf_in_fwd=self.apply_dcf_in_fwd, dcf_in_adj=self.apply_dcf_in_adj,
                         maps=self._maps, like=like, strip_cols=strip_cols,
                         beta=_kb_beta(self.kb_width, self.osf))

    def _plan(self, like: int) -> NufftPlan:
        """The plan for ``like`` images per frame.  With ``autotune`` and ``strip_cols = 0`` every
        strip width the library builds for this grid and KB width (4, 8, 16 grid columns) gets a
        plan whose coil chunk is tuned; the fastest plan is kept (``plan.tuning`` lists all times)."""
        p = self._plans.get(like)
        if p is not None:
            return p
        if self.ndim == 3:                       # no strip kernels to tune
            p = self._new_plan(like, 0)
            self._plans[like] = p
            return p
        if not self.autotune:
            p = self._new_plan(like, self.strip_cols)
        elif self.strip_cols:
            p = self._new_plan(like, self.strip_cols)
            p.autotune(like=like)
        else:
            G0 = _grid_size(self._im_shape[1], self.osf)
            widths = [8] + ([16] if G0 <= 512 else []) + ([4] if self.kb_width == 4 else [])
            qs, create_ms = [], {}
            t_setup = time.perf_counter()
            for tc in widths:
                try:
                    q = self._new_plan(like, tc)
                except (NotImplementedError, ValueError):
                    if tc == widths[0]:           # the default width: the plan itself is invalid
                        raise
                    continue                      # a width this grid cannot use (G1 % tc, KB terms)
                qs.append((tc, q))
                create_ms[tc] = round(q.create_ms, 2)
            t_tune = time.perf_counter()
            tuned = tune_coil_chunks([q for _, q in qs], like=like)
            tune_ms = round((time.perf_counter() - t_tune) * 1e3, 2)
            times = {tc: t for (tc, _), t in zip(qs, tuned)}
            k = min(range(len(qs)), key=lambda j: min(tuned[j].values()))
            p = qs[k][1]
            for j, (_, q) in enumerate(qs):
                if j != k:
                    q.close()
            p.tuning = {"strip_cols": int(p.info()["strip_cols"]), "coil_chunk": p.coil_chunk,
                        "ms_per_pair": dict(tuned[k]),
                        "ms_per_pair_by_width": {tc: dict(t) for tc, t in times.items()},
                        "plan_create_ms_by_width": create_ms, "autotune_ms": tune_ms,
                        "setup_ms": (time.perf_counter() - t_setup) * 1e3}
        self._plans[like] = p
        return p

    @property
    def plan(self) -> NufftPlan:
        return self._plan(1)

    def k_per_frame(self) -> int:
        return int(self._K)

    @property
    def n_frames(self) -> int:
        return self._B

    # ------------------------------------------------------------ helpers
    def _like_of(self, Bx: int) -> int:
        if Bx % self._B != 0:
            raise ValueError(f"omega batch {self._B} incompatible with requested {Bx}.")
        return Bx // self._B

    @staticmethod
    def _as_c64(t: torch.Tensor) -> torch.Tensor:
        t = t if t.dtype == torch.complex64 else t.to(torch.complex64)
        return t.contiguous()

    def _check_dev(self, t: torch.Tensor) -> None:
        if t.device.type != "cuda":
            raise RuntimeError(f"'hip' backend needs GPU tensors, got {t.device}")

    # ------------------------------------------------------------ operators
    @torch.no_grad()
    def A(self, x: torch.Tensor, out: Optional[torch.Tensor] = None, *, scratch=None, arena=None) -> torch.Tensor:
        """Forward NUFFT: image -> per-coil k-space (maps applied when Cin == 1)."""
        if self._im_shape is None:
            raise RuntimeError("prepare() was not called")
        self._check_dev(x)
        if self.ndim == 3:
            return self._A3(x, out)
        C = self._im_shape[0]
        spatial = self._im_shape[1:]
        if x.ndim == 4:
            Bx, Cin = int(x.shape[0]), int(x.shape[1])
            L = 1
        elif x.ndim == 5:
            Bx, L, Cin = int(x.shape[0]) * int(x.shape[1]), int(x.shape[1]), int(x.shape[2])
        else:
            raise ValueError(f"x must be (B,C,*S) or (B,L,C,*S); got shape {tuple(x.shape)} for ndim=2")
        if tuple(int(s) for s in x.shape[-2:]) != tuple(spatial):
            raise ValueError(f"x spatial shape {tuple(x.shape[-2:])} != {tuple(spatial)}")
        if Cin not in (1, C):
            raise ValueError(f"image coil slot must be 1 or {C}; got {Cin}")
        if self._K < 0:
            need = (self._plan(1).n_k * (Bx // self._B),)
        else:
            need = (int(x.shape[0]), C, self._K) if x.ndim == 4 else (int(x.shape[0]), L, C, self._K)
        like = self._like_of(Bx)
        if out is None:
            out = torch.empty(need, dtype=x.dtype, device=x.device)
        elif tuple(out.shape) != tuple(need):
            raise ValueError(f"out has shape {tuple(out.shape)} but expected {tuple(need)}")
        xc = self._as_c64(x)
        direct = out.dtype == torch.complex64 and out.is_contiguous()
        yb = out if direct else torch.empty(need, dtype=torch.complex64, device=x.device)
        # Cin == 1 always multiplies by the maps (torchkb.py:243-246)
        self._plan(like).forward(xc, yb, like=like, x_has_coils=(Cin == C and C != 1))
        if not direct:
            out.copy_(yb)
        return out

    @torch.no_grad()
    def AH(self, y: torch.Tensor, out: Optional[torch.Tensor] = None, *, scratch=None, arena=None) -> torch.Tensor:
        """Adjoint NUFFT: per-coil k-space -> coil-combined image."""
        if self._im_shape is None:
            raise RuntimeError("prepare() was not called")
        self._check_dev(y)
        if self.ndim == 3:
            return self._AH3(y, out)
        C = self._im_shape[0]
        spatial = tuple(self._im_shape[1:])
        if self._K < 0:
            if y.ndim != 1 or int(y.numel()) % self._plan(1).n_k != 0:
                raise ValueError("ragged k-space must be a flat vector of per-frame (C, K_b) blocks")
            like = int(y.numel()) // self._plan(1).n_k
            need = (self._B * like, 1, *spatial)
        elif y.ndim == 3:
            if int(y.shape[1]) != C:
                raise ValueError(f"y coil dim {int(y.shape[1])} must equal maps C {C}")
            like = self._like_of(int(y.shape[0]))
            need = (int(y.shape[0]), 1, *spatial)
        elif y.ndim == 4:
            if int(y.shape[2]) != C:
                raise ValueError(f"y coil dim {int(y.shape[2])} must equal maps C {C}")
            like = self._like_of(int(y.shape[0]) * int(y.shape[1]))
            need = (int(y.shape[0]), int(y.shape[1]), 1, *spatial)
        else:
            raise ValueError("y must be (B,C,K) or (B,L,C,K)")
        if y.ndim >= 3 and int(y.shape[-1]) != self._K:
            raise ValueError(f"y has K={int(y.shape[-1])}, plan has K={self._K}")
        if out is None:
            out = torch.empty(need, dtype=y.dtype, device=y.device)
        elif tuple(out.shape) != tuple(need):
            raise ValueError(f"out has shape {tuple(out.shape)} but expected {tuple(need)}")
        yc = self._as_c64(y)
        direct = out.dtype == torch.complex64 and out.is_contiguous()
        xb = out if direct else torch.empty(need, dtype=torch.complex64, device=y.device)
        self._plan(like).adjoint(yc, xb, like=like)
        if not direct:
            out.copy_(xb)
        return out

    # ------------------------------------------------------------ 3-D operators
    def _A3(self, x: torch.Tensor, out: Optional[torch.Tensor]) -> torch.Tensor:
        """(B, 1, X, Y, Z) volumes -> (B, C, K) (flat per-frame (C, K_b) blocks when ragged)."""
        C, spatial = self._im_shape[0], tuple(self._im_shape[1:])
        if x.ndim != 5 or int(x.shape[1]) != 1 or tuple(int(v) for v in x.shape[2:]) != spatial:
            raise ValueError(f"x must be (B, 1, *{spatial}) for ndim=3; got {tuple(x.shape)}")
        if int(x.shape[0]) != self._B:
            raise ValueError(f"x batch {int(x.shape[0])} != trajectory frames {self._B} (one like image)")
        plan = self._plan(1)
        need = (plan.n_k,) if self._K < 0 else (self._B, C, self._K)
        if out is None:
            out = torch.empty(need, dtype=x.dtype, device=x.device)
        elif tuple(out.shape) != tuple(need):
            raise ValueError(f"out has shape {tuple(out.shape)} but expected {tuple(need)}")
        direct = out.dtype == torch.complex64 and out.is_contiguous()
        yb = out if direct else torch.empty(need, dtype=torch.complex64, device=x.device)
        plan.forward(self._as_c64(x), yb)
        if not direct:
            out.copy_(yb)
        return out

    def _AH3(self, y: torch.Tensor, out: Optional[torch.Tensor]) -> torch.Tensor:
        """(B, C, K) or flat ragged k-space -> (B, 1, X, Y, Z)."""
        C, spatial = self._im_shape[0], tuple(self._im_shape[1:])
        plan = self._plan(1)
        if int(y.numel()) != plan.n_k:
            raise ValueError(f"y has {int(y.numel())} samples, the plan expects {plan.n_k}")
        if y.ndim == 3 and (int(y.shape[0]) != self._B or int(y.shape[1]) != C):
            raise ValueError(f"y must be ({self._B}, {C}, K)")
        need = (self._B, 1) + spatial
        if out is None:
            out = torch.empty(need, dtype=y.dtype, device=y.device)
        elif tuple(out.shape) != tuple(need):
            raise ValueError(f"out has shape {tuple(out.shape)} but expected {tuple(need)}")
        direct = out.dtype == torch.complex64 and out.is_contiguous()
        xb = out if direct else torch.empty(need, dtype=torch.complex64, device=y.device)
        plan.adjoint(self._as_c64(y), xb)
        if not direct:
            out.copy_(xb)
        return out

    # ------------------------------------------------------------ calibration
    @torch.no_grad()
    def diag_AHA_profile(self) -> torch.Tensor:
        """Per-frame alpha_b ~ sum_k dcf[b, k] (or K_b without DCF) (torchkb.py:329-340)."""
        off = self._k_offsets
        if self._dcf_flat is None:
            v = np.diff(off).astype(np.float32)
        else:
            v = np.array([self._dcf_flat[off[b]:off[b + 1]].sum() for b in range(self._B)], np.float32)
        return torch.from_numpy(v).to(self._dev)

    @torch.no_grad()
    def diag_AHA_scalar(self) -> float:
        """alpha = A^H A(delta_centre) / sum_c |S_c|^2(centre), cached (torchkb.py:342-364)."""
        if self._alpha_scalar is not None:
            return float(self._alpha_scalar)
        if self.ndim == 3:                       # maps applied (one image slot): A^H A delta = alpha sum |S_c|^2
            spatial = tuple(self._im_shape[1:])
            ctr = tuple(v // 2 for v in spatial)
            x0 = torch.zeros((self._B, 1) + spatial, device=self._dev, dtype=torch.complex64)
            x0[(0, 0) + ctr] = 1.0
            z = self.AH(self.A(x0))
            zc = float(z[(0, 0) + ctr].real.item())
            sos = float((self._maps.abs() ** 2).sum(dim=0)[ctr].clamp_min(1e-20).item())
            self._alpha_scalar = zc / sos
            return self._alpha_scalar
        C, S0, S1 = self._im_shape
        ctr = (S0 // 2, S1 // 2)
        x0 = torch.zeros((self._B, C, S0, S1), device=self._dev, dtype=torch.complex64)
        x0[0, 0][ctr] = 1.0
        z = self.AH(self.A(x0))
        zc = float(z[0, 0][ctr].real.item())
        sos = float((self._maps.abs() ** 2).sum(dim=0)[ctr].clamp_min(1e-20).item())
        self._alpha_scalar = zc / sos
        return self._alpha_scalar

    def close(self) -> None:
        for p in self._plans.values():
            p.close()
        self._plans = {}
