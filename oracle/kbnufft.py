"""ORACLE (test infrastructure only) -- CPU restatement of the KB-gridding NUFFT.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module, and only as the checker / the timed CPU baseline.  The product
path (mrirecon_amd) never imports it.

What it restates
----------------
* The operator contract of the reference's default backend
  (graspcg/nufft/backends/torchkb.py:203-323): forward multiplies the
  coil-combined image by each coil map, NUFFTs it and applies sqrt(dcf) in
  'balanced' mode (torchkb.py:185-199, 245-254); the adjoint applies sqrt(dcf),
  NUFFT^H's each coil and combines with conj(maps) (torchkb.py:305-313).
* The NUFFT semantics of the reference's ground-truth test: centred
  coordinates r = i - S//2 and kernel exp(-i w.r) in radians
  (graspcg/tests/test_nufft_precision_gt.py:48-121) -> :func:`nudft_forward`,
  :func:`nudft_adjoint`, used as the closed-form pin.
* Kaiser-Bessel gridding with the kernel and beta formula of the reference's
  in-repo gridder (graspcg/nufft/backends/ptnufft.py:79-90 and :125):
  phi(d) = I0(beta sqrt(1-(2d/W)^2)) / I0(beta), beta = pi sqrt((W/osf (osf-1/2))^2 - 0.8).
  Unlike ptnufft (whose grid placement and missing forward deapodisation make it
  no faithful NUFFT), DC sits at grid index 0, the image is placed at
  (i - S//2) mod G, and deapodisation uses the analytic Fourier transform of
  the KB window (Jackson et al. 1991):
  Phi(xi) = W sinh(sqrt(beta^2 - (pi W xi)^2)) / sqrt(beta^2 - (pi W xi)^2) / I0(beta).
* Footprint: for grid coordinate u = w G / (2 pi), the W integer nodes
  m0 = floor(u - W/2) + 1, ..., m0 + W - 1 (so every distance |u - m| <= W/2).

Pinning (tests/test_oracle_nufft.py): closeness to the FP64 NUDFT at the
reference test's problem, adjointness <Ax, y> = <x, A^H y> to FP64 round-off,
the closed-form KB transform against numerical quadrature, delta-image KATs
(centre -> S_c(ctr) sqrt(dcf); shifted delta -> plane wave), FFT placement
against numpy's direct DFT on tiny grids.
"""
from __future__ import annotations

import math
from typing import Optional, Sequence

import numpy as np


# --------------------------------------------------------------- KB window

def kb_beta(W: int, osf: float) -> float:
    """Beatty 2005 beta, as written at graspcg/nufft/backends/ptnufft.py:125."""
    return math.pi * math.sqrt(((W / float(osf)) * (float(osf) - 0.5)) ** 2 - 0.8)


def kb_phi(d, W: int, beta: float):
    """KB window at distance d (grid units), zero outside |d| <= W/2 (ptnufft.py:79-90)."""
    d = np.asarray(d, dtype=np.float64)
    arg = 1.0 - (2.0 * d / W) ** 2
    out = np.i0(beta * np.sqrt(np.clip(arg, 0.0, None))) / np.i0(beta)
    return np.where(np.abs(d) <= W / 2.0, out, 0.0)


def kb_ft(xi, W: int, beta: float):
    """Continuous Fourier transform of :func:`kb_phi` at frequency xi (cycles / grid unit)."""
    xi = np.asarray(xi, dtype=np.float64)
    z = beta * beta - (math.pi * W * xi) ** 2
    sz = np.sqrt(np.abs(z))
    with np.errstate(invalid="ignore", divide="ignore"):
        pos = np.sinh(sz) / sz
        neg = np.sin(sz) / sz
    val = np.where(z > 0, pos, np.where(z < 0, neg, 1.0))
    return W * val / np.i0(beta)


def deapod_1d(S: int, G: int, W: int, beta: float) -> np.ndarray:
    """1 / Phi((i - S//2) / G) for i = 0..S-1 (float64)."""
    r = np.arange(S, dtype=np.float64) - S // 2
    return 1.0 / kb_ft(r / G, W, beta)


def footprint(u: np.ndarray, W: int, beta: float):
    """Start node m0 (int64) and weights (n, W) of the W-point footprint at coordinate u."""
    u = np.asarray(u, dtype=np.float64)
    m0 = np.floor(u - W / 2.0).astype(np.int64) + 1
    d = u[:, None] - (m0[:, None] + np.arange(W)[None, :])
    return m0, kb_phi(d, W, beta)


# --------------------------------------------------------------- operator

class KBNufft:
    """2-D KB-gridding SENSE operator on (B, S0, S1) images and flat k-space.

    k-space layout: frame b owns the block ``y[C*k_offsets[b] : C*k_offsets[b+1]]``
    viewed as (C, K_b); with uniform K this is exactly (B, C, K) contiguous.
    ``dtype`` selects float64 (oracle) or float32 (CPU baseline port).
    """

    def __init__(self, maps: np.ndarray, traj: np.ndarray, k_offsets: np.ndarray,
                 dcf: Optional[np.ndarray], W: int, osf: float = 2.0,
                 dcf_mode: str = "balanced", like: int = 1, dtype=np.float64,
                 dcf_in_fwd: bool = False, dcf_in_adj: bool = True):
        self.rdt = np.dtype(dtype)
        self.cdt = np.complex128 if self.rdt == np.float64 else np.complex64
        self.maps = np.asarray(maps).astype(self.cdt)
        self.C, self.S0, self.S1 = self.maps.shape
        self.G0 = int(math.ceil(osf * self.S0))
        self.G1 = int(math.ceil(osf * self.S1))
        self.W = int(W)
        self.beta = kb_beta(W, osf)
        self.off = np.asarray(k_offsets, dtype=np.int64)
        self.B = len(self.off) - 1
        self.like = int(like)
        self.traj = np.asarray(traj, dtype=np.float64)
        if dcf is None or dcf_mode == "none":
            self.wA = self.wAH = None
        elif dcf_mode == "balanced":
            w = np.sqrt(np.clip(np.asarray(dcf, dtype=np.float64), 0.0, None))
            self.wA = self.wAH = w
        elif dcf_mode == "standard":     # dcf on the sides the flags name (torchkb.py:192-197)
            d = np.asarray(dcf, dtype=np.float64)
            self.wA = d if dcf_in_fwd else None
            self.wAH = d if dcf_in_adj else None
        else:
            raise ValueError(dcf_mode)
        dx = deapod_1d(self.S0, self.G0, self.W, self.beta)
        dy = deapod_1d(self.S1, self.G1, self.W, self.beta)
        self.deapod = (dx[:, None] * dy[None, :]).astype(self.rdt)
        self._fp = []
        for b in range(self.B):
            s = slice(self.off[b], self.off[b + 1])
            u = self.traj[0, s] * self.G0 / (2.0 * math.pi)
            v = self.traj[1, s] * self.G1 / (2.0 * math.pi)
            mx, wx = footprint(u, self.W, self.beta)
            my, wy = footprint(v, self.W, self.beta)
            ix = (mx[:, None] + np.arange(self.W)[None, :]) % self.G0
            iy = (my[:, None] + np.arange(self.W)[None, :]) % self.G1
            flat = (ix[:, :, None] * self.G1 + iy[:, None, :]).reshape(len(u), self.W * self.W)
            wgt = (wx[:, :, None] * wy[:, None, :]).reshape(len(u), self.W * self.W).astype(self.rdt)
            self._fp.append((flat, wgt))
        r0 = (np.arange(self.S0) - self.S0 // 2) % self.G0
        r1 = (np.arange(self.S1) - self.S1 // 2) % self.G1
        self._rows, self._cols = r0, r1

    @property
    def n_k(self) -> int:
        return int(self.off[-1]) * self.C * self.like

    def _kblock(self, b: int, l: int):
        K = int(self.off[b + 1] - self.off[b])
        base = self.C * (int(self.off[b]) * self.like + l * K)
        return base, K

    def forward(self, x: np.ndarray) -> np.ndarray:
        """x (B*like, S0, S1) -> flat k-space; like index l of frame b is batch b*like + l."""
        x = np.asarray(x).astype(self.cdt).reshape(self.B * self.like, self.S0, self.S1)
        y = np.zeros(self.n_k, dtype=self.cdt)
        for bb in range(self.B * self.like):
            b, l = divmod(bb, self.like)
            flat, wgt = self._fp[b]
            img = x[bb][None] * self.maps * self.deapod[None]
            grid = np.zeros((self.C, self.G0, self.G1), dtype=self.cdt)
            grid[:, self._rows[:, None], self._cols[None, :]] = img
            grid = np.fft.fft2(grid, axes=(1, 2)).reshape(self.C, -1)
            vals = np.einsum("ckw,kw->ck", grid[:, flat], wgt)
            s = slice(self.off[b], self.off[b + 1])
            if self.wA is not None:
                vals = vals * self.wA[s].astype(self.rdt)[None]
            base, K = self._kblock(b, l)
            y[base:base + self.C * K] = vals.reshape(-1)
        return y

    def adjoint(self, y: np.ndarray) -> np.ndarray:
        """flat k-space -> (B*like, S0, S1) coil-combined image."""
        y = np.asarray(y).astype(self.cdt).reshape(-1)
        out = np.zeros((self.B * self.like, self.S0, self.S1), dtype=self.cdt)
        G = self.G0 * self.G1
        for bb in range(self.B * self.like):
            b, l = divmod(bb, self.like)
            flat, wgt = self._fp[b]
            base, K = self._kblock(b, l)
            yk = y[base:base + self.C * K].reshape(self.C, K)
            s = slice(self.off[b], self.off[b + 1])
            if self.wAH is not None:
                yk = yk * self.wAH[s].astype(self.rdt)[None]
            acc = np.zeros((self.S0, self.S1), dtype=self.cdt)
            for c in range(self.C):
                contrib = (yk[c][:, None] * wgt).reshape(-1)
                fl = flat.reshape(-1)
                grid = (np.bincount(fl, weights=contrib.real, minlength=G)
                        + 1j * np.bincount(fl, weights=contrib.imag, minlength=G))
                grid = grid.astype(self.cdt).reshape(self.G0, self.G1)
                img = np.fft.ifft2(grid) * G
                acc += img[self._rows[:, None], self._cols[None, :]] * np.conj(self.maps[c])
            out[bb] = acc * self.deapod
        return out


class KBNufft3:
    """3-D KB-gridding SENSE operator on (B, S0, S1, S2) volumes (the ndim=3 case of the
    reference front end, graspcg/nufft/api.py:85-89, with the 2-D conventions above applied
    per axis).  k-space: frame b owns ``y[C*k_offsets[b] : C*k_offsets[b+1]]`` as (C, K_b);
    ``traj`` is (3, sum K_b) in radians."""

    def __init__(self, maps: np.ndarray, traj: np.ndarray, k_offsets: np.ndarray,
                 dcf: Optional[np.ndarray], W: int, osf: float = 2.0, dcf_mode: str = "balanced"):
        self.maps = np.asarray(maps).astype(np.complex128)
        self.C = self.maps.shape[0]
        self.S = tuple(int(v) for v in self.maps.shape[1:])
        self.G = tuple(int(math.ceil(osf * v)) for v in self.S)
        self.W = int(W)
        self.beta = kb_beta(W, osf)
        self.off = np.asarray(k_offsets, dtype=np.int64)
        self.B = len(self.off) - 1
        self.traj = np.asarray(traj, dtype=np.float64)
        if dcf is None or dcf_mode == "none":
            self.wA = self.wAH = None
        elif dcf_mode == "balanced":
            self.wA = self.wAH = np.sqrt(np.clip(np.asarray(dcf, dtype=np.float64), 0.0, None))
        elif dcf_mode == "standard":
            self.wA, self.wAH = None, np.asarray(dcf, dtype=np.float64)
        else:
            raise ValueError(dcf_mode)
        d = [deapod_1d(S, G, self.W, self.beta) for S, G in zip(self.S, self.G)]
        self.deapod = d[0][:, None, None] * d[1][None, :, None] * d[2][None, None, :]
        self.idx = [(np.arange(S) - S // 2) % G for S, G in zip(self.S, self.G)]
        self._fp = []
        for b in range(self.B):
            sl = slice(self.off[b], self.off[b + 1])
            ms, ws = [], []
            for a in range(3):
                m, w = footprint(self.traj[a, sl] * self.G[a] / (2.0 * math.pi), self.W, self.beta)
                ms.append((m[:, None] + np.arange(self.W)[None, :]) % self.G[a])
                ws.append(w)
            n, W3 = len(ms[0]), self.W ** 3
            flat = ((ms[0][:, :, None, None] * self.G[1] + ms[1][:, None, :, None]) * self.G[2]
                    + ms[2][:, None, None, :]).reshape(n, W3)
            wgt = (ws[0][:, :, None, None] * ws[1][:, None, :, None] * ws[2][:, None, None, :]).reshape(n, W3)
            self._fp.append((flat, wgt))

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.asarray(x, dtype=np.complex128).reshape(self.B, *self.S)
        y = np.zeros(int(self.off[-1]) * self.C, dtype=np.complex128)
        i0, i1, i2 = np.ix_(*self.idx)
        for b in range(self.B):
            flat, wgt = self._fp[b]
            grid = np.zeros((self.C,) + self.G, dtype=np.complex128)
            grid[:, i0, i1, i2] = x[b][None] * self.maps * self.deapod[None]
            grid = np.fft.fftn(grid, axes=(1, 2, 3)).reshape(self.C, -1)
            vals = np.einsum("ckw,kw->ck", grid[:, flat], wgt)
            sl = slice(self.off[b], self.off[b + 1])
            if self.wA is not None:
                vals = vals * self.wA[sl][None]
            y[self.C * self.off[b]:self.C * self.off[b + 1]] = vals.reshape(-1)
        return y

    def adjoint(self, y: np.ndarray) -> np.ndarray:
        y = np.asarray(y, dtype=np.complex128).reshape(-1)
        out = np.zeros((self.B,) + self.S, dtype=np.complex128)
        Gt = int(np.prod(self.G))
        i0, i1, i2 = np.ix_(*self.idx)
        for b in range(self.B):
            flat, wgt = self._fp[b]
            K = int(self.off[b + 1] - self.off[b])
            yk = y[self.C * self.off[b]:self.C * self.off[b + 1]].reshape(self.C, K)
            sl = slice(self.off[b], self.off[b + 1])
            if self.wAH is not None:
                yk = yk * self.wAH[sl][None]
            acc = np.zeros(self.S, dtype=np.complex128)
            for c in range(self.C):
                contrib = (yk[c][:, None] * wgt).reshape(-1)
                fl = flat.reshape(-1)
                grid = (np.bincount(fl, weights=contrib.real, minlength=Gt)
                        + 1j * np.bincount(fl, weights=contrib.imag, minlength=Gt)).reshape(self.G)
                img = np.fft.ifftn(grid) * Gt
                acc += img[i0, i1, i2] * np.conj(self.maps[c])
            out[b] = acc * self.deapod
        return out


def nudft3_forward(maps: np.ndarray, x: np.ndarray, traj_b: np.ndarray) -> np.ndarray:
    """FP64 3-D NUDFT of one frame (the 2-D ground truth of graspcg/tests/test_nufft_precision_gt.py:55-89
    with a third centred coordinate): maps (C, S0, S1, S2), x (S0, S1, S2), traj_b (3, K) -> (C, K)."""
    C = maps.shape[0]
    S = maps.shape[1:]
    r = [np.arange(s, dtype=np.float64) - s // 2 for s in S]
    R = np.stack([g.reshape(-1) for g in np.meshgrid(*r, indexing="ij")], axis=1)
    E = np.exp(-1j * (np.asarray(traj_b, dtype=np.float64).T @ R.T))
    img = (np.asarray(maps, dtype=np.complex128) * np.asarray(x, dtype=np.complex128)[None]).reshape(C, -1)
    return img @ E.T


# --------------------------------------------------------------- NUDFT (closed form)

def nudft_forward(maps: np.ndarray, x: np.ndarray, traj_b: np.ndarray,
                  dcf_b: Optional[np.ndarray] = None) -> np.ndarray:
    """FP64 NUDFT of one frame, graspcg/tests/test_nufft_precision_gt.py:55-89.

    maps (C, S0, S1); x (L, S0, S1); traj_b (2, K) radians; returns (L, C, K)."""
    C, S0, S1 = maps.shape
    r0 = np.arange(S0, dtype=np.float64) - S0 // 2
    r1 = np.arange(S1, dtype=np.float64) - S1 // 2
    R0, R1 = np.meshgrid(r0, r1, indexing="ij")
    coords = np.stack([R0.reshape(-1), R1.reshape(-1)], axis=1)
    E = np.exp(-1j * (np.asarray(traj_b, dtype=np.float64).T @ coords.T))       # (K, R)
    xl = np.asarray(x, dtype=np.complex128).reshape(x.shape[0], -1)
    m = np.asarray(maps, dtype=np.complex128).reshape(C, -1)
    y = np.einsum("lr,cr,kr->lck", xl, m, E)
    if dcf_b is not None:
        y = y * np.sqrt(np.clip(np.asarray(dcf_b, dtype=np.float64), 0.0, None))[None, None]
    return y


def nudft_adjoint(maps: np.ndarray, y: np.ndarray, traj_b: np.ndarray,
                  dcf_b: Optional[np.ndarray] = None) -> np.ndarray:
    """FP64 adjoint NUDFT of one frame, graspcg/tests/test_nufft_precision_gt.py:91-121.

    y (L, C, K) -> (L, S0, S1) coil-combined."""
    C, S0, S1 = maps.shape
    r0 = np.arange(S0, dtype=np.float64) - S0 // 2
    r1 = np.arange(S1, dtype=np.float64) - S1 // 2
    R0, R1 = np.meshgrid(r0, r1, indexing="ij")
    coords = np.stack([R0.reshape(-1), R1.reshape(-1)], axis=1)
    Ec = np.exp(1j * (np.asarray(traj_b, dtype=np.float64).T @ coords.T))         # (K, R)
    yy = np.asarray(y, dtype=np.complex128)
    if dcf_b is not None:
        yy = yy * np.sqrt(np.clip(np.asarray(dcf_b, dtype=np.float64), 0.0, None))[None, None]
    m = np.asarray(maps, dtype=np.complex128).reshape(C, -1)
    x = np.einsum("lck,kr,cr->lr", yy, Ec, np.conj(m))
    return x.reshape(yy.shape[0], S0, S1)


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.complex128).reshape(-1)
    b = np.asarray(b, dtype=np.complex128).reshape(-1)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
