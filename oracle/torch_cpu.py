"""ORACLE (test infrastructure only) -- the pure-torch CPU port timed as the CPU baseline.

Only tests/ and bench.py's cpu_baseline leg may use this module.  It restates the same KB
gridding operator as oracle/kbnufft.py in torch CPU ops (complex64, ``torch.fft``, gathers
and ``index_add_``), the way the reference's default backend runs on a CPU
(graspcg/nufft/backends/torchkb.py:203-323: maps applied in A, balanced sqrt(dcf) on both
sides, coil combine with conj(maps) in A^H), and the reference NLCG loop in torch with FP32
reductions as the reference has them (graspcg/solvers/cg.py:162-246,
graspcg/numerics/line_search.py:24-79, graspcg/numerics/directions.py:75-162,
graspcg/regularization/tv_nd.py:54-88 with the stencils of graspcg/ops/ndops.py:28-149).

Pinned (tests/test_oracle_torch_cpu.py) against the FP64 numpy oracle, which the NUDFT and the
reference's closed forms pin: operator rel L2 <= 1e-5 on ragged frames, TV energy and gradient,
and the NLCG step sequence.
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence

import numpy as np
import torch

from .kbnufft import deapod_1d, footprint, kb_beta


class TorchKBNufft:
    """KB-gridding SENSE operator on (B, S0, S1) images and flat (sum_b C K_b) k-space, torch CPU."""

    def __init__(self, maps: np.ndarray, traj: np.ndarray, k_offsets: np.ndarray, dcf: Optional[np.ndarray],
                 W: int, osf: float = 2.0, dcf_mode: str = "balanced"):
        m = torch.from_numpy(np.ascontiguousarray(maps).astype(np.complex64))
        self.C, self.S0, self.S1 = (int(v) for v in m.shape)
        self.maps = m
        self.G0, self.G1 = int(math.ceil(osf * self.S0)), int(math.ceil(osf * self.S1))
        self.W = int(W)
        beta = kb_beta(W, osf)
        self.off = np.asarray(k_offsets, dtype=np.int64)
        self.B = len(self.off) - 1
        d0 = deapod_1d(self.S0, self.G0, self.W, beta)
        d1 = deapod_1d(self.S1, self.G1, self.W, beta)
        self.deapod = torch.from_numpy((d0[:, None] * d1[None, :]).astype(np.float32))
        traj = np.asarray(traj, dtype=np.float64)
        self.idx: List[torch.Tensor] = []
        self.wgt: List[torch.Tensor] = []
        self.wA: List[Optional[torch.Tensor]] = []
        self.wAH: List[Optional[torch.Tensor]] = []
        for b in range(self.B):
            s = slice(self.off[b], self.off[b + 1])
            u = traj[0, s] * self.G0 / (2.0 * math.pi)
            v = traj[1, s] * self.G1 / (2.0 * math.pi)
            mx, wx = footprint(u, self.W, beta)
            my, wy = footprint(v, self.W, beta)
            ix = (mx[:, None] + np.arange(self.W)[None, :]) % self.G0
            iy = (my[:, None] + np.arange(self.W)[None, :]) % self.G1
            flat = (ix[:, :, None] * self.G1 + iy[:, None, :]).reshape(-1)
            self.idx.append(torch.from_numpy(flat.astype(np.int64)))
            self.wgt.append(torch.from_numpy((wx[:, :, None] * wy[:, None, :]).reshape(len(u), -1).astype(np.float32)))
            if dcf is None or dcf_mode == "none":
                self.wA.append(None); self.wAH.append(None)
            elif dcf_mode == "balanced":
                w = torch.from_numpy(np.sqrt(np.clip(np.asarray(dcf, dtype=np.float64)[s], 0, None)).astype(np.float32))
                self.wA.append(w); self.wAH.append(w)
            else:
                raise ValueError(dcf_mode)

    @property
    def n_k(self) -> int:
        return int(self.off[-1]) * self.C

    def _pad(self, img: torch.Tensor) -> torch.Tensor:
        """(..., S0, S1) -> (..., G0, G1) with image index i at grid index (i - S//2) mod G."""
        g = torch.zeros(img.shape[:-2] + (self.G0, self.G1), dtype=torch.complex64)
        g[..., :self.S0, :self.S1] = img
        return torch.roll(g, shifts=(-(self.S0 // 2), -(self.S1 // 2)), dims=(-2, -1))

    def _crop(self, g: torch.Tensor) -> torch.Tensor:
        g = torch.roll(g, shifts=(self.S0 // 2, self.S1 // 2), dims=(-2, -1))
        return g[..., :self.S0, :self.S1]

    @torch.no_grad()
    def forward(self, x: torch.Tensor) -> torch.Tensor:
        x = x.reshape(self.B, self.S0, self.S1)
        grid = torch.fft.fft2(self._pad(x[:, None] * self.maps[None] * self.deapod)).reshape(self.B, self.C, -1)
        out = []
        for b in range(self.B):
            K = int(self.off[b + 1] - self.off[b])
            vals = (grid[b][:, self.idx[b]].reshape(self.C, K, -1) * self.wgt[b][None]).sum(-1)
            if self.wA[b] is not None:
                vals = vals * self.wA[b][None]
            out.append(vals.reshape(-1))
        return torch.cat(out)

    @torch.no_grad()
    def adjoint(self, y: torch.Tensor) -> torch.Tensor:
        y = y.reshape(-1)
        G = self.G0 * self.G1
        grid = torch.zeros(self.B, self.C, G, dtype=torch.complex64)
        for b in range(self.B):
            K = int(self.off[b + 1] - self.off[b])
            yk = y[self.C * int(self.off[b]):self.C * int(self.off[b + 1])].reshape(self.C, K)
            if self.wAH[b] is not None:
                yk = yk * self.wAH[b][None]
            contrib = (yk[:, :, None] * self.wgt[b][None]).reshape(self.C, -1)
            grid[b].index_add_(1, self.idx[b], contrib)
        img = self._crop(torch.fft.ifft2(grid.reshape(self.B, self.C, self.G0, self.G1)) * G)
        return (img * self.maps.conj()[None]).sum(1) * self.deapod


# --------------------------------------------------------------- TV and NLCG (torch, FP32)

def _fwd_diff(x: torch.Tensor, ax: int) -> torch.Tensor:
    out = torch.roll(x, -1, dims=ax) - x
    out.select(ax, x.shape[ax] - 1).zero_()
    return out


def _bwd_diff(p: torch.Tensor, ax: int) -> torch.Tensor:
    out = p - torch.roll(p, 1, dims=ax)
    out.select(ax, 0).copy_(p.select(ax, 0))
    return out


def tv_energy_grad(x: torch.Tensor, weight: float, axes: Sequence[int], eps: float = 1e-3,
                   g: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Isotropic Huber TV: E = lambda sum(sqrt(sum_a |D_a x|^2 + eps^2) - eps); g += -lambda div(flux)."""
    grads = [_fwd_diff(x, ax) for ax in axes]
    ss = sum(gi.real * gi.real + gi.imag * gi.imag for gi in grads)
    den = torch.sqrt(ss + eps * eps)
    if g is not None:
        dcl = torch.clamp(den, min=1e-12)
        div = sum(_bwd_diff(gi / dcl, ax) for gi, ax in zip(grads, axes))
        g.add_(div, alpha=-weight)
    return ((den - eps) * weight).sum()


def nlcg(A, AH, y: torch.Tensor, x0: torch.Tensor, regs: Sequence[tuple], max_iter: int = 50,
         direction: str = "prplus", c1: float = 1e-4, ls_beta: float = 0.5, ls_max_iter: int = 20,
         rel_tol: Optional[float] = None, restart_c: Optional[float] = None):
    """Reference-loop NLCG in torch: every Armijo trial evaluates f and the gradient at x + t d
    (one A^H per trial, two A per iteration); ``regs`` = [(weight, axes), ...].  Returns
    (x, history of dicts with f, step, f_step)."""
    x = x0.clone()
    hist = []

    def fg(Ax, Ad, d, t):
        r = Ad * t + Ax - y
        f = 0.5 * (r.real * r.real + r.imag * r.imag).sum()
        g = AH(r)
        xe = x + d * t
        for w, axes in regs:
            f = f + tv_energy_grad(xe, w, axes, g=g)
        return f, torch.vdot(g.reshape(-1), d.reshape(-1)).real, g

    d = torch.zeros_like(x)
    _, _, g = fg(A(x), A(d), d, 0.0)
    d = -g
    g_prev, rho_prev, f_prev = g.clone(), None, None
    for k in range(max_iter):
        Ax, Ad = A(x), A(d)
        f0, g0d, g = fg(Ax, Ad, d, 0.0)
        if not bool(g0d < 0) or not bool(torch.isfinite(g0d)):
            break
        t, ok = 1.0, False
        for _ in range(ls_max_iter):
            ft, _, g = fg(Ax, Ad, d, t)
            if bool(ft <= f0 + c1 * t * g0d):
                ok = True
                break
            t *= ls_beta
        if not ok:
            break
        x = x + d * t
        hist.append(dict(iter=k, f=float(f0), step=t, f_step=float(ft)))
        if rel_tol is not None and f_prev is not None and abs(float(ft) - f_prev) <= rel_tol * abs(f_prev):
            break
        f_prev = float(ft)
        rho = torch.vdot(g.reshape(-1), g.reshape(-1)).real
        if direction == "fr":
            rho_prev = rho if rho_prev is None else rho_prev
            beta = rho / torch.clamp(rho_prev, min=1e-20)
            rho_prev = rho
        else:
            num = rho - torch.vdot(g_prev.reshape(-1), g.reshape(-1)).real
            beta = torch.clamp(num / torch.clamp(torch.vdot(g_prev.reshape(-1), g_prev.reshape(-1)).real, min=1e-20),
                               min=0.0)
            g_prev = g.clone()
        if restart_c is not None and bool(-rho + beta * torch.vdot(g.reshape(-1), d.reshape(-1)).real > -restart_c * rho):
            beta = torch.zeros(())
        d = d * beta - g
    return x, hist
