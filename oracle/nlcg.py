"""ORACLE (test infrastructure only) -- CPU restatement of the reference's NLCG.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module.  It follows the reference solver stack line by line, quirks
included, on any operator exposing ``forward`` / ``adjoint`` over
(B, S0, S1) images and flat k-space vectors:

* Huber TV stencils: graspcg/ops/ndops.py:28-59 (forward difference with the
  last slice zeroed, backward-difference divergence with the first slice kept),
  isotropic / anisotropic energy and flux ndops.py:66-135, accumulation
  ``g += -lambda * div`` and ``E = sum(lambda * e)`` at
  graspcg/regularization/tv_nd.py:54-88;
* objective: residual ``r = t*Ad + Ax - y`` in that order
  (graspcg/ops/objective.py:31-45), f_data = 0.5 ||r||^2 (:229), grad = A^H r
  (:225), regulariser evaluated at x + t d with x mutated and restored
  (:256-266), slope Re<g, d> (:268-277);
* Armijo backtracking from t = 1 with factor 0.5, c1 = 1e-4, at most 20 trials,
  rejecting a non-descent / non-finite slope (graspcg/numerics/line_search.py:24-79);
  or the strong-Wolfe search (doubling from t = 1, bisection zoom, c2 = 0.9;
  line_search.py:89-204) with ``ls_name="wolfe"``;
* CG loop: prime d = -g (graspcg/solvers/cg.py:178-190), accept x += t d
  (:205-207), stats norms (:213-220), the reference convergence test
  (graspcg/solvers/utils.py:40-43), direction update FR with the first beta = 1
  quirk (graspcg/numerics/directions.py:141-162), PR+ (:75-109) or Dai-Yuan
  (:111-139), D = 1 (cg.py:95) or, with ``precond``, the regularisers' diagonal
  majoriser (graspcg/regularization/manager.py:145-218, tv_nd.py:113-215);
* graph-Laplacian term along frames (graspcg/regularization/graph_laplacian.py:22-321).

Pinning (tests/test_oracle_nlcg.py): closed-form quadratic with a unitary
operator (graspcg/tests/test_numerics_ns.py:110-194), rejection of an ascent
direction (:229-248), the Armijo inequality and monotone history
(graspcg/tests/test_cg_descend_compile_ns.py:75-123), the Richardson
finite-difference slope (:180-229), TV of a constant, the fwd/bwd adjoint
identity and the hand-computable step-edge energy; Dai-Yuan by its identity
Re<g_k, d_k> = beta_k Re<g_(k-1), d_(k-1)>, which no other beta rule satisfies
in general; strong Wolfe by the reference's zoom-to-1/2 case
(graspcg/tests/test_numerics_ns.py:195-225), hand-traced steps of the expansion
and of both zoom branches on scaled quadratics, its trial limits, and the two
Wolfe conditions along TV-regularised runs of every direction rule
(graspcg/tests/test_cg_descend_compile_ns.py:125-176); the diagonal majoriser by
hand-evaluated entries (constant image, step edge, axis weight, single-frame
profile) and the preconditioned rules by two properties: a constant D = 2 only
doubles the accepted steps, and DY's identity holds with D from the majoriser;
the graph-Laplacian term by the chain graph's sum of squared frame differences,
finite-difference gradients, the reference's degree-diagonal and FD tests
(graspcg/tests/test_regularization_ns.py:195-275) and the cycle's L_sym = L_rw = L / 2.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np


# --------------------------------------------------------------- TV (ndops.py)

def fwd_diff(x: np.ndarray, ax: int) -> np.ndarray:
    out = np.roll(x, -1, axis=ax) - x
    idx = [slice(None)] * x.ndim
    idx[ax] = -1
    out[tuple(idx)] = 0
    return out


def bwd_diff(p: np.ndarray, ax: int) -> np.ndarray:
    out = p - np.roll(p, 1, axis=ax)
    idx = [slice(None)] * p.ndim
    idx[ax] = 0
    out[tuple(idx)] = p[tuple(idx)]
    return out


@dataclass
class TV:
    """One Huber-TV term: weight lambda, knee eps, axes of the (B, S0, S1) image."""
    weight: float
    axes: Sequence[int]
    eps: float = 1e-3
    isotropic: bool = True
    axis_weights: Optional[Sequence[float]] = None


def _abs2(z):
    return z.real * z.real + z.imag * z.imag


def tv_energy_grad(x: np.ndarray, reg: TV, g: Optional[np.ndarray] = None):
    """Energy of one TV term at x; when ``g`` is given, g += -lambda * div(flux).

    Slices are independent problems only when the caller passes one slice;
    this routine differences across the whole array like the reference."""
    rdt = x.real.dtype
    if reg.weight == 0.0 or len(reg.axes) == 0:
        return rdt.type(0.0)
    ws = reg.axis_weights or [1.0] * len(reg.axes)
    grads = [rdt.type(w) * fwd_diff(x, ax) for w, ax in zip(ws, reg.axes)]
    eps = rdt.type(reg.eps)
    lam = rdt.type(reg.weight)
    if reg.isotropic:
        ss = sum(_abs2(gi) for gi in grads)
        den = np.sqrt(ss + eps * eps)
        e = den - eps
        dcl = np.maximum(den, rdt.type(1e-12))
        flux = [gi / dcl for gi in grads]
    else:
        e = 0
        flux = []
        for gi in grads:
            den = np.sqrt(_abs2(gi) + eps * eps)
            e = e + (den - eps)
            flux.append(gi / np.maximum(den, rdt.type(1e-12)))
    if g is not None:
        div = sum(bwd_diff(pi, ax) for pi, ax in zip(flux, reg.axes))
        g += (-div) * lam
    return (e * lam).sum(dtype=rdt)


@dataclass
class GraphLap:
    """Graph-Laplacian term along the frame axis (graspcg/regularization/graph_laplacian.py:22-321).

    W (F, F) real adjacency; normalize "none" (L = Deg - W), "rw" (I - Deg^-1 W) or "sym"
    (I - Deg^-1/2 W Deg^-1/2)."""
    weight: float
    W: np.ndarray
    normalize: str = "none"

    def laplacian(self) -> np.ndarray:
        """W symmetrised, its diagonal zeroed, degrees = row sums clamped at 0 (:43-78); the
        normalised forms divide by the degree clamped at 1e-20 (:101-106, :140-154)."""
        W = np.asarray(self.W, dtype=np.float64)
        W = 0.5 * (W + W.T)
        np.fill_diagonal(W, 0.0)
        deg = np.maximum(W.sum(axis=1), 0.0)
        n = W.shape[0]
        norm = (self.normalize or "none").lower()
        if norm == "none":
            return np.diag(deg) - W
        if norm == "rw":
            return np.eye(n) - (1.0 / np.maximum(deg, 1e-20))[:, None] * W
        if norm == "sym":
            dmh = 1.0 / np.sqrt(np.maximum(deg, 1e-20))
            return np.eye(n) - dmh[:, None] * W * dmh[None, :]
        raise ValueError(self.normalize)

    def profile(self) -> np.ndarray:
        """majorizer_profile (:283-321): lambda deg (none) or lambda [deg > 0] (rw / sym)."""
        W = np.asarray(self.W, dtype=np.float64)
        W = 0.5 * (W + W.T)
        np.fill_diagonal(W, 0.0)
        deg = np.maximum(W.sum(axis=1), 0.0)
        v = deg if (self.normalize or "none").lower() == "none" else (deg > 0).astype(np.float64)
        return self.weight * v


def gl_energy_grad(x: np.ndarray, reg: GraphLap, g: Optional[np.ndarray] = None):
    """E = (lambda / 2) sum_p Re<x[:, p], (L x)[:, p]> and g += lambda L x over the frame axis
    (graph_laplacian.py:188-242).  For "rw" L is not symmetric and lambda L x is the reference's
    gradient, not the derivative of E."""
    rdt = x.real.dtype
    if reg.weight == 0.0:
        return rdt.type(0.0)
    F = x.shape[0]
    X = x.reshape(F, -1)
    Y = reg.laplacian().astype(rdt) @ X
    lam = rdt.type(reg.weight)
    if g is not None:
        g += (lam * Y).reshape(x.shape)
    return rdt.type(0.5) * lam * rdt.type(np.sum((X.conj() * Y).real))


def reg_energy_grad(x: np.ndarray, reg, g: Optional[np.ndarray] = None):
    return gl_energy_grad(x, reg, g) if isinstance(reg, GraphLap) else tv_energy_grad(x, reg, g)


def _degree(n: int) -> np.ndarray:
    """Ends of an axis count once, interior positions twice (tv_nd.py:150-159, :207-209)."""
    v = np.ones(n)
    if n > 2:
        v[1:-1] = 2.0
    return v


def majoriser_diag(x: np.ndarray, regs: Sequence[TV]) -> np.ndarray:
    """Diagonal preconditioner of graspcg/regularization/manager.py:145-218 for one problem (F, S0, S1).

    Starts from the workspace's ones (graspcg/solvers/cg.py:95).  A term whose axes include the
    frame axis contributes its axis profile only (TVND.majorizer_profile, tv_nd.py:181-215):
    2 lambda w_0^2 times the frame degree.  Any other term adds the lagged-diffusivity weights
    (TVND.add_diag, tv_nd.py:113-162): per axis, lambda w_a^2 / sqrt(|D_a x|^2 + eps^2) with D_a
    the unweighted forward difference (zero on the last slice), times the position's degree."""
    rdt = x.real.dtype
    D = np.ones(x.shape, dtype=rdt)
    for reg in regs:
        if isinstance(reg, GraphLap):       # its frame profile (graph_laplacian.py:283-321)
            if reg.weight != 0.0:
                D += reg.profile().astype(rdt).reshape(-1, 1, 1)
            continue
        if reg.weight == 0.0 or len(reg.axes) == 0:
            continue
        ws = list(reg.axis_weights) if reg.axis_weights is not None else [1.0] * len(reg.axes)
        lam, eps = rdt.type(reg.weight), rdt.type(reg.eps)
        if 0 in reg.axes:
            w0 = ws[list(reg.axes).index(0)]
            D += (rdt.type(2.0 * reg.weight * w0 * w0) * _degree(x.shape[0]).astype(rdt)).reshape(-1, 1, 1)
            continue
        for ax, w in zip(reg.axes, ws):
            gi = fwd_diff(x, ax)
            wi = rdt.type(1.0) / np.sqrt(_abs2(gi) + eps * eps)
            shape = [1] * x.ndim
            shape[ax] = -1
            D += lam * rdt.type(w * w) * wi * _degree(x.shape[ax]).astype(rdt).reshape(shape)
    return D


# --------------------------------------------------------------- objective

class Objective:
    """phi(t) = 0.5 ||A(x + t d) - y||^2 + sum_r TV_r(x + t d), reference evaluation order."""

    def __init__(self, A: Callable, AH: Callable, y: np.ndarray, regs: Sequence[TV],
                 mutate_x: bool = True):
        self.A, self.AH, self.y, self.regs = A, AH, y, list(regs)
        self.mutate_x = mutate_x
        self.Ax = self.Ad = None
        self.n_A = self.n_AH = 0

    def begin_linesearch(self, x, d):
        self.Ax = self.A(x)
        self.Ad = self.A(d)
        self.n_A += 2

    def f_g(self, x: np.ndarray, d: np.ndarray, t: float):
        rdt = x.real.dtype
        tt = rdt.type(t)
        r = self.Ad * tt
        r = r + self.Ax
        r = r - self.y
        f_data = rdt.type(0.5) * (_abs2(r).sum(dtype=rdt))
        g = self.AH(r).astype(x.dtype)
        self.n_AH += 1
        if self.mutate_x:
            x += d * tt
            xe = x
        else:
            xe = x + d * tt
        f_reg = rdt.type(0.0)
        for reg in self.regs:
            f_reg = f_reg + reg_energy_grad(xe, reg, g)
        if self.mutate_x:
            x += d * (-tt)
        gdot = rdt.type(np.vdot(g.reshape(-1), d.reshape(-1)).real)
        return rdt.type(f_data + f_reg), gdot, g


# --------------------------------------------------------------- line search

def armijo(obj: "Objective", x: np.ndarray, d: np.ndarray, c1: float = 1e-4, ls_beta: float = 0.5,
           ls_max_iter: int = 20):
    """Backtracking Armijo from t = 1 (graspcg/numerics/line_search.py:24-79).

    Returns (ok, t, f0, g0d, f_t, gdot_t, g) with g the gradient at the last
    evaluated point; a non-descent or non-finite slope returns ok=False, t=0."""
    rdt = x.real.dtype
    obj.begin_linesearch(x, d)
    f0, g0d, g = obj.f_g(x, d, 0.0)
    if not (g0d < 0) or not np.isfinite(g0d):
        return False, rdt.type(0.0), f0, g0d, f0, g0d, g
    t = rdt.type(1.0)
    ft, gdt = f0, g0d
    for _ in range(ls_max_iter):
        ft, gdt, g = obj.f_g(x, d, t)
        if ft <= f0 + rdt.type(c1) * t * g0d:
            return True, t, f0, g0d, ft, gdt, g
        t = rdt.type(t * ls_beta)
    return False, t, f0, g0d, ft, gdt, g


def wolfe(obj: "Objective", x: np.ndarray, d: np.ndarray, c1: float = 1e-4, c2: float = 0.9,
          ls_max_iter: int = 20, t_max: float = 1e6):
    """Strong-Wolfe search with doubling and bisection zoom (graspcg/numerics/line_search.py:89-204).

    Expansion from t = 1 (:121-152): a step that breaks Armijo, or (after the first trial)
    does not lower f below the previous trial, opens a zoom between the previous trial and it;
    a step meeting |g'(t)| <= c2 |g'(0)| is accepted; a non-negative slope opens a zoom with the
    ends swapped; otherwise t doubles (capped at t_max).  Out of trials: failure (ok False).
    Zoom (:161-204): bisect; Armijo break or f(mid) >= f(lo) moves hi; otherwise accept on the
    curvature test, else hi <- lo when g'(mid) (hi - lo) >= 0, then lo <- mid.  After
    ``ls_max_iter`` bisections the lo end is accepted (its f and g evaluated again, :197-204),
    even at t = 0.  Returns (ok, t, f0, g0d, f_t, gdot_t, g) like ``armijo``."""
    rdt = x.real.dtype
    obj.begin_linesearch(x, d)
    f0, g0d, g = obj.f_g(x, d, 0.0)
    if not (g0d < 0) or not np.isfinite(g0d):
        return False, rdt.type(0.0), f0, g0d, f0, g0d, g
    c1, c2 = rdt.type(c1), rdt.type(c2)

    def zoom(t_lo, f_lo, t_hi):
        for _ in range(ls_max_iter):
            t_mid = rdt.type(0.5) * (t_lo + t_hi)
            f_mid, g_mid, gm = obj.f_g(x, d, t_mid)
            if f_mid > f0 + c1 * t_mid * g0d or f_mid >= f_lo:
                t_hi = t_mid
            else:
                if abs(g_mid) <= c2 * abs(g0d):
                    return True, t_mid, f0, g0d, f_mid, g_mid, gm
                if g_mid * (t_hi - t_lo) >= 0:
                    t_hi = t_lo
                t_lo, f_lo = t_mid, f_mid
        f_lo, g_lo, gl = obj.f_g(x, d, t_lo)
        return True, t_lo, f0, g0d, f_lo, g_lo, gl

    t = rdt.type(1.0)
    t_prev, f_prev, g_prev = rdt.type(0.0), f0, g0d
    for i in range(ls_max_iter):
        ft, gdt, g = obj.f_g(x, d, t)
        if ft > f0 + c1 * t * g0d or (i > 0 and ft >= f_prev):
            return zoom(t_prev, f_prev, t)
        if abs(gdt) <= c2 * abs(g0d):
            return True, t, f0, g0d, ft, gdt, g
        if gdt >= 0:
            return zoom(t, ft, t_prev)
        t_prev, f_prev, g_prev = t, ft, gdt
        t = rdt.type(min(2.0 * t, t_max))
    return False, t_prev, f0, g0d, f_prev, g_prev, g


# --------------------------------------------------------------- solver

@dataclass
class CGResult:
    x: np.ndarray
    history: List[dict] = field(default_factory=list)
    f_final: float = float("nan")
    n_iter: int = 0
    stop: str = ""
    n_A: int = 0
    n_AH: int = 0


def _norm(v, rdt):
    return rdt.type(math.sqrt(max(float(np.vdot(v.reshape(-1), v.reshape(-1)).real), 0.0)))


def nlcg(A: Callable, AH: Callable, y: np.ndarray, x0: np.ndarray, regs: Sequence[TV],
         max_iter: int = 50, direction: str = "fr", c1: float = 1e-4, ls_beta: float = 0.5,
         ls_max_iter: int = 20, tol_g: float = 1e-4, tol_step: float = 1e-9,
         rel_tol: Optional[float] = None, mutate_x: bool = True,
         restart_c: Optional[float] = None, ls_name: str = "armijo", c2: float = 0.9,
         precond: bool = False) -> CGResult:
    """Reference-semantics NLCG (graspcg/solvers/cg.py:162-246) with Armijo search, or the
    strong-Wolfe search with ``ls_name="wolfe"`` (line_search.py:12-17, CGConfig.ls_name / c2
    at cg.py:37-39).

    ``rel_tol`` adds the stop rule |f_k - f_{k-1}| / |f_{k-1}| < rel_tol on the
    accepted objective values (not a reference rule; used for time-to-tolerance).
    ``restart_c`` (not a reference rule) sets beta = 0 whenever the new direction
    -g + beta d would have Re<g, d_new> > -restart_c |g|^2, so every searched
    direction descends; None keeps the reference behaviour (such a direction
    ends the run with "not_descent").  ``precond`` builds the diagonal majoriser D of the
    regularisers at x0 (``majoriser_diag``, RegManager.add_diag) before the prime step and runs
    the Jacobi-preconditioned rules of graspcg/numerics/directions.py: d = beta d - g / D, the
    beta numerators / FR and PR+ denominators and the reported gradient norm weighted by 1 / D
    (ops/dot.py:5-71 with diag, solvers/utils.py:29-37); D stays fixed for the run."""
    x = np.array(x0, copy=True)
    rdt = x.real.dtype
    obj = Objective(A, AH, y, regs, mutate_x=mutate_x)
    d = np.zeros_like(x)
    res = CGResult(x=x)
    Dinv = (rdt.type(1.0) / majoriser_diag(x, regs)) if precond else None
    pdot = ((lambda a, b: rdt.type(np.sum((a.conj() * (b * Dinv)).real))) if precond else
            (lambda a, b: rdt.type(np.vdot(a.reshape(-1), b.reshape(-1)).real)))
    # prime (cg.py:178-190): d = -g / D
    obj.begin_linesearch(x, d)
    _, _, g = obj.f_g(x, d, 0.0)
    d = -g * Dinv if precond else -g
    rho_prev = None
    g_prev = g.copy()
    f_prev = None
    for k in range(max_iter):
        if ls_name == "wolfe":
            ok, t, f0, g0d, ft, gdt, g_new = wolfe(obj, x, d, c1=c1, c2=c2, ls_max_iter=ls_max_iter)
        elif ls_name == "armijo":
            ok, t, f0, g0d, ft, gdt, g_new = armijo(obj, x, d, c1=c1, ls_beta=ls_beta, ls_max_iter=ls_max_iter)
        else:
            raise ValueError(ls_name)
        if g_new is not None:
            g = g_new
        if not ok:
            res.stop = "not_descent" if t == 0 and not (g0d < 0 and np.isfinite(g0d)) else "ls_failed"
            break
        x += d * t
        xn, sn = _norm(x, rdt), _norm(d, rdt)
        gn = rdt.type(math.sqrt(max(float(pdot(g, g)), 0.0)))
        res.history.append(dict(iter=k, f=float(f0), gdot=float(g0d), gdot_at_step=float(gdt),
                                step=float(t), f_step=float(ft), xnorm=float(xn),
                                gnorm=float(gn), stepnorm=float(sn)))
        res.n_iter = k + 1
        if gn <= (1 + xn) * tol_g and sn <= (1 + xn) * tol_step:
            res.stop = "converged"
            break
        if rel_tol is not None and f_prev is not None and abs(ft - f_prev) <= rel_tol * abs(f_prev):
            f_prev = ft
            res.stop = "rel_tol"
            break
        f_prev = ft
        # direction update (directions.py)
        rho = pdot(g, g)
        if direction == "fr":
            if rho_prev is None:
                rho_prev = rho
            beta = rho / max(rho_prev, rdt.type(1e-20))
            rho_prev = rho
        elif direction == "prplus":
            num = rho - pdot(g_prev, g)
            den = pdot(g_prev, g_prev)
            beta = max(num / max(den, rdt.type(1e-20)), rdt.type(0.0))
            g_prev = g.copy()
        elif direction == "dy":
            # beta_DY = <g, g> / Re<d, g - g_prev> (directions.py:120-130)
            den = (rdt.type(np.vdot(d.reshape(-1), g.reshape(-1)).real)
                   - rdt.type(np.vdot(d.reshape(-1), g_prev.reshape(-1)).real))
            beta = rho / max(den, rdt.type(1e-20))
            g_prev = g.copy()
        else:
            raise ValueError(direction)
        if restart_c is not None:
            gd_old = rdt.type(np.vdot(g.reshape(-1), d.reshape(-1)).real)
            if -rho + beta * gd_old > -restart_c * rho:
                beta = rdt.type(0.0)
        res.history[-1]["beta"] = float(beta)
        d = d * rdt.type(beta) - (g * Dinv if precond else g)
    if not res.stop:
        res.stop = "max_iter"
    res.x = x
    res.f_final = float(objective_value(A, y, x, regs))
    res.n_A, res.n_AH = obj.n_A, obj.n_AH
    return res


def objective_value(A: Callable, y: np.ndarray, x: np.ndarray, regs: Sequence[TV]) -> float:
    """f(x) evaluated in float64 from the given iterate."""
    r = np.asarray(A(x), dtype=np.complex128) - np.asarray(y, dtype=np.complex128)
    f = 0.5 * float(np.sum(_abs2(r)))
    x64 = np.asarray(x, dtype=np.complex128)
    for reg in regs:
        f += float(reg_energy_grad(x64, reg))
    return f
