"""Non-linear CG on the device: batched engine and the reference's CGSolver surface.

``NLCGEngine`` runs P independent problems (slices) of F frames each with
ragged per-frame k-space in one set of launches.  One iteration is

    line search  : a = |Ax-y|^2, b = Re<Ax-y, Ad>, c = |Ad|^2 and TV(x + t d) for
                   every candidate t in {1, 1/2, ..., 2^-19} in one pass, first t
                   with f(t) <= f(0) + c1 t Re<g, d> selected on the device
                   (``ls_name="wolfe"``: the strong-Wolfe search, TV values and
                   slopes at seven steps per pass, the sequential search replayed
                   over them on the device);
    accept       : x += t d, Ax += t Ad, r = Ax - y;
    gradient     : g = A^H r + grad TV(x), |g|^2, beta (FR, first beta = 1; or PR+);
    direction    : d = beta d - g;  Ad = A d.

i.e. 1 forward + 1 adjoint NUFFT per iteration, against 2 A + (1 + trials) A^H
for the reference loop (graspcg/solvers/cg.py:193-244 driving
graspcg/numerics/line_search.py:24-79 and graspcg/ops/objective.py:114-291).
The data term along d is exactly quadratic, so the accepted step equals the
sequential backtracking result; Ax is carried forward instead of recomputed
(identical in exact arithmetic).  The iteration has no host synchronisation and
is captured once in a HIP graph and replayed.

``CGSolver`` / ``CGConfig`` keep the reference names (cg.py:30-288):
``CGSolver(y, nufft_op, regm, cfg).run(max_iter)``, ``result()``, ``history()``.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .regularization import RegManager

_STATUS = {0: "active", 1: "ls_failed", 2: "not_descent", 3: "rel_tol", 4: "target"}


# direction rules of the reference (numerics/directions.py): FR (:141-162, first beta = 1),
# PR+ (:75-109) and Dai-Yuan (:111-139)
_DIRECTIONS = {"fr": _lib.CG_DIR_FR, "prplus": _lib.CG_DIR_PRPLUS, "dy": _lib.CG_DIR_DY}
# line searches (numerics/line_search.py): backtracking Armijo (:24-79), strong Wolfe (:89-204)
_SEARCHES = {"armijo": _lib.CG_LS_ARMIJO, "wolfe": _lib.CG_LS_WOLFE}


class NLCGEngine:
    """Device NLCG for P problems x F frames (image (P*F, S0, S1), flat ragged k-space)."""

    def __init__(self, plan, y: torch.Tensor, *, problems: int, frames: int, regm: Optional[RegManager],
                 direction: str = "fr", c1: float = 1e-4, ls_beta: float = 0.5, ls_max_iter: int = 20,
                 rel_tol: float = 0.0, max_hist: int = 256, restart_c: Optional[float] = None,
                 skip_stopped: bool = True, ls_name: str = "armijo", c2: float = 0.9,
                 precond: bool = False, iters_per_replay: int = 4, fp32_reductions: bool = False):
        if direction not in _DIRECTIONS:
            raise ValueError(f"direction must be one of {sorted(_DIRECTIONS)}")
        if ls_name not in _SEARCHES:
            raise ValueError(f"ls_name must be one of {sorted(_SEARCHES)}")
        self.plan = plan
        self.P, self.F = int(problems), int(frames)
        if self.P * self.F != plan.B:
            raise ValueError(f"problems*frames = {self.P * self.F} != trajectory frames {plan.B}")
        self.S0, self.S1 = plan.S
        self.C = plan.C
        dev = y.device
        self.device = dev
        if y.dtype != torch.complex64:
            y = y.to(torch.complex64)
        self.y = y.reshape(-1).contiguous()
        if self.y.numel() != plan.n_k:
            raise ValueError(f"y has {self.y.numel()} samples, plan expects {plan.n_k}")
        self.direction = direction
        shape = (self.P * self.F, self.S0, self.S1)
        z = lambda s: torch.zeros(s, dtype=torch.complex64, device=dev)
        self.x, self.d, self.g = z(shape), z(shape), z(shape)
        self.g_prev = z(shape) if direction == "prplus" else None
        self.Ax, self.Ad, self.r = z(self.y.shape), z(self.y.shape), z(self.y.shape)
        koff = plan.k_offsets
        self.k_prob = np.ascontiguousarray(
            [self.C * int(koff[p * self.F]) for p in range(self.P + 1)], dtype=np.int64)
        desc = _lib.CGDesc()
        desc.tv = (regm or RegManager([])).tv_desc(self.P, self.F, self.S0, self.S1)
        desc.k_prob = self.k_prob.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        desc.direction = _DIRECTIONS[direction]
        desc.c1, desc.ls_beta, desc.ls_max_iter = float(c1), float(ls_beta), int(ls_max_iter)
        desc.rel_tol, desc.max_hist = float(rel_tol or 0.0), int(max_hist)
        desc.restart_c = -1.0 if restart_c is None else float(restart_c)
        desc.ls_kind, desc.c2 = _SEARCHES[ls_name], float(c2)
        # all-FP32 inner products / beta / line-search values, as the reference computes them
        # (cg.py:89-90, ops/dot.py:54-71): a precision-study variant of the FP64 default
        desc.fp32_reductions = 1 if fp32_reductions else 0
        self.ls_name = ls_name
        # diagonal preconditioner D (the reference workspace's "diag"): the regularisers'
        # majoriser at x0, built by prime() (RegManager.add_diag); None: D = 1
        self.diag = torch.ones(shape, dtype=torch.float32, device=dev) if precond else None
        desc.diag = self.diag.data_ptr() if precond else None
        gl = (regm or RegManager([])).graph()
        if gl is not None:
            if gl.frames != self.F:
                raise ValueError(f"graph term has {gl.frames} frames, the problems have {self.F}")
            self._gl_L = np.ascontiguousarray(gl.laplacian(), dtype=np.float32)
            self._gl_prof = np.ascontiguousarray(gl.profile(), dtype=np.float32)
            desc.gl_weight, desc.gl_F = float(gl.params.weight), int(gl.frames)
            desc.gl_L = self._gl_L.ctypes.data
            desc.gl_prof = self._gl_prof.ctypes.data
        if restart_c is not None and restart_c < 0:
            raise ValueError("restart_c must be >= 0 (None keeps the reference behaviour)")
        self.max_hist = int(max_hist)
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        with torch.cuda.device(dev):
            _lib.check(self._lib.mrh_cg_create(ctypes.byref(desc), ctypes.byref(h)))
        self.handle = h
        flags = ctypes.c_void_p()
        _lib.check(self._lib.mrh_cg_active_flags(h, ctypes.byref(flags)))
        self._active = flags.value
        # the iteration's NUFFTs skip the frames of problems that have stopped
        self.skip_stopped = bool(skip_stopped)
        if int(iters_per_replay) < 1:
            raise ValueError("iters_per_replay must be >= 1")
        # iterations captured in one graph replay; the host reads the problems' status once per
        # replay.  A problem that stops inside a replay is skipped by every later kernel of it
        # (status check in the solver kernels, masked NUFFTs), so the results equal one replay
        # per iteration; without skip_stopped the NUFFTs still run for stopped problems.
        self.iters_per_replay = int(iters_per_replay)
        self._graph = None
        self._graph_k = None
        self.n_steps = 0

    # ------------------------------------------------------------------ steps
    def _st(self) -> int:
        return _lib.stream_of(self.x)

    def prime(self, x0: Optional[torch.Tensor] = None, Ax0: Optional[torch.Tensor] = None) -> None:
        """Start a run at x0 (or the current x): g = grad f(x0), d = -g / D, Ad = A d.
        ``Ax0`` (= A x0, already computed by the caller) saves the forward NUFFT.  With a
        preconditioner, D is rebuilt here from x0 and kept for the run."""
        lib, e = self._lib, self.handle
        if x0 is not None:
            self.x.copy_(x0.reshape(self.x.shape))
        _lib.check(lib.mrh_cg_reset(e, self._st()))
        if self.diag is not None:
            _lib.check(lib.mrh_cg_build_diag(e, self.x.data_ptr(), self._st()))
        if Ax0 is not None:
            self.Ax.copy_(Ax0.reshape(self.Ax.shape))
        else:
            self.plan.forward(self.x, self.Ax)
        _lib.check(lib.mrh_cg_residual(e, self.Ax.data_ptr(), self.y.data_ptr(), self.r.data_ptr(), self._st()))
        self.plan.adjoint(self.r, self.g)
        _lib.check(lib.mrh_cg_gradient(e, self.x.data_ptr(), self.g.data_ptr(), 0, _lib.ptr(self.g_prev), 1,
                                       self._st()))
        _lib.check(lib.mrh_cg_direction(e, self.d.data_ptr(), self.g.data_ptr(), _lib.ptr(self.g_prev), 1,
                                        self._st()))
        self.plan.forward(self.d, self.Ad)
        self.n_steps = 0

    def step(self) -> None:
        """One NLCG iteration for every active problem (no host synchronisation)."""
        lib, e, s = self._lib, self.handle, self._st()
        # r = Ax - y is current (prime and every accept leave it), so the line search reads r alone
        _lib.check(lib.mrh_cg_linesearch(e, self.x.data_ptr(), self.d.data_ptr(), self.g.data_ptr(),
                                         self.r.data_ptr(), self.Ad.data_ptr(), None, s))
        _lib.check(lib.mrh_cg_accept(e, self.x.data_ptr(), self.d.data_ptr(), self.Ax.data_ptr(),
                                     self.Ad.data_ptr(), self.y.data_ptr(), self.r.data_ptr(), s))
        if self.skip_stopped:
            self.plan.adjoint_masked(self.r, self.g, self._active, self.F)
        else:
            self.plan.adjoint(self.r, self.g)
        _lib.check(lib.mrh_cg_gradient(e, self.x.data_ptr(), self.g.data_ptr(), self.d.data_ptr(),
                                       _lib.ptr(self.g_prev), 0, s))
        _lib.check(lib.mrh_cg_direction(e, self.d.data_ptr(), self.g.data_ptr(), _lib.ptr(self.g_prev), 0, s))
        if self.skip_stopped:
            self.plan.forward_masked(self.d, self.Ad, self._active, self.F)
        else:
            self.plan.forward(self.d, self.Ad)

    def capture(self) -> None:
        """Capture :meth:`step` in a HIP graph, and ``iters_per_replay`` steps in a second one
        (replayed by :meth:`run`)."""
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.step()
        self._graph = g
        self._graph_k = None
        if self.iters_per_replay > 1:
            gk = torch.cuda.CUDAGraph()
            with torch.cuda.graph(gk):
                for _ in range(self.iters_per_replay):
                    self.step()
            self._graph_k = gk

    def run(self, max_iter: int, *, use_graph: bool = True, check_every: int = 1) -> int:
        """Iterate until every problem stopped or ``max_iter`` steps; returns the steps launched.
        With the graph, ``iters_per_replay`` steps go per replay while at least that many remain.
        The status is read once ``check_every`` steps have passed since the last read (so at most
        once per replay; 0 = never)."""
        if use_graph and self._graph is None:
            self.capture()
        k = 0
        last_check = 0
        while k < max_iter:
            if use_graph and self._graph_k is not None and max_iter - k >= self.iters_per_replay:
                self._graph_k.replay()
                n = self.iters_per_replay
            elif use_graph:
                self._graph.replay()
                n = 1
            else:
                self.step()
                n = 1
            k += n
            self.n_steps += n
            if check_every and (k - last_check >= check_every or k == max_iter):
                last_check = k
                if not self.active().any():
                    break
        return k

    def set_targets(self, targets: Optional[Sequence[float]]) -> None:
        """Stop problem p once an accepted step reaches f <= targets[p] (None clears)."""
        if targets is None:
            _lib.check(self._lib.mrh_cg_set_targets(self.handle, None, self._st()))
            return
        t = np.ascontiguousarray(np.asarray(targets, dtype=np.float64).reshape(-1))
        if t.size != self.P:
            raise ValueError(f"need {self.P} targets, got {t.size}")
        _lib.check(self._lib.mrh_cg_set_targets(self.handle, t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                self._st()))

    # ------------------------------------------------------------------ readout
    def status(self):
        st = (ctypes.c_int * self.P)()
        it = (ctypes.c_int * self.P)()
        _lib.check(self._lib.mrh_cg_read_status(self.handle, st, it, self._st()))
        return np.array(st[:], dtype=np.int32), np.array(it[:], dtype=np.int32)

    def latest(self):
        """Status, iteration counts and each problem's latest history record (gnorm, xnorm,
        stepnorm) from one copy and one stream synchronisation."""
        st = (ctypes.c_int * self.P)()
        it = (ctypes.c_int * self.P)()
        last = (_lib.CGHist * self.P)()
        _lib.check(self._lib.mrh_cg_read_latest(self.handle, st, it, last, self._st()))
        rows = [dict(gnorm=math.sqrt(max(h.rho, 0.0)), xnorm=math.sqrt(max(h.xnorm2, 0.0)),
                     stepnorm=math.sqrt(max(h.dnorm2, 0.0)), f=h.f0, f_step=h.ft, step=h.t) for h in last]
        return np.array(st[:], dtype=np.int32), np.array(it[:], dtype=np.int32), rows

    def active(self) -> np.ndarray:
        return self.status()[0] == _lib.CG_ACTIVE

    def history(self) -> List[List[Dict[str, float]]]:
        """Per problem, the accepted-iteration records (reference history keys plus extras)."""
        status, iters = self.status()
        n = int(min(self.max_hist, iters.max() if len(iters) else 0))
        buf = (_lib.CGHist * max(1, n * self.P))()
        _lib.check(self._lib.mrh_cg_read_history(self.handle, buf, n, self._st()))
        out: List[List[Dict[str, float]]] = []
        for p in range(self.P):
            rows = []
            for k in range(min(n, int(iters[p]))):
                h = buf[k * self.P + p]
                rows.append(dict(iter=k, f=h.f0, gdot=h.g0d, gdot_at_step=h.gdot_step, step=h.t,
                                 f_step=h.ft, xnorm=math.sqrt(max(h.xnorm2, 0.0)),
                                 gnorm=math.sqrt(max(h.rho, 0.0)), stepnorm=math.sqrt(max(h.dnorm2, 0.0)),
                                 beta=h.beta, f_reg=h.freg))
            out.append(rows)
        return out

    def stop_reasons(self) -> List[str]:
        return [_STATUS.get(int(s), str(s)) for s in self.status()[0]]

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            self._lib.mrh_cg_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------- reference surface

@dataclass
class CGConfig:
    """graspcg/solvers/cg.py:30-52, plus ``rel_tol`` (relative objective-change stop rule),
    ``restart_c`` (beta = 0 restart when the next direction would not descend; None = reference),
    ``precond`` (the regularisers' diagonal majoriser at x0 as the preconditioner D of the direction
    rules, built before the prime step; False = the reference's D = 1),
    ``use_graph`` (HIP-graph replay of the iteration) and ``problems`` (independent slices)."""
    devices: Optional[object] = None
    direction: str = "fr"
    ls_name: str = "armijo"
    c1: float = 1e-4
    c2: float = 0.9
    ls_beta: float = 0.5
    ls_max_iter: int = 20
    ls_zoom: bool = True
    max_iter: int = 50
    tol_g: float = 1e-4
    tol_step: float = 1e-9
    init_mode: str = "keep"
    backproj_xfactor: float = 1.0
    record_history: bool = True
    verbose: bool = False
    rel_tol: Optional[float] = None
    restart_c: Optional[float] = None
    precond: bool = False
    use_graph: bool = True
    problems: int = 1


class Workspace:
    """Named buffers with the reference's canonical names and aliases
    (graspcg/compat/names.py:13-27): var/x, grad/g, dir/dx, fwd_var_sh, fwd_dir_sh, resid_sh."""

    _ALIAS = {"x": "var", "g": "grad", "dx": "dir", "Ax_sh": "fwd_var_sh", "Ad_sh": "fwd_dir_sh",
              "r_sh": "resid_sh"}

    def __init__(self, eng: NLCGEngine):
        self._b = {"var": eng.x, "grad": eng.g, "dir": eng.d, "fwd_var_sh": eng.Ax,
                   "fwd_dir_sh": eng.Ad, "resid_sh": eng.r}
        if eng.g_prev is not None:
            self._b["g_prev"] = eng.g_prev
        if eng.diag is not None:
            self._b["diag"] = eng.diag

    def has(self, name: str) -> bool:
        return self._ALIAS.get(name, name) in self._b

    def list_bufs(self):
        return list(self._b)

    def get(self, name: str, i: int = 0) -> torch.Tensor:
        if i != 0:
            raise IndexError("single-shard workspace")
        return self._b[self._ALIAS.get(name, name)]

    def bind(self, i: int, *names: str):
        return tuple(self.get(n, i) for n in names)

    def iter_shards(self):
        yield self, 0


class CGSolver:
    """Reference-shaped NLCG driver (graspcg/solvers/cg.py:56-288) on the device engine."""

    def __init__(self, y: torch.Tensor, nufft_op, regm: RegManager, cfg: Optional[CGConfig] = None):
        self.cfg = cfg or CGConfig()
        self.y, self.nufft_op, self.regm = y, nufft_op, regm
        be = getattr(nufft_op, "backend", nufft_op)
        plan = be.plan
        P = int(self.cfg.problems)
        self.engine = NLCGEngine(plan, y, problems=P, frames=plan.B // P, regm=regm,
                                 direction=self.cfg.direction.lower(), c1=self.cfg.c1,
                                 ls_beta=self.cfg.ls_beta, ls_max_iter=self.cfg.ls_max_iter,
                                 rel_tol=self.cfg.rel_tol or 0.0, max_hist=max(64, self.cfg.max_iter),
                                 restart_c=self.cfg.restart_c, ls_name=self.cfg.ls_name.lower(),
                                 c2=self.cfg.c2, precond=self.cfg.precond)
        self.ws = Workspace(self.engine)
        if self.cfg.init_mode == "zero":
            self.engine.x.zero_()
        elif self.cfg.init_mode == "backproj":
            plan.adjoint(self.engine.y, self.engine.x)
            if self.cfg.backproj_xfactor != 1.0:
                self.engine.x.mul_(self.cfg.backproj_xfactor)
        elif self.cfg.init_mode != "keep":
            raise ValueError(f"unknown init_mode {self.cfg.init_mode!r}")
        self.iter = 0

    @torch.no_grad()
    def run(self, max_iter: Optional[int] = None, tol_g: Optional[float] = None,
            tol_step: Optional[float] = None) -> "CGSolver":
        max_iter = int(max_iter or self.cfg.max_iter)
        tol_g = float(self.cfg.tol_g if tol_g is None else tol_g)
        tol_step = float(self.cfg.tol_step if tol_step is None else tol_step)
        eng = self.engine
        eng.prime()
        for k in range(max_iter):
            eng.run(1, use_graph=self.cfg.use_graph, check_every=0)
            self.iter = k
            # one read per iteration: status plus each problem's latest record
            status, iters, last = eng.latest()
            if not (status == _lib.CG_ACTIVE).any():
                break
            # reference convergence test (graspcg/solvers/utils.py:40-43)
            if all(n > 0 and r["gnorm"] <= (1 + r["xnorm"]) * tol_g and r["stepnorm"] <= (1 + r["xnorm"]) * tol_step
                   for n, r in zip(iters, last)):
                break
        return self

    @torch.no_grad()
    def result(self) -> torch.Tensor:
        e = self.engine
        return e.x.reshape(-1, 1, e.S0, e.S1).detach()

    @torch.no_grad()
    def history(self) -> List[Dict[str, float]]:
        h = self.engine.history()
        return h[0] if len(h) == 1 else [r for rows in h for r in rows]

