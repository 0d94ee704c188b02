"""Parity of the device objective / TV / NLCG kernels with the CPU oracle.

NLCG bar: objective after a fixed iteration count within 1e-4 relative of the
reference-semantics oracle run (oracle/nlcg.py, FP64) on the same data.
"""
import ctypes

import numpy as np
import pytest
import torch

from mrirecon_amd import (CGConfig, CGSolver, GraphLaplacian, GraphLapParams, NLCGEngine, RegManager, TVND,
                          TVParams, _lib, synth)
from mrirecon_amd.nufft.backends.hip import HipNUFFTAdapter
from oracle.kbnufft import KBNufft, rel_l2
from oracle.nlcg import TV, GraphLap, majoriser_diag, nlcg, objective_value, reg_energy_grad, tv_energy_grad

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def c64(a):
    return torch.from_numpy(np.ascontiguousarray(a).astype(np.complex64)).to(DEV)


def crand(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def workspace():
    return torch.empty(int(_lib.load().mrh_reduce_workspace_len()), dtype=torch.float64, device=DEV)


def test_residual_energy_and_dot():
    lib = _lib.load()
    n = 300_001
    Ax, Ad, y = (c64(crand(n, s)) for s in (1, 2, 3))
    r = torch.empty_like(Ax)
    out = torch.zeros(1, dtype=torch.float64, device=DEV)
    w = workspace()
    _lib.check(lib.mrh_residual_energy(r.data_ptr(), Ax.data_ptr(), Ad.data_ptr(), y.data_ptr(), 0.37, n,
                                       w.data_ptr(), out.data_ptr(), _lib.stream_of(r)))
    rr = (0.37 * Ad.cpu().numpy().astype(np.complex128) + Ax.cpu().numpy() - y.cpu().numpy())
    assert abs(out.item() - np.sum(np.abs(rr) ** 2)) <= 1e-6 * np.sum(np.abs(rr) ** 2)
    assert rel_l2(r.cpu().numpy(), rr) <= 1e-6
    _lib.check(lib.mrh_dot_real(Ax.data_ptr(), Ad.data_ptr(), n, w.data_ptr(), out.data_ptr(), _lib.stream_of(r)))
    ref = np.vdot(Ax.cpu().numpy().astype(np.complex128), Ad.cpu().numpy().astype(np.complex128)).real
    assert abs(out.item() - ref) <= 1e-9 * np.sqrt(n) * 2


@pytest.mark.parametrize("iso", [True, False])
def test_tv_energy_grad_matches_oracle(iso):
    P, F, S0, S1 = 2, 3, 16, 20
    x = crand((P * F, S0, S1), 4)
    regs = [TVND("t", TVParams(weight=0.02, axes="temporal")),
            TVND("s", TVParams(weight=0.01, axes="spatial", isotropic=iso)),
            TVND("w", TVParams(weight=0.005, axes=(0, 2, 3), isotropic=iso, axis_weights=[0.5, 1.0, 2.0]))]
    regm = RegManager(regs)
    xt = c64(x)
    g = c64(np.zeros_like(x))
    e = regm.energy_and_grad_tensors(xt, g, problems=P).cpu().numpy()
    oregs = [TV(0.02, (0,)), TV(0.01, (1, 2), isotropic=iso),
             TV(0.005, (0, 1, 2), isotropic=iso, axis_weights=[0.5, 1.0, 2.0])]
    for p in range(P):
        xp = x[p * F:(p + 1) * F].astype(np.complex128)
        gp = np.zeros_like(xp)
        ep = sum(tv_energy_grad(xp, r, gp) for r in oregs)
        assert abs(e[p] - ep) <= 2e-6 * abs(ep), (p, e[p], ep)
        assert rel_l2(g.cpu().numpy()[p * F:(p + 1) * F], gp) <= 2e-6


@pytest.mark.parametrize("P,F,S0,S1", [(1, 34, 64, 48), (2, 3, 16, 20), (3, 1, 37, 23), (1, 7, 128, 96)])
def test_engine_gradient_equals_standalone_tv(P, F, S0, S1):
    """The solver's gradient pass (16 x 16 tiles walking their frames, each flux evaluated once) adds to g
    the bits of the standalone TV kernel (every backward flux re-evaluated from memory) for temporal,
    spatial iso / aniso and a weighted mixed-axes term, on tiles cut by the image edge, one frame, several
    problems and several frame groups; its energy partials sum to the same TV energy."""
    lib = _lib.load()
    regs = [TVND("t", TVParams(weight=0.02, axes="temporal")),
            TVND("s", TVParams(weight=0.01, axes="spatial", isotropic=True)),
            TVND("w", TVParams(weight=0.005, axes=(0, 2, 3), isotropic=False, axis_weights=[0.5, 1.0, 2.0]))]
    regm = RegManager(regs)
    x = c64(crand((P * F, S0, S1), 7 + F))
    g0 = crand((P * F, S0, S1), 8 + F)
    g_ref = c64(g0)
    e_ref = regm.energy_and_grad_tensors(x, g_ref, problems=P).cpu().numpy()
    desc = _lib.CGDesc()
    desc.tv = regm.tv_desc(P, F, S0, S1)
    k_prob = np.zeros(P + 1, dtype=np.int64)
    desc.k_prob = k_prob.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    desc.direction, desc.c1, desc.ls_beta, desc.ls_max_iter, desc.max_hist = 0, 1e-4, 0.5, 20, 4
    desc.restart_c, desc.ls_kind = -1.0, 0
    h = ctypes.c_void_p()
    _lib.check(lib.mrh_cg_create(ctypes.byref(desc), ctypes.byref(h)))
    try:
        g = c64(g0)           # the first gradient of a solve (prime)
        _lib.check(lib.mrh_cg_gradient(h, x.data_ptr(), g.data_ptr(), None, None, 1, _lib.stream_of(g)))
        g2 = c64(g0)          # an iteration's gradient: records the TV energy in the history
        _lib.check(lib.mrh_cg_gradient(h, x.data_ptr(), g2.data_ptr(), None, None, 0, _lib.stream_of(g)))
        buf = (_lib.CGHist * P)()
        _lib.check(lib.mrh_cg_read_history(h, buf, 1, _lib.stream_of(g)))
        torch.cuda.synchronize()
    finally:
        lib.mrh_cg_destroy(h)
    assert torch.equal(g, g_ref) and torch.equal(g2, g_ref)
    freg = np.array([buf[p].freg for p in range(P)])
    assert np.all(np.abs(freg - e_ref) <= 1e-9 * np.abs(e_ref)), (freg, e_ref)


def small_dynamic(seed=21, F=5, S=64, C=4, spokes=13, R=128):
    traj, dcf, off = synth.radial_frames([spokes] * F, R)
    x = synth.dynamic_phantom(S, F, seed=seed, n_ellipses=6)
    return synth.Workload("dyn", S, 1, F, C, R, 6, 2.0, [spokes] * F, synth.gaussian_maps(C, S), traj, dcf,
                          off, x, lam_t=0.02, lam_s=0.002, n_iter=15, noise_seed=seed + 1)


def prepare(wl):
    be = HipNUFFTAdapter(ndim=2, traj_units="rad", kb_width=wl.W, osf=wl.osf)
    be.prepare((wl.n_coils, wl.S, wl.S), torch.from_numpy(wl.traj), torch.from_numpy(wl.dcf), c64(wl.maps),
               device=DEV, k_offsets=wl.k_offsets)
    ora = KBNufft(wl.maps, wl.traj, wl.k_offsets, wl.dcf, wl.W, wl.osf)
    synth.measure(wl, ora.forward)
    return be, ora


def regs_for(wl):
    g, o = [], []
    if wl.lam_t:
        g.append(TVND("tv_t", TVParams(weight=wl.lam_t, axes="temporal")))
        o.append(TV(wl.lam_t, (0,)))
    if wl.lam_s:
        g.append(TVND("tv_s", TVParams(weight=wl.lam_s, axes="spatial")))
        o.append(TV(wl.lam_s, (1, 2)))
    return RegManager(g), o


def run_pair(wl, n_iter, direction="fr", use_graph=True, rel_tol=None, ls_name="armijo", c2=0.9, precond=False):
    be, ora = prepare(wl)
    regm, oregs = regs_for(wl)
    x0 = ora.adjoint(wl.y)
    eng = NLCGEngine(be.plan, c64(wl.y), problems=1, frames=wl.B, regm=regm, direction=direction,
                     rel_tol=rel_tol or 0.0, ls_name=ls_name, c2=c2, precond=precond)
    eng.prime(c64(x0))
    eng.run(n_iter, use_graph=use_graph)
    torch.cuda.synchronize()
    hist = eng.history()[0]
    x_gpu = eng.x.cpu().numpy()
    ref = nlcg(ora.forward, ora.adjoint, wl.y.astype(np.complex128), x0.astype(np.complex128), oregs,
               max_iter=n_iter, direction=direction, mutate_x=False, rel_tol=rel_tol, ls_name=ls_name, c2=c2,
               precond=precond)
    f_gpu = objective_value(ora.forward, wl.y, x_gpu.astype(np.complex128), oregs)
    return eng, hist, ref, f_gpu


def test_nlcg_config1_matches_reference_semantics():
    """c1: 20 FR iterations, spatial TV 0.01 -- objective within 1e-4 of the oracle run."""
    wl = synth.config_c1()
    eng, hist, ref, f_gpu = run_pair(wl, 20)
    fh = np.array([h["f"] for h in hist])
    fr = np.array([h["f"] for h in ref.history])
    steps = [h["step"] for h in hist]
    print(f"[c1] f_gpu={f_gpu:.8e} f_ref={ref.f_final:.8e} rel={abs(f_gpu - ref.f_final) / ref.f_final:.2e} "
          f"steps={steps}")
    assert len(hist) == 20 and ref.n_iter == 20
    assert np.max(np.abs(fh - fr) / np.abs(fr)) <= 1e-4
    assert abs(f_gpu - ref.f_final) <= 1e-4 * abs(ref.f_final)
    assert [h["step"] for h in ref.history] == steps
    for h in hist:       # Armijo inequality on the device-side selection
        assert h["f_step"] <= h["f"] + 1e-4 * h["step"] * h["gdot"] + 1e-9 * abs(h["f"])


@pytest.mark.parametrize("direction", ["fr", "prplus", "dy"])
def test_nlcg_dynamic_temporal_spatial_tv(direction):
    wl = small_dynamic()
    eng, hist, ref, f_gpu = run_pair(wl, 15, direction=direction)
    rel = abs(f_gpu - ref.f_final) / ref.f_final
    print(f"[dyn {direction}] f_gpu={f_gpu:.8e} f_ref={ref.f_final:.8e} rel={rel:.2e} iters={len(hist)}")
    assert rel <= 1e-4
    fh = np.array([h["f"] for h in hist])
    fr = np.array([h["f"] for h in ref.history])
    assert len(fh) == len(fr)
    assert np.max(np.abs(fh - fr) / np.abs(fr)) <= 1e-4
    if direction == "dy":      # Re<g_k, d_k> = beta_k Re<g_(k-1), d_(k-1)> on the device history too
        for a, b in zip(hist, hist[1:]):
            assert abs(b["gdot"] - a["beta"] * a["gdot"]) <= 1e-6 * abs(b["gdot"])


def check_wolfe_run(hist, ref, f_gpu, c2, tag):
    steps = [h["step"] for h in hist]
    print(f"[wolfe {tag}] f_gpu={f_gpu:.8e} f_ref={ref.f_final:.8e} steps={steps} stop={ref.stop}")
    assert [h["step"] for h in ref.history] == steps
    fh = np.array([h["f"] for h in hist])
    fr = np.array([h["f"] for h in ref.history])
    assert np.max(np.abs(fh - fr) / np.abs(fr)) <= 1e-4
    assert abs(f_gpu - ref.f_final) <= 1e-4 * abs(ref.f_final)
    for h, r in zip(hist, ref.history):
        # Armijo and the strong curvature condition at every accepted step (the slope at the step
        # from the gradient pass, float rounding aside), as graspcg/tests/test_cg_descend_compile_ns.py:125-176
        assert h["f_step"] <= h["f"] + 1e-4 * h["step"] * h["gdot"] + 1e-9 * abs(h["f"])
        assert abs(h["gdot_at_step"]) <= c2 * abs(h["gdot"]) * (1 + 1e-4)
        assert abs(h["gdot_at_step"] - r["gdot_at_step"]) <= 1e-3 * abs(r["gdot"])
    return steps


@pytest.mark.parametrize("direction", ["fr", "prplus", "dy"])
def test_nlcg_wolfe_matches_oracle(direction):
    """Strong-Wolfe search (graspcg/numerics/line_search.py:89-204, c2 = 0.1) with temporal + spatial TV:
    the device replay of the search accepts the oracle's steps (zooms from (0, 1) two passes deep),
    objectives within 1e-4."""
    wl = small_dynamic(seed=71)
    eng, hist, ref, f_gpu = run_pair(wl, 12, direction=direction, ls_name="wolfe", c2=0.1)
    assert len(hist) == 12 and ref.n_iter == 12
    steps = check_wolfe_run(hist, ref, f_gpu, 0.1, direction)
    assert any(s * 32 != round(s * 32) for s in steps)   # a bisection below 1/32: a second zoom pass


def test_wolfe_expansion_and_zoom_over_several_passes():
    """Coil maps scaled by 1/16 (|A|^2 / 256), no regulariser: the search doubles to t = 256, past the
    first pass's six trials, and zooms back (160, 64, 48, 72, ...); steps as in the oracle run."""
    wl = small_dynamic(seed=61)
    wl.maps = wl.maps / 16.0
    wl.lam_t = wl.lam_s = 0.0
    eng, hist, ref, f_gpu = run_pair(wl, 6, ls_name="wolfe", c2=0.1)
    steps = check_wolfe_run(hist, ref, f_gpu, 0.1, "expansion")
    assert max(steps) > 64


def test_wolfe_small_steps_zoom_deep():
    """Coil maps scaled by 8: accepted steps near 2^-8, about ten bisections of (0, 1) (four passes)."""
    wl = small_dynamic(seed=61)
    wl.maps = wl.maps * 8.0
    eng, hist, ref, f_gpu = run_pair(wl, 6, ls_name="wolfe", c2=0.1)
    steps = check_wolfe_run(hist, ref, f_gpu, 0.1, "deep zoom")
    assert min(steps) < 2.0 ** -8


def test_majoriser_diag_matches_oracle():
    """mrh_cg_build_diag (via prime) == the oracle's RegManager.add_diag restatement: temporal profile,
    an anisotropic lagged-diffusivity term with axis weights, and a term over frames and image axes
    together (its frame profile only), and a graph-Laplacian term (lambda x degree per frame).  (The
    isotropic spatial term is covered by the preconditioned NLCG runs below.)"""
    wl = small_dynamic(seed=81, F=4)
    be, ora = prepare(wl)
    regs = [TVND("t", TVParams(weight=0.02, axes="temporal")),
            TVND("a", TVParams(weight=0.001, axes="spatial", isotropic=False, eps=2e-3, axis_weights=[0.5, 2.0])),
            TVND("m", TVParams(weight=0.004, axes=(0, 2, 3), axis_weights=[1.5, 1.0, 1.0]))]
    oregs = [TV(0.02, (0,)), TV(0.001, (1, 2), isotropic=False, eps=2e-3, axis_weights=[0.5, 2.0]),
             TV(0.004, (0, 1, 2), axis_weights=[1.5, 1.0, 1.0])]
    x0 = crand((wl.B, wl.S, wl.S), 82) * 0.01
    x0[:, :10, :] = 0.25        # flat rows: weights near 1 / eps
    W = graph_W(wl.B, 83)
    regs = regs + [GraphLaplacian("gl", W, GraphLapParams(weight=0.3))]
    oregs = oregs + [GraphLap(0.3, W)]
    eng = NLCGEngine(be.plan, c64(wl.y), problems=1, frames=wl.B, regm=RegManager(regs), precond=True)
    eng.prime(c64(x0))
    torch.cuda.synchronize()
    D = eng.diag.cpu().numpy().astype(np.float64)
    ref = majoriser_diag(x0.astype(np.complex128), oregs)
    err = np.max(np.abs(D - ref) / ref)
    print(f"[majoriser] max rel err {err:.2e}, D in [{ref.min():.3f}, {ref.max():.3f}]")
    assert err <= 2e-6
    assert ref.max() > 5.0 and ref.min() < 1.7     # flat rows and edges both present


@pytest.mark.parametrize("direction", ["fr", "prplus", "dy"])
def test_preconditioned_nlcg_matches_oracle(direction):
    """Directions preconditioned by the regularisers' majoriser at x0 (d = beta d - g / D, 1/D-weighted
    beta inner products): accepted steps equal the oracle's, objectives within 1e-4."""
    wl = small_dynamic(seed=71)
    eng, hist, ref, f_gpu = run_pair(wl, 12, direction=direction, precond=True)
    steps = [h["step"] for h in hist]
    rel = abs(f_gpu - ref.f_final) / ref.f_final
    print(f"[precond {direction}] f_gpu={f_gpu:.8e} f_ref={ref.f_final:.8e} rel={rel:.2e} steps={steps} "
          f"stop={ref.stop}")
    assert [h["step"] for h in ref.history] == steps
    fh = np.array([h["f"] for h in hist])
    fr = np.array([h["f"] for h in ref.history])
    assert np.max(np.abs(fh - fr) / np.abs(fr)) <= 1e-4
    assert rel <= 1e-4
    gn = np.array([h["gnorm"] for h in hist])
    gr = np.array([h["gnorm"] for h in ref.history])
    assert np.max(np.abs(gn - gr) / gr) <= 1e-3      # the 1/D-weighted gradient norm


def graph_W(F, seed):
    rng = np.random.default_rng(seed)
    W = np.zeros((F, F))
    for t in range(F - 1):
        W[t, t + 1] = W[t + 1, t] = 1.0
    W[0, F - 1] = W[F - 1, 0] = rng.random()     # a weighted wrap-around edge
    return W


@pytest.mark.parametrize("normalize,ls_name", [("none", "armijo"), ("sym", "armijo"), ("rw", "armijo"),
                                               ("none", "wolfe")])
def test_graph_laplacian_nlcg_matches_oracle(normalize, ls_name):
    """Graph-Laplacian term along frames (graph_laplacian.py:22-321) with temporal + spatial TV:
    k_gl_grad (g += lambda L x, its energy) and k_gl_ls (the term's quadratic along d) give the
    oracle's steps and objectives within 1e-4, for each normalisation and both line searches."""
    wl = small_dynamic(seed=91)
    be, ora = prepare(wl)
    regm, oregs = regs_for(wl)
    W = graph_W(wl.B, 92)
    lam = 0.05
    regm = RegManager(regm.regs + [GraphLaplacian("gl", W, GraphLapParams(weight=lam, normalize=normalize))])
    oregs = oregs + [GraphLap(lam, W, normalize=normalize)]
    x0 = ora.adjoint(wl.y)
    eng = NLCGEngine(be.plan, c64(wl.y), problems=1, frames=wl.B, regm=regm, direction="prplus", restart_c=0.0,
                     ls_name=ls_name, c2=0.1)
    eng.prime(c64(x0))
    eng.run(10)
    torch.cuda.synchronize()
    hist = eng.history()[0]
    ref = nlcg(ora.forward, ora.adjoint, wl.y.astype(np.complex128), x0.astype(np.complex128), oregs,
               max_iter=10, direction="prplus", mutate_x=False, restart_c=0.0, ls_name=ls_name, c2=0.1)
    f_gpu = objective_value(ora.forward, wl.y, eng.x.cpu().numpy().astype(np.complex128), oregs)
    steps = [h["step"] for h in hist]
    print(f"[graph {normalize} {ls_name}] f_gpu={f_gpu:.8e} f_ref={ref.f_final:.8e} steps={steps}")
    assert [h["step"] for h in ref.history] == steps
    fh = np.array([h["f"] for h in hist])
    fr = np.array([h["f"] for h in ref.history])
    assert np.max(np.abs(fh - fr) / np.abs(fr)) <= 1e-4
    assert abs(f_gpu - ref.f_final) <= 1e-4 * abs(ref.f_final)
    # the term's energy is part of the regulariser energy at each accepted iterate
    x1 = eng.x.cpu().numpy().astype(np.complex128)
    e_ref = sum(float(reg_energy_grad(x1, r)) for r in oregs)
    assert abs(hist[-1]["f_reg"] - e_ref) <= 1e-4 * abs(e_ref)


def test_small_steps_take_the_second_line_search_pass():
    """Coil maps scaled by 8 (|A|^2 x 64): Armijo stops near t = 2^-6, past the five steps of the
    first image pass, so the second pass selects; steps and objectives as in the oracle run."""
    wl = small_dynamic(seed=61)
    wl.maps = wl.maps * 8.0
    eng, hist, ref, f_gpu = run_pair(wl, 8)
    steps = [h["step"] for h in hist]
    print(f"[small steps] {steps}")
    assert [h["step"] for h in ref.history] == steps
    assert min(steps) <= 0.5 ** 5
    fh = np.array([h["f"] for h in hist])
    fr = np.array([h["f"] for h in ref.history])
    assert np.max(np.abs(fh - fr) / np.abs(fr)) <= 1e-4
    assert abs(f_gpu - ref.f_final) <= 1e-4 * abs(ref.f_final)


def test_mixed_axes_tv_term_matches_oracle():
    """A TV term over (frame, x, y) together, with axis weights, inside NLCG: objectives as in
    the oracle run."""
    wl = small_dynamic(seed=71)
    be, ora = prepare(wl)
    regm = RegManager([TVND("w", TVParams(weight=0.01, axes=(0, 2, 3), isotropic=True,
                                          axis_weights=[0.5, 1.0, 2.0]))])
    oregs = [TV(0.01, (0, 1, 2), isotropic=True, axis_weights=[0.5, 1.0, 2.0])]
    x0 = ora.adjoint(wl.y)
    eng = NLCGEngine(be.plan, c64(wl.y), problems=1, frames=wl.B, regm=regm, direction="prplus", rel_tol=0.0)
    eng.prime(c64(x0))
    eng.run(10)
    torch.cuda.synchronize()
    hist = eng.history()[0]
    ref = nlcg(ora.forward, ora.adjoint, wl.y.astype(np.complex128), x0.astype(np.complex128), oregs,
               max_iter=10, direction="prplus", mutate_x=False)
    fh = np.array([h["f"] for h in hist])
    fr = np.array([h["f"] for h in ref.history])
    assert len(fh) == len(fr) == 10
    assert np.max(np.abs(fh - fr) / np.abs(fr)) <= 1e-4


@pytest.mark.parametrize("ls_name", ["armijo", "wolfe"])
def test_graph_replay_equals_eager(ls_name):
    wl = small_dynamic(seed=31)
    _, h_graph, _, f_graph = run_pair(wl, 8, use_graph=True, ls_name=ls_name, c2=0.1)
    _, h_eager, _, f_eager = run_pair(wl, 8, use_graph=False, ls_name=ls_name, c2=0.1)
    for a, b in zip(h_graph, h_eager):
        assert a["step"] == b["step"]
        assert abs(a["f"] - b["f"]) <= 1e-6 * abs(a["f"])
    assert abs(f_graph - f_eager) <= 1e-6 * abs(f_eager)


class _AxYLineSearchEngine(NLCGEngine):
    """The engine with the line search reading Ax and y (the entry point's other input form)."""

    calls = 0

    def step(self) -> None:
        lib = self._lib
        real = lib.mrh_cg_linesearch

        def with_ax_y(eng, x, d, g, r, Ad, y, st):
            assert y is None                      # the engine's own call passes the residual
            type(self).calls += 1
            return real(eng, x, d, g, self.Ax.data_ptr(), Ad, self.y.data_ptr(), st)

        lib.mrh_cg_linesearch = with_ax_y
        try:
            super().step()
        finally:
            lib.mrh_cg_linesearch = real


@pytest.mark.parametrize("ls_name", ["armijo", "wolfe"])
def test_line_search_from_residual_equals_from_ax_and_y(ls_name):
    """mrh_cg_linesearch with y null reads the residual r that prime and accept leave (the engine's
    form); with Ax and y it forms r itself: the same FP32 subtraction, so both trajectories agree bit
    for bit (steps, objectives and the final image)."""
    wl = small_dynamic(seed=33)
    be, ora = prepare(wl)
    regm, _ = regs_for(wl)
    x0 = c64(ora.adjoint(wl.y))
    runs = []
    _AxYLineSearchEngine.calls = 0
    for cls in (NLCGEngine, _AxYLineSearchEngine):
        eng = cls(be.plan, c64(wl.y), problems=1, frames=wl.B, regm=regm, direction="prplus", ls_name=ls_name,
                  c2=0.1)
        eng.prime(x0)
        eng.run(8, use_graph=False)
        torch.cuda.synchronize()
        runs.append((eng.history()[0], eng.x.clone()))
    (h_r, x_r), (h_a, x_a) = runs
    assert _AxYLineSearchEngine.calls >= 8
    assert len(h_r) == len(h_a) == 8
    for a, b in zip(h_r, h_a):
        assert a["step"] == b["step"] and a["f"] == b["f"] and a["f_step"] == b["f_step"]
    assert torch.equal(x_r, x_a)


def test_rel_tol_stop_rule():
    wl = small_dynamic(seed=41)
    eng, hist, ref, _ = run_pair(wl, 60, rel_tol=1e-4)
    print(f"[dyn rel_tol] gpu iters={len(hist)} stop={eng.stop_reasons()} ref iters={ref.n_iter} stop={ref.stop}")
    assert eng.stop_reasons()[0] == "rel_tol" and ref.stop == "rel_tol"
    assert abs(len(hist) - ref.n_iter) <= 1


@pytest.mark.parametrize("ls_name,extra", [("armijo", ""), ("wolfe", ""), ("armijo", "graph+precond")])
def test_batched_problems_equal_separate_runs(ls_name, extra):
    """Two slices with uneven spokes in one engine == each slice alone (with either search; the
    strong-Wolfe passes skip the slice whose search has ended; the graph term and the majoriser
    preconditioner work per slice)."""
    S, C, R, F = 48, 3, 96, 4
    spokes = [[13, 21, 17, 30], [25, 14, 40, 19]]
    ws = []
    for i, sp in enumerate(spokes):
        traj, dcf, off = synth.radial_frames(sp, R)
        ws.append(synth.Workload(f"s{i}", S, 1, F, C, R, 6, 2.0, sp, synth.gaussian_maps(C, S), traj, dcf, off,
                                 synth.dynamic_phantom(S, F, seed=50 + i, n_ellipses=5), lam_t=0.02,
                                 noise_seed=60 + i))
    kw = dict(ls_name=ls_name, c2=0.1)
    W = graph_W(F, 64)

    def regs(wl):
        regm = regs_for(wl)[0]
        if extra:
            regm = RegManager(regm.regs + [GraphLaplacian("gl", W, GraphLapParams(weight=0.05, normalize="sym"))])
        return regm

    if extra:
        kw["precond"] = True
    solo = []
    for wl in ws:
        be, ora = prepare(wl)
        regm = regs(wl)
        eng = NLCGEngine(be.plan, c64(wl.y), problems=1, frames=F, regm=regm, **kw)
        eng.prime(c64(ora.adjoint(wl.y)))
        eng.run(6)
        solo.append((eng.history()[0], eng.x.cpu().numpy()))
    traj = np.concatenate([w.traj for w in ws], axis=1)
    dcf = np.concatenate([w.dcf for w in ws])
    off = np.concatenate([[0], np.cumsum([o for w in ws for o in np.diff(w.k_offsets)])])
    be = HipNUFFTAdapter(ndim=2, traj_units="rad", kb_width=6)
    be.prepare((C, S, S), torch.from_numpy(traj), torch.from_numpy(dcf), c64(ws[0].maps), device=DEV,
               k_offsets=off)
    y = np.concatenate([w.y for w in ws])
    x0 = np.concatenate([KBNufft(w.maps, w.traj, w.k_offsets, w.dcf, 6).adjoint(w.y) for w in ws])
    eng = NLCGEngine(be.plan, c64(y), problems=2, frames=F, regm=regs(ws[0]), **kw)
    eng.prime(c64(x0))
    eng.run(6)
    hb = eng.history()
    xb = eng.x.cpu().numpy()
    for p in range(2):
        assert [h["step"] for h in hb[p]] == [h["step"] for h in solo[p][0]]
        for a, b in zip(hb[p], solo[p][0]):
            assert abs(a["f"] - b["f"]) <= 1e-6 * abs(b["f"])
        assert rel_l2(xb[p * F:(p + 1) * F], solo[p][1]) <= 1e-5


def test_cgsolver_reference_surface():
    """CGSolver(y, nufft_op, regm, cfg).run().result()/history() with the reference's keys."""
    wl = synth.config_c1()
    be, ora = prepare(wl)
    regm, _ = regs_for(wl)
    y = c64(wl.y).reshape(1, wl.n_coils, -1)
    solver = CGSolver(y, be, regm, CGConfig(max_iter=5, init_mode="backproj"))
    solver.run()
    h = solver.history()
    assert len(h) == 5
    for k in ("iter", "f", "gdot", "gdot_at_step", "step", "xnorm", "gnorm", "stepnorm"):
        assert k in h[0]
    assert tuple(solver.result().shape) == (1, 1, wl.S, wl.S)
    assert solver.ws.get("x").data_ptr() == solver.ws.get("var").data_ptr()
    with pytest.raises(ValueError):
        CGSolver(y, be, regm, CGConfig(ls_name="more-thuente"))
    psolver = CGSolver(y, be, regm, CGConfig(max_iter=3, init_mode="backproj", precond=True, restart_c=0.0))
    psolver.run()
    assert len(psolver.history()) == 3 and float(psolver.ws.get("diag").min()) >= 1.0
    wsolver = CGSolver(y, be, regm, CGConfig(max_iter=4, init_mode="backproj", ls_name="wolfe", c2=0.5))
    wsolver.run()
    for h in wsolver.history():
        assert abs(h["gdot_at_step"]) <= 0.5 * abs(h["gdot"]) * (1 + 1e-4)


def test_restart_safeguard_matches_oracle():
    """The device beta restart (restart_c = 0) follows the oracle's restarted PR+ run to the
    relative-change stop: same iteration count (+-1), objective within 1e-4."""
    wl = small_dynamic(seed=21, F=5)
    be, ora = prepare(wl)
    regm, oregs = regs_for(wl)
    x0 = ora.adjoint(wl.y)
    eng = NLCGEngine(be.plan, c64(wl.y), problems=1, frames=wl.B, regm=regm, direction="prplus", rel_tol=1e-4,
                     restart_c=0.0, max_hist=80)
    eng.prime(c64(x0))
    eng.run(60)
    torch.cuda.synchronize()
    hist = eng.history()[0]
    ref = nlcg(ora.forward, ora.adjoint, wl.y.astype(np.complex128), x0.astype(np.complex128), oregs,
               max_iter=60, direction="prplus", mutate_x=False, rel_tol=1e-4, restart_c=0.0)
    f_gpu = objective_value(ora.forward, wl.y, eng.x.cpu().numpy().astype(np.complex128), oregs)
    print(f"[restart] gpu iters={len(hist)} stop={eng.stop_reasons()} ref iters={ref.n_iter} stop={ref.stop} "
          f"f_gpu={f_gpu:.8e} f_ref={ref.f_final:.8e}")
    assert eng.stop_reasons()[0] == "rel_tol" and ref.stop == "rel_tol"
    assert abs(len(hist) - ref.n_iter) <= 1
    assert abs(f_gpu - ref.f_final) <= 1e-4 * abs(ref.f_final)
    assert all(h["gdot"] < 0 for h in hist)


def test_objective_target_stop():
    """mrh_cg_set_targets: a run stops at the first accepted step with f <= target, i.e. the
    iteration where an identical run without a target first reached that value."""
    wl = small_dynamic(seed=23, F=5)
    be, ora = prepare(wl)
    regm, _ = regs_for(wl)
    x0 = c64(ora.adjoint(wl.y))
    eng = NLCGEngine(be.plan, c64(wl.y), problems=1, frames=wl.B, regm=regm, direction="prplus",
                     restart_c=0.0, max_hist=40)
    eng.prime(x0)
    eng.run(12)
    h = eng.history()[0]
    assert eng.stop_reasons() == ["active"] and len(h) == 12
    target = h[6]["f_step"]
    eng.set_targets([target])
    eng.prime(x0)
    eng.run(12)
    h2 = eng.history()[0]
    assert eng.stop_reasons() == ["target"] and len(h2) == 7
    assert h2[-1]["f_step"] == target
    eng.set_targets(None)
    eng.prime(x0)
    eng.run(12)
    assert eng.stop_reasons() == ["active"]


def test_prime_with_precomputed_forward():
    """prime(x0, Ax0=A x0) runs exactly the iterations prime(x0) runs."""
    wl = small_dynamic(seed=27, F=4)
    be, ora = prepare(wl)
    regm, _ = regs_for(wl)
    x0 = c64(ora.adjoint(wl.y))
    eng = NLCGEngine(be.plan, c64(wl.y), problems=1, frames=wl.B, regm=regm, direction="prplus", restart_c=0.0)
    eng.prime(x0)
    eng.run(6)
    h1 = eng.history()[0]
    Ax0 = torch.empty_like(eng.Ax)
    be.plan.forward(x0, Ax0)
    eng.prime(x0, Ax0)
    eng.run(6)
    h2 = eng.history()[0]
    assert [r["step"] for r in h1] == [r["step"] for r in h2]
    assert all(abs(a["f_step"] - b["f_step"]) <= 1e-9 * abs(a["f_step"]) for a, b in zip(h1, h2))


def test_skipping_stopped_problems_changes_nothing():
    """With the masked NUFFTs (frames of stopped slices skipped) a batch whose slices stop at
    different iterations gives the same histories and images as without them."""
    S, C, R, F = 48, 3, 96, 4
    spokes = [[13, 21, 17, 30], [25, 14, 40, 19], [33, 12, 28, 16]]
    traj, dcf, off = synth.radial_frames([s for sp in spokes for s in sp], R)
    x_true = np.concatenate([synth.dynamic_phantom(S, F, seed=80 + i, n_ellipses=5) for i in range(3)])
    maps = synth.gaussian_maps(C, S)
    be = HipNUFFTAdapter(ndim=2, traj_units="rad", kb_width=6)
    be.prepare((C, S, S), torch.from_numpy(traj), torch.from_numpy(dcf), c64(maps), device=DEV, k_offsets=off)
    y = torch.empty(be.plan.n_k, dtype=torch.complex64, device=DEV)
    be.plan.forward(c64(x_true), y)
    y += 0.01 * y.abs().max() * torch.randn(y.shape, dtype=torch.complex64, device=DEV,
                                            generator=torch.Generator(DEV).manual_seed(5))
    x0 = torch.empty((3 * F, S, S), dtype=torch.complex64, device=DEV)
    be.plan.adjoint(y, x0)
    regm = RegManager([TVND("t", TVParams(weight=0.02, axes="temporal"))])
    out = []
    # (skip stopped slices, iterations per graph replay): one replay per iteration with and without
    # the masked NUFFTs, and 4 or 7 iterations per replay (slices stop inside a replay)
    for skip, ipr in ((False, 1), (True, 1), (True, 4), (True, 7), (False, 4)):
        eng = NLCGEngine(be.plan, y, problems=3, frames=F, regm=regm, direction="prplus", rel_tol=1e-3,
                         restart_c=0.0, skip_stopped=skip, max_hist=80, iters_per_replay=ipr)
        eng.prime(x0)
        eng.run(60)
        torch.cuda.synchronize()
        out.append((eng.history(), eng.x.cpu().numpy(), eng.status()[1]))
    h0, x0n, it0 = out[0]
    assert len(set(it0.tolist())) > 1, it0                                 # slices stop at different iterations
    for h1, x1n, it1 in out[1:]:
        assert list(it0) == list(it1)
        for p in range(3):
            assert [r["step"] for r in h0[p]] == [r["step"] for r in h1[p]]
            assert [r["f_step"] for r in h0[p]] == [r["f_step"] for r in h1[p]]
        assert np.array_equal(x0n, x1n)


def test_solver_option_limits_are_reported():
    """Engine limits fail loudly at creation: a graph term whose frame count differs from the problems',
    an unknown line search, a negative c2."""
    wl = small_dynamic(seed=95, F=5)
    be, ora = prepare(wl)
    regm, _ = regs_for(wl)
    bad = RegManager(regm.regs + [GraphLaplacian("gl", np.ones((4, 4)), GraphLapParams(weight=0.1))])
    with pytest.raises(ValueError):
        NLCGEngine(be.plan, c64(wl.y), problems=1, frames=wl.B, regm=bad)
    with pytest.raises(ValueError):
        NLCGEngine(be.plan, c64(wl.y), problems=1, frames=wl.B, regm=regm, ls_name="backtrack")
    with pytest.raises(ValueError):
        NLCGEngine(be.plan, c64(wl.y), problems=1, frames=wl.B, regm=regm, ls_name="wolfe", c2=-0.5)


def test_latest_record_read_equals_the_history_tail():
    """mrh_cg_read_latest: status, counts and each problem's last record in one read, equal to
    the last row of the full history (two problems that stop at different iterations)."""
    wl = small_dynamic(seed=33, F=4)
    be, ora = prepare(wl)
    regm, _ = regs_for(wl)
    y = np.concatenate([wl.y, wl.y * 0.5])
    traj = np.concatenate([wl.traj, wl.traj], axis=1)
    off = np.concatenate([wl.k_offsets, wl.k_offsets[1:] + wl.k_offsets[-1]])
    be2 = HipNUFFTAdapter(ndim=2, traj_units="rad", kb_width=wl.W, osf=wl.osf)
    be2.prepare((wl.n_coils, wl.S, wl.S), torch.from_numpy(traj), torch.from_numpy(np.concatenate([wl.dcf] * 2)),
                c64(wl.maps), device=DEV, k_offsets=off)
    eng = NLCGEngine(be2.plan, c64(y), problems=2, frames=wl.B, regm=regm, direction="prplus", rel_tol=1e-3,
                     restart_c=0.0)
    st0, it0, last0 = eng.latest()
    assert list(it0) == [0, 0] and last0[0]["gnorm"] == 0.0
    x0 = np.concatenate([ora.adjoint(wl.y), ora.adjoint(wl.y * 0.5)])
    eng.prime(c64(x0))
    eng.run(40)
    st, it, last = eng.latest()
    hist = eng.history()
    s2, i2 = eng.status()
    assert list(st) == list(s2) and list(it) == list(i2)
    for p in range(2):
        h = hist[p][-1]
        for k in ("gnorm", "xnorm", "stepnorm", "f", "f_step", "step"):
            assert last[p][k] == h[k], (p, k)


def test_fp32_reduction_variant():
    """The all-FP32 reduction variant (the reference's precision) on c1: it runs the same loop
    (20 FR iterations), its objective stays within 1e-3 of the FP64 engine's, and its reductions
    are really FP32 (the FP64 and FP32 histories differ)."""
    wl = synth.config_c1()
    be, ora = prepare(wl)
    regm, oregs = regs_for(wl)
    x0 = c64(ora.adjoint(wl.y))
    fs = []
    for fp32 in (False, True):
        eng = NLCGEngine(be.plan, c64(wl.y), problems=1, frames=1, regm=regm, fp32_reductions=fp32)
        eng.prime(x0)
        eng.run(20)
        torch.cuda.synchronize()
        h = eng.history()[0]
        assert len(h) == 20
        f = objective_value(ora.forward, wl.y, eng.x.cpu().numpy().astype(np.complex128), oregs)
        fs.append((f, [r["f"] for r in h]))
        eng.close()
    (f64, h64), (f32, h32) = fs
    print(f"[c1] f after 20 iterations: FP64 {f64:.8e}, FP32 {f32:.8e}, rel gap {(f32 - f64) / f64:.2e}")
    assert abs(f32 - f64) <= 1e-3 * abs(f64)
    assert h64 != h32
    with pytest.raises(NotImplementedError):
        NLCGEngine(be.plan, c64(wl.y), problems=1, frames=1, regm=regm, fp32_reductions=True, ls_name="wolfe")
