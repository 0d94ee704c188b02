"""Pins of the CPU oracle's KB-gridding NUFFT (oracle/kbnufft.py).

The reference's only numeric pin for the operator is the FP64 NUDFT of
graspcg/tests/test_nufft_precision_gt.py:48-183 (tolerance 2e-2); the checks
below add closed forms (KB Fourier transform against quadrature, delta-image
known answers), adjointness, and planted mistakes that must fail.
"""
import math

import numpy as np
import pytest
from scipy import integrate

from oracle.kbnufft import (KBNufft, deapod_1d, footprint, kb_beta, kb_ft, kb_phi, nudft_adjoint,
                            nudft_forward, rel_l2)


def reference_precision_problem(seed=0):
    """Problem of graspcg/tests/test_nufft_precision_gt.py:136-160: B=2, L=2, C=3, 32x32, K=1024."""
    rng = np.random.default_rng(seed)
    B, L, C, S, K = 2, 2, 3, 32, 1024
    maps = (rng.standard_normal((C, S, S)) + 1j * rng.standard_normal((C, S, S))) / math.sqrt(C)
    traj = rng.uniform(-math.pi, math.pi, (B, 2, K))
    dcf = rng.uniform(0.1, 1.1, (B, K))
    x = rng.standard_normal((B, L, S, S)) + 1j * rng.standard_normal((B, L, S, S))
    return B, L, C, S, K, maps, traj, dcf, x


def nudft_pair(maps, traj, dcf, x):
    B = traj.shape[0]
    ygt = np.stack([nudft_forward(maps, x[b], traj[b], dcf[b]) for b in range(B)])
    xgt = np.stack([nudft_adjoint(maps, ygt[b], traj[b], dcf[b]) for b in range(B)])
    return ygt, xgt


def flat_traj(traj):
    """(B, 2, K) -> (2, B*K) in frame order (the oracle's flat layout)."""
    return traj.transpose(1, 0, 2).reshape(2, -1)


@pytest.mark.parametrize("W,bound", [(4, 2e-3), (6, 3e-5), (8, 5e-7)])
def test_kb_oracle_matches_nudft(W, bound):
    B, L, C, S, K, maps, traj, dcf, x = reference_precision_problem()
    op = KBNufft(maps, flat_traj(traj), np.arange(B + 1) * K, dcf.reshape(-1), W, 2.0, like=L)
    ygt, xgt = nudft_pair(maps, traj, dcf, x)
    relA = rel_l2(op.forward(x.reshape(B * L, S, S)), ygt)
    relAH = rel_l2(op.adjoint(ygt.reshape(-1)), xgt)
    assert relA <= bound and relAH <= bound, (relA, relAH)
    assert relA <= 2e-2 and relAH <= 2e-2        # the reference test's own bar


def test_kb_oracle_adjoint_identity():
    B, L, C, S, K, maps, traj, dcf, x = reference_precision_problem(1)
    op = KBNufft(maps, flat_traj(traj), np.arange(B + 1) * K, dcf.reshape(-1), 6, 2.0, like=L)
    rng = np.random.default_rng(5)
    yr = rng.standard_normal(op.n_k) + 1j * rng.standard_normal(op.n_k)
    lhs = np.vdot(op.forward(x.reshape(B * L, S, S)), yr)
    rhs = np.vdot(x.reshape(-1), op.adjoint(yr).reshape(-1))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


@pytest.mark.parametrize("W", [4, 6])
def test_kb_fourier_transform_closed_form(W):
    beta = kb_beta(W, 2.0)
    for xi in (0.0, 0.05, 0.13, 0.25, 0.31):
        num, _ = integrate.quad(lambda s: float(kb_phi(s, W, beta)) * math.cos(2 * math.pi * xi * s),
                                -W / 2, W / 2, epsabs=1e-14, epsrel=1e-13, limit=200)
        assert abs(num - float(kb_ft(xi, W, beta))) <= 1e-10 * abs(num), xi


def test_beta_formula_values():
    # beta = pi sqrt((W/osf (osf - 1/2))^2 - 0.8) (graspcg/nufft/backends/ptnufft.py:125)
    assert abs(kb_beta(4, 2.0) - 8.9955) < 1e-3
    assert abs(kb_beta(6, 2.0) - 13.8551) < 1e-3


def test_footprint_covers_kernel_support():
    rng = np.random.default_rng(2)
    for W in (4, 5, 6):
        u = rng.uniform(-300, 300, 1000)
        m0, w = footprint(u, W, kb_beta(W, 2.0))
        d = u[:, None] - (m0[:, None] + np.arange(W))
        assert np.all(np.abs(d) <= W / 2 + 1e-12)
        assert np.all(w > 0)


def test_delta_known_answers():
    """delta at the centre -> y_ck = S_c(ctr) sqrt(dcf_k); delta at r0 -> plane wave."""
    rng = np.random.default_rng(3)
    S, C, K = 32, 2, 512
    maps = rng.standard_normal((C, S, S)) + 1j * rng.standard_normal((C, S, S))
    traj = rng.uniform(-math.pi, math.pi, (2, K))
    dcf = rng.uniform(0.2, 1.0, K)
    op = KBNufft(maps, traj, np.array([0, K]), dcf, 6, 2.0)
    for r0 in [(S // 2, S // 2), (S // 2 + 5, S // 2 - 7)]:
        x = np.zeros((1, S, S), complex)
        x[0][r0] = 1.0
        y = op.forward(x).reshape(C, K)
        ph = np.exp(-1j * (traj[0] * (r0[0] - S // 2) + traj[1] * (r0[1] - S // 2)))
        exact = maps[:, r0[0], r0[1]][:, None] * ph[None] * np.sqrt(dcf)[None]
        assert rel_l2(y, exact) < 2e-4


def test_deapodisation_is_reciprocal_of_window_transform():
    S, G, W = 64, 128, 6
    beta = kb_beta(W, 2.0)
    d = deapod_1d(S, G, W, beta)
    r = np.arange(S) - S // 2
    np.testing.assert_allclose(d * kb_ft(r / G, W, beta), 1.0, rtol=1e-14)
    assert np.argmin(d) == S // 2          # the window transform peaks at the centre


@pytest.mark.parametrize("mistake", ["swap_axes", "conj_phase", "no_deapod"])
def test_planted_mistakes_fail_the_nudft_pin(mistake):
    """A plausible slip in the oracle (transposed trajectory, wrong sign, dropped term)
    must be caught by the NUDFT comparison that pins it."""
    B, L, C, S, K, maps, traj, dcf, x = reference_precision_problem(4)
    t = flat_traj(traj)
    if mistake == "swap_axes":
        t = t[::-1].copy()
    elif mistake == "conj_phase":
        t = -t
    op = KBNufft(maps, t, np.arange(B + 1) * K, dcf.reshape(-1), 6, 2.0, like=L)
    if mistake == "no_deapod":
        op.deapod[:] = 1.0
    ygt, _ = nudft_pair(maps, traj, dcf, x)
    assert rel_l2(op.forward(x.reshape(B * L, S, S)), ygt) > 1e-1


def test_ragged_frames_match_per_frame_nudft():
    rng = np.random.default_rng(6)
    S, C = 24, 2
    ks = [300, 123, 517]
    off = np.concatenate([[0], np.cumsum(ks)])
    traj = rng.uniform(-math.pi, math.pi, (2, off[-1]))
    dcf = rng.uniform(0.1, 1.0, off[-1])
    maps = rng.standard_normal((C, S, S)) + 1j * rng.standard_normal((C, S, S))
    x = rng.standard_normal((3, S, S)) + 1j * rng.standard_normal((3, S, S))
    op = KBNufft(maps, traj, off, dcf, 6, 2.0)
    y = op.forward(x)
    for b in range(3):
        s = slice(off[b], off[b + 1])
        ref = nudft_forward(maps, x[b:b + 1], traj[:, s], dcf[s])[0]
        assert rel_l2(y[C * off[b]:C * off[b + 1]].reshape(C, -1), ref) < 3e-5


# ------------------------------------------------------------------ 3-D

def _vol_problem(seed=5, S=(12, 10, 8), C=2, K=300):
    rng = np.random.default_rng(seed)
    maps = (rng.standard_normal((C,) + S) + 1j * rng.standard_normal((C,) + S)) * 0.5
    x = rng.standard_normal(S) + 1j * rng.standard_normal(S)
    traj = rng.uniform(-math.pi, math.pi, size=(3, K))
    return maps, x, traj


def test_kb_nufft3_matches_the_nudft():
    """The 3-D restatement equals the FP64 3-D NUDFT to the KB approximation error (W = 6, osf 2),
    and is adjoint to FP64 round-off."""
    from oracle.kbnufft import KBNufft3, nudft3_forward
    maps, x, traj = _vol_problem()
    op = KBNufft3(maps, traj, np.array([0, traj.shape[1]]), None, 6, 2.0)
    y = op.forward(x[None])
    ref = nudft3_forward(maps, x, traj).reshape(-1)
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 2e-3
    yr = np.random.default_rng(9).standard_normal(y.shape) + 0j
    lhs = np.vdot(y, yr)
    rhs = np.vdot(x[None], op.adjoint(yr))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def test_kb_nufft3_reduces_to_2d_on_a_flat_axis():
    """A volume one voxel deep along the last axis with k_z = 0 samples is the 2-D problem: the 3-D
    restatement equals the 2-D one times the z-axis KB factor (deapodisation and window at k = 0)."""
    from oracle.kbnufft import KBNufft, KBNufft3, kb_beta, kb_ft, kb_phi, footprint
    rng = np.random.default_rng(11)
    S, C, K, W = (12, 10), 2, 200, 6
    maps2 = rng.standard_normal((C,) + S) + 1j * rng.standard_normal((C,) + S)
    x2 = rng.standard_normal(S) + 1j * rng.standard_normal(S)
    t2 = rng.uniform(-math.pi, math.pi, size=(2, K))
    y2 = KBNufft(maps2, t2, np.array([0, K]), None, W, 2.0).forward(x2[None])
    maps3 = maps2[..., None]
    t3 = np.concatenate([t2, np.zeros((1, K))])
    y3 = KBNufft3(np.repeat(maps3, 1, axis=3), t3, np.array([0, K]), None, W, 2.0).forward(x2[..., None][None])
    beta = kb_beta(W, 2.0)
    # z: S = 1 (deapod 1 / Phi(0)), G = 2, footprint of u = 0 over nodes -2..3 folded onto 2 grid points
    m0, w = footprint(np.zeros(1), W, beta)
    fz = (1.0 / kb_ft(0.0, W, beta)) * w.sum()
    assert np.linalg.norm(y3 - fz * y2) / np.linalg.norm(y2) < 1e-12


@pytest.mark.parametrize("mode,fwd,adj", [("none", False, True), ("standard", False, True),
                                          ("standard", True, True), ("standard", True, False),
                                          ("standard", False, False)])
def test_dcf_policy_matches_the_weighted_nudft(mode, fwd, adj):
    """DCF policy of torchkb.py:185-199: 'none' weights neither side; 'standard' multiplies by
    dcf (not its square root) on the sides apply_dcf_in_fwd / apply_dcf_in_adj name.  Pinned by
    the unweighted NUDFT with the weights applied by hand (W = 6 gridding error <= 3e-5)."""
    B, L, C, S, K, maps, traj, dcf, x = reference_precision_problem(2)
    op = KBNufft(maps, flat_traj(traj), np.arange(B + 1) * K, dcf.reshape(-1), 6, 2.0, like=L,
                 dcf_mode=mode, dcf_in_fwd=fwd, dcf_in_adj=adj)
    use_f = mode == "standard" and fwd
    use_a = mode == "standard" and adj
    y_un = np.stack([nudft_forward(maps, x[b], traj[b]) for b in range(B)])          # (B, L, C, K)
    ygt = y_un * (dcf[:, None, None, :] if use_f else 1.0)
    rng = np.random.default_rng(3)
    yr = rng.standard_normal(y_un.shape) + 1j * rng.standard_normal(y_un.shape)
    xgt = np.stack([nudft_adjoint(maps, yr[b] * (dcf[b][None, None] if use_a else 1.0), traj[b])
                    for b in range(B)])
    assert rel_l2(op.forward(x.reshape(B * L, S, S)), ygt) <= 3e-5
    assert rel_l2(op.adjoint(yr.reshape(-1)), xgt) <= 3e-5
    # planted mistake: the balanced square root on a 'standard' side fails the pin
    if use_f:
        bad = y_un * np.sqrt(dcf)[:, None, None, :]
        assert rel_l2(op.forward(x.reshape(B * L, S, S)), bad) > 1e-2
