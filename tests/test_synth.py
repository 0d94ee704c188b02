"""Invariants of the seeded synthetic inputs (no GPU): coil maps with sum-of-squares 1, the
golden-angle spoke sequence across frames, the readout grid, the radial DCF, seed
reproducibility and the uneven spoke draw of the 48-slice batched workload."""
import math

import numpy as np
import pytest

from mrirecon_amd import synth


@pytest.mark.parametrize("C,S", [(8, 128), (12, 256), (16, 384), (3, 31)])
def test_coil_maps_have_unit_sum_of_squares(C, S):
    m = synth.gaussian_maps(C, S)
    assert m.shape == (C, S, S) and m.dtype == np.complex64
    sos = np.sqrt((np.abs(m.astype(np.complex128)) ** 2).sum(0))
    assert np.abs(sos - 1.0).max() <= 5e-7
    # the phase of coil c is the constant 2 pi c / C
    for c in range(C):
        ph = np.angle(m[c].astype(np.complex128) * np.exp(-2j * math.pi * c / C))
        assert np.abs(ph).max() <= 1e-6


def test_golden_angle_spokes_continue_across_frames():
    R = 64
    spokes = [5, 3, 7]
    traj, dcf, off = synth.radial_frames(spokes, R)
    assert list(off) == [0, 5 * R, 8 * R, 15 * R]
    t = traj.astype(np.float64).reshape(2, -1, R)        # (2, spokes, R)
    ang = np.arctan2(t[1, :, -1], t[0, :, -1])            # direction of the positive end
    inc = np.mod(np.diff(ang), 2 * math.pi)
    assert np.allclose(inc, math.fmod(synth.GOLDEN_ANGLE, 2 * math.pi), atol=2e-6)
    assert abs(math.degrees(synth.GOLDEN_ANGLE) - 111.2461) < 1e-4
    # readout grid w_j = pi (2 j - R) / R along each spoke, |w| <= pi
    wr = math.pi * (2 * np.arange(R) - R) / R
    rad = np.sign(t[0] * np.cos(ang)[:, None] + t[1] * np.sin(ang)[:, None]) * np.hypot(t[0], t[1])
    assert np.allclose(rad, wr[None, :], atol=2e-6)
    assert np.abs(traj).max() <= math.pi
    # DCF: radial ramp max(|w|, pi / 2R) / (2 R N_spokes) per frame
    for b, ns in enumerate(spokes):
        d = dcf[off[b]:off[b + 1]].reshape(ns, R)
        ref = np.maximum(np.abs(wr), math.pi / (2 * R)) / (2 * R * ns)
        assert np.allclose(d, ref[None, :], rtol=1e-6)
    # first_spoke shifts the sequence
    t2, _, _ = synth.radial_frames([3], R, first_spoke=5)
    assert np.array_equal(t2, traj[:, 5 * R:8 * R])


def test_uniform_layout_matches_the_ragged_one():
    traj, dcf = synth.radial_uniform(4, 6, 32)
    tr, dr, off = synth.radial_frames([6] * 4, 32)
    assert traj.shape == (4, 2, 6 * 32) and dcf.shape == (4, 6 * 32)
    for b in range(4):
        assert np.array_equal(traj[b], tr[:, off[b]:off[b + 1]])
        assert np.array_equal(dcf[b], dr[off[b]:off[b + 1]])


def test_seeds_reproduce_and_differ():
    a, b = synth.dynamic_phantom(64, 6, seed=5), synth.dynamic_phantom(64, 6, seed=5)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, synth.dynamic_phantom(64, 6, seed=6))
    s1, s2 = synth.shepp_logan(64, 101), synth.shepp_logan(64, 101)
    assert np.array_equal(s1, s2) and not np.array_equal(s1, synth.shepp_logan(64, 102))
    # the phase only: magnitudes do not depend on the seed
    assert np.allclose(np.abs(s1), np.abs(synth.shepp_logan(64, 7)), atol=1e-6)
    n1, n2 = synth.complex_noise((4000,), 0.5, 3), synth.complex_noise((4000,), 0.5, 3)
    assert np.array_equal(n1, n2)
    assert abs(np.mean(np.abs(n1.astype(np.complex128)) ** 2) - 0.25) < 0.02   # E|n|^2 = sigma^2
    w1, w2 = synth.config_c1(), synth.config_c1()
    for f in ("traj", "dcf", "k_offsets", "maps", "x_true"):
        assert np.array_equal(getattr(w1, f), getattr(w2, f))


def test_shepp_logan_intensities():
    """Toft's modified Shepp-Logan: background 0, the skull ring 1.0, the brain 0.2 and the
    small features 0.1 above or below it; every magnitude is one of those sums."""
    mag = np.abs(synth.shepp_logan(256, 101).astype(np.complex128))
    vals = np.unique(np.round(mag, 4))
    allowed = {0.0, 0.1, 0.2, 0.3, 0.4, 1.0}
    assert set(vals.tolist()) <= allowed
    assert (np.isclose(mag, 0.2, atol=1e-4).mean() > 0.25) and (np.isclose(mag, 1.0, atol=1e-4).mean() > 0.02)


def test_batched_uneven_spoke_draw():
    """48 slices x 20 frames, spokes per frame drawn from U{13..55} with seed 401."""
    wl = synth.config_c4()
    sp = np.asarray(wl.spokes)
    assert sp.shape == (960,) and wl.B == 960
    assert np.array_equal(sp, np.random.default_rng(401).integers(13, 56, size=960))
    assert sp.min() == 13 and sp.max() == 55 and len(np.unique(sp)) == 43
    assert abs(sp.mean() - 34.0) < 1.5                     # E = 34 spokes (17 408 samples per frame)
    assert list(np.diff(wl.k_offsets)) == list(sp * 512)
    assert wl.x_true.shape == (960, 256, 256)
    # slice s uses the phantom seed 402 + s
    assert np.array_equal(wl.x_true[20:40], synth.dynamic_phantom(256, 20, seed=403))
    assert wl.uniform_K is None and synth.config_c2().uniform_K == 21 * 512
