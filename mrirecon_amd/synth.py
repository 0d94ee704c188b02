"""Seeded synthetic workloads for golden-angle radial MRI.

The reference library takes maps / trajectory / DCF / k-space data as inputs
(graspcg/nufft/api.py:74-84) and ships no generator, so the inputs of every
workload here are produced by this module, with numpy PCG64 generators so that
the CPU and GPU sides see bit-identical arrays.

Conventions (shared by the GPU path and the oracle):

* trajectory in radians; axis 0 pairs with the first image axis (X, rows),
  axis 1 with the second (Y, columns); NUDFT phase exp(-i w.r) with centred
  r = (ix - S//2, iy - S//2)  (graspcg/tests/test_nufft_precision_gt.py:48-53);
* golden-angle radial: spoke n has angle n * pi*(sqrt(5)-1)/2 and the spoke
  sequence continues across frames; readout j = 0..R-1 sits at
  w_r = pi*(2j - R)/R;
* DCF: analytic radial ramp normalised so that A^H A ~ I on band-limited
  images, dcf = max(|w_r|, pi/(2R)) / (2 R N_spokes) with N_spokes the number of
  spokes of that frame;
* coil maps: Gaussian magnitude exp(-|r - r_c|^2 / (2 sigma^2)) with centres on a
  circle of radius S/4 and sigma = 0.2 S, a constant phase 2 pi c / C per coil,
  normalised to sum-of-squares 1.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

GOLDEN_ANGLE = math.pi * (math.sqrt(5.0) - 1.0) / 2.0   # ~111.246 degrees


# ----------------------------------------------------------------- trajectory

def radial_frames(spokes_per_frame: Sequence[int], n_read: int, first_spoke: int = 0
                  ) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Golden-angle radial sampling for frames with (possibly uneven) spoke counts.

    Returns ``traj`` (2, sum K_b) float32 radians, ``dcf`` (sum K_b,) float32 and
    ``k_offsets`` (B+1,) int64, where frame b owns samples
    ``k_offsets[b]:k_offsets[b+1]`` (spoke-major, readout-minor).
    """
    R = int(n_read)
    j = np.arange(R, dtype=np.float64)
    wr = math.pi * (2.0 * j - R) / R
    trajs, dcfs, offs = [], [], [0]
    n = int(first_spoke)
    for ns in spokes_per_frame:
        ns = int(ns)
        ang = (n + np.arange(ns, dtype=np.float64)) * GOLDEN_ANGLE
        n += ns
        wx = (np.cos(ang)[:, None] * wr[None, :]).reshape(-1)
        wy = (np.sin(ang)[:, None] * wr[None, :]).reshape(-1)
        trajs.append(np.stack([wx, wy]))
        d = np.maximum(np.abs(wr), math.pi / (2.0 * R)) / (2.0 * R * ns)
        dcfs.append(np.tile(d, ns))
        offs.append(offs[-1] + ns * R)
    traj = np.concatenate(trajs, axis=1).astype(np.float32)
    dcf = np.concatenate(dcfs).astype(np.float32)
    return traj, dcf, np.asarray(offs, dtype=np.int64)


def radial_uniform(n_frames: int, spokes: int, n_read: int) -> Tuple[np.ndarray, np.ndarray]:
    """Uniform-K golden-angle radial: traj (B, 2, K) float32, dcf (B, K) float32."""
    traj, dcf, _ = radial_frames([spokes] * n_frames, n_read)
    K = spokes * n_read
    traj = traj.reshape(2, n_frames, K).transpose(1, 0, 2).copy()
    return traj, dcf.reshape(n_frames, K).copy()


# ----------------------------------------------------------------- coil maps

def gaussian_maps(n_coils: int, S: int) -> np.ndarray:
    """(C, S, S) complex64 Gaussian sensitivities, sum-of-squares 1."""
    r = np.arange(S, dtype=np.float64) - S // 2
    X, Y = np.meshgrid(r, r, indexing="ij")
    rad, sig = 0.5 * S / 2.0, 0.4 * S / 2.0
    maps = np.empty((n_coils, S, S), dtype=np.complex128)
    for c in range(n_coils):
        a = 2.0 * math.pi * c / n_coils
        cx, cy = rad * math.cos(a), rad * math.sin(a)
        mag = np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2.0 * sig * sig))
        maps[c] = mag * np.exp(1j * a)
    sos = np.sqrt(np.sum(np.abs(maps) ** 2, axis=0))
    return (maps / sos[None]).astype(np.complex64)


# ----------------------------------------------------------------- phantoms

# Modified Shepp-Logan ellipses (Toft): intensity, semi-axes a, b, centre x0, y0, angle (deg)
_SHEPP_LOGAN = [
    (1.00, 0.6900, 0.9200, 0.00, 0.0000, 0.0),
    (-0.80, 0.6624, 0.8740, 0.00, -0.0184, 0.0),
    (-0.20, 0.1100, 0.3100, 0.22, 0.0000, -18.0),
    (-0.20, 0.1600, 0.4100, -0.22, 0.0000, 18.0),
    (0.10, 0.2100, 0.2500, 0.00, 0.3500, 0.0),
    (0.10, 0.0460, 0.0460, 0.00, 0.1000, 0.0),
    (0.10, 0.0460, 0.0460, 0.00, -0.1000, 0.0),
    (0.10, 0.0460, 0.0230, -0.08, -0.6050, 0.0),
    (0.10, 0.0230, 0.0230, 0.00, -0.6060, 0.0),
    (0.10, 0.0230, 0.0460, 0.06, -0.6050, 0.0),
]


def _unit_coords(S: int) -> Tuple[np.ndarray, np.ndarray]:
    r = (np.arange(S, dtype=np.float64) - S // 2 + 0.5) / (S / 2.0)
    yy, xx = np.meshgrid(-r, r, indexing="ij")   # rows go down, columns go right
    return xx, yy


def _ellipse_mask(xx, yy, a, b, x0, y0, deg):
    th = math.radians(deg)
    c, s = math.cos(th), math.sin(th)
    xr = (xx - x0) * c + (yy - y0) * s
    yr = -(xx - x0) * s + (yy - y0) * c
    return (xr / a) ** 2 + (yr / b) ** 2 <= 1.0


def _smooth_phase(S: int, rng: np.random.Generator) -> np.ndarray:
    xx, yy = _unit_coords(S)
    coef = rng.normal(0.0, math.pi / 4.0, size=6)
    return (coef[0] + coef[1] * xx + coef[2] * yy + coef[3] * xx * yy
            + coef[4] * xx * xx + coef[5] * yy * yy)


def shepp_logan(S: int, seed: int) -> np.ndarray:
    """(S, S) complex64 modified Shepp-Logan with a seeded smooth random phase."""
    xx, yy = _unit_coords(S)
    img = np.zeros((S, S), dtype=np.float64)
    for (A, a, b, x0, y0, deg) in _SHEPP_LOGAN:
        img[_ellipse_mask(xx, yy, a, b, x0, y0, deg)] += A
    rng = np.random.default_rng(seed)
    return (img * np.exp(1j * _smooth_phase(S, rng))).astype(np.complex64)


def dynamic_phantom(S: int, n_frames: int, seed: int, n_ellipses: int = 20) -> np.ndarray:
    """(F, S, S) complex64: random ellipses with sinusoid / gamma-variate time curves."""
    rng = np.random.default_rng(seed)
    xx, yy = _unit_coords(S)
    t = np.arange(n_frames, dtype=np.float64)
    out = np.zeros((n_frames, S, S), dtype=np.float64)
    for _ in range(n_ellipses):
        rad = 0.7 * math.sqrt(rng.uniform())
        ang = rng.uniform(0.0, 2.0 * math.pi)
        x0, y0 = rad * math.cos(ang), rad * math.sin(ang)
        a, b = rng.uniform(0.05, 0.3, size=2)
        deg = rng.uniform(0.0, 180.0)
        mask = _ellipse_mask(xx, yy, a, b, x0, y0, deg).astype(np.float64)
        if rng.uniform() < 0.5:
            base = rng.uniform(0.2, 1.0)
            amp = rng.uniform(0.0, 0.5) * base
            f = rng.uniform(0.5, 3.0)
            ph = rng.uniform(0.0, 2.0 * math.pi)
            curve = base + amp * np.sin(2.0 * math.pi * f * t / n_frames + ph)
        else:
            t0 = rng.uniform(0.0, 0.4) * n_frames
            alpha = rng.uniform(1.5, 4.0)
            beta = rng.uniform(0.05, 0.2) * n_frames
            tt = np.maximum(t - t0, 0.0)
            g = tt ** alpha * np.exp(-tt / beta)
            peak = (alpha * beta) ** alpha * math.exp(-alpha)
            curve = 0.1 + rng.uniform(0.5, 1.0) * g / peak
        out += curve[:, None, None] * mask[None]
    phase = _smooth_phase(S, rng)
    return (out * np.exp(1j * phase)[None]).astype(np.complex64)


def complex_noise(shape, sigma: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    n = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return (sigma / math.sqrt(2.0) * n).astype(np.complex64)


# ----------------------------------------------------------------- workloads

@dataclass
class Workload:
    """Inputs of one reconstruction problem set (all numpy, CPU).

    ``x_true`` is (B, S, S) with B = n_slices * n_frames; ``traj`` is (2, sum K);
    ``k_offsets`` (B+1,) delimits each frame's samples; ``y`` is filled by
    :func:`measure` as the concatenation over frames of (C, K_b) blocks.
    """
    name: str
    S: int
    n_slices: int
    n_frames: int
    n_coils: int
    n_read: int
    W: int
    osf: float
    spokes: List[int]
    maps: np.ndarray
    traj: np.ndarray
    dcf: np.ndarray
    k_offsets: np.ndarray
    x_true: np.ndarray
    lam_t: float = 0.0
    lam_s: float = 0.0
    n_iter: int = 20
    noise_seed: int = 0
    y: Optional[np.ndarray] = None

    @property
    def B(self) -> int:
        return self.n_slices * self.n_frames

    @property
    def uniform_K(self) -> Optional[int]:
        k = np.diff(self.k_offsets)
        return int(k[0]) if np.all(k == k[0]) else None


def config_c1() -> Workload:
    """128^2, 8 coils, 64 golden-angle spokes x 256 samples, W=4, osf=2, spatial TV 0.01."""
    S, C, R, ns = 128, 8, 256, 64
    traj, dcf, off = radial_frames([ns], R)
    x = shepp_logan(S, seed=101)[None]
    return Workload("c1", S, 1, 1, C, R, 4, 2.0, [ns], gaussian_maps(C, S), traj, dcf, off,
                    x, lam_t=0.0, lam_s=0.01, n_iter=20, noise_seed=102)


def config_c2(seed: int = 201, noise_seed: int = 202) -> Workload:
    """GRASP: 256^2 x 34 frames, 12 coils, 21 spokes/frame x 512, W=6, temporal TV 0.02 + spatial 0.002."""
    S, F, C, R, ns = 256, 34, 12, 512, 21
    traj, dcf, off = radial_frames([ns] * F, R)
    x = dynamic_phantom(S, F, seed=seed)
    return Workload("c2", S, 1, F, C, R, 6, 2.0, [ns] * F, gaussian_maps(C, S), traj, dcf, off,
                    x, lam_t=0.02, lam_s=0.002, n_iter=30, noise_seed=noise_seed)


def config_c3() -> Workload:
    """Precision study: 384^2, 16 coils, 402 spokes x 768, W=6; x ~ CN(0,1) (seed 301)."""
    S, C, R, ns = 384, 16, 768, 402
    traj, dcf, off = radial_frames([ns], R)
    rng = np.random.default_rng(301)
    x = (rng.standard_normal((1, S, S)) + 1j * rng.standard_normal((1, S, S))).astype(np.complex64)
    return Workload("c3", S, 1, 1, C, R, 6, 2.0, [ns], gaussian_maps(C, S), traj, dcf, off,
                    x, n_iter=20, noise_seed=302)


def config_c4(n_slices: int = 48, n_frames: int = 20, seed: int = 401) -> Workload:
    """Batched uneven: n_slices x n_frames at 256^2, 8 coils, spokes/frame ~ U{13..55} x 512, W=6."""
    S, C, R = 256, 8, 512
    rng = np.random.default_rng(seed)
    spokes = [int(v) for v in rng.integers(13, 56, size=n_slices * n_frames)]
    traj, dcf, off = radial_frames(spokes, R)
    x = np.concatenate([dynamic_phantom(S, n_frames, seed=402 + s) for s in range(n_slices)])
    return Workload("c4", S, n_slices, n_frames, C, R, 6, 2.0, spokes, gaussian_maps(C, S),
                    traj, dcf, off, x, lam_t=0.02, lam_s=0.0, n_iter=60, noise_seed=403)


def measure(wl: Workload, forward, noise_frac: float = 0.01) -> np.ndarray:
    """y = A x_true + complex Gaussian noise with sigma = noise_frac * max|A x_true|.

    ``forward`` maps (B, S, S) complex64 to the flat k-space vector (sum_b C*K_b,)."""
    y = np.asarray(forward(wl.x_true)).astype(np.complex64).reshape(-1)
    sigma = noise_frac * float(np.abs(y).max())
    wl.y = (y + complex_noise(y.shape, sigma, wl.noise_seed)).astype(np.complex64)
    return wl.y
