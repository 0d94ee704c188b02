"""3-D NUFFT (ndim=3) on the GPU against the FP64 oracle (oracle/kbnufft.py KBNufft3).

Bar: 1e-5 relative L2 for the forward and adjoint (complex64 vs FP64), adjointness
<Ax, y> = <x, A^H y> to 1e-5, the reference adapter surface (shapes, errors)."""
import math

import numpy as np
import pytest
import torch

from mrirecon_amd.nufft.backends.hip import HipNUFFTAdapter
from oracle.kbnufft import KBNufft3, nudft3_forward, rel_l2

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def problem(S, C, Ks, seed, W=6, osf=2.0):
    rng = np.random.default_rng(seed)
    maps = ((rng.standard_normal((C,) + S) + 1j * rng.standard_normal((C,) + S)) * 0.5).astype(np.complex64)
    off = np.concatenate([[0], np.cumsum(Ks)]).astype(np.int64)
    traj = rng.uniform(-math.pi, math.pi, size=(3, int(off[-1]))).astype(np.float32)
    dcf = rng.uniform(0.2, 1.5, size=int(off[-1])).astype(np.float32)
    x = (rng.standard_normal((len(Ks),) + S) + 1j * rng.standard_normal((len(Ks),) + S)).astype(np.complex64)
    be = HipNUFFTAdapter(ndim=3, traj_units="rad", kb_width=W, osf=osf)
    be.prepare((C,) + S, torch.from_numpy(traj), torch.from_numpy(dcf), torch.from_numpy(maps).to(DEV),
               device=DEV, k_offsets=off)
    ora = KBNufft3(maps, traj, off, dcf, W, osf)
    return be, ora, x, off


@pytest.mark.parametrize("S,C,Ks,W,osf", [((16, 16, 16), 3, [400, 700], 6, 2.0),
                                          ((16, 32, 48), 2, [300, 0, 500], 6, 2.0),
                                          ((16, 16, 32), 2, [600], 4, 2.0),
                                          ((25, 51, 51), 2, [400, 300], 6, 1.25)])
def test_nufft3_forward_adjoint_match_the_oracle(S, C, Ks, W, osf):
    """Odd image sizes (osf 1.25: 25 x 51 x 51 on a 32 x 64 x 64 grid) put S - S/2 and S/2 image rows
    on the two sides of the grid origin, the case the row-pruned FFT passes must get right."""
    be, ora, x, off = problem(S, C, Ks, seed=sum(S) + C, W=W, osf=osf)
    xg = torch.from_numpy(x).to(DEV).reshape(len(Ks), 1, *S)
    y = be.A(xg)
    torch.cuda.synchronize()
    y_ref = ora.forward(x.astype(np.complex128))
    ef = rel_l2(y.cpu().numpy().reshape(-1), y_ref)
    yr = (np.random.default_rng(3).standard_normal(y_ref.shape)
          + 1j * np.random.default_rng(4).standard_normal(y_ref.shape)).astype(np.complex64)
    xa = be.AH(torch.from_numpy(yr).to(DEV))
    torch.cuda.synchronize()
    xa_ref = ora.adjoint(yr.astype(np.complex128))
    ea = rel_l2(xa.cpu().numpy().reshape(len(Ks), *S), xa_ref)
    print(f"[3-D {S} C={C} K={Ks} W={W} osf={osf}] forward rel={ef:.2e} adjoint rel={ea:.2e}")
    assert ef <= 1e-5 and ea <= 1e-5
    if 0 in Ks:                                   # an empty frame adjoints to zeros
        b = Ks.index(0)
        assert float(xa[b].abs().max()) == 0.0


def test_nufft3_adjointness_and_nudft():
    """<A x, y> = <x, A^H y> in complex64, and the GPU forward within the KB error of the FP64 NUDFT."""
    S, C, Ks = (16, 16, 16), 2, [900]
    be, ora, x, off = problem(S, C, Ks, seed=77)
    xg = torch.from_numpy(x).to(DEV).reshape(1, 1, *S)
    y = be.A(xg)
    yr = torch.randn(y.shape, dtype=torch.complex64, device=DEV)
    lhs = torch.vdot(y.reshape(-1), yr.reshape(-1)).item()
    rhs = torch.vdot(xg.reshape(-1), be.AH(yr).reshape(-1)).item()
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
    be2 = HipNUFFTAdapter(ndim=3, traj_units="rad", kb_width=6, osf=2.0, dcf_mode="none")
    rng = np.random.default_rng(78)
    maps = ora.maps.astype(np.complex64)
    traj = rng.uniform(-math.pi, math.pi, size=(3, 500)).astype(np.float32)
    be2.prepare((C,) + S, torch.from_numpy(traj), None, torch.from_numpy(maps).to(DEV), device=DEV)
    yg = be2.A(xg).cpu().numpy().reshape(C, -1)
    ref = nudft3_forward(maps, x[0], traj)
    assert rel_l2(yg, ref) <= 2e-3


def test_nufft3_adapter_surface():
    """Uniform trajectories give (B, C, K) k-space and (B, 1, X, Y, Z) images; wrong shapes are refused."""
    S, C, B, K = (16, 16, 16), 2, 2, 256
    rng = np.random.default_rng(79)
    maps = torch.from_numpy((rng.standard_normal((C,) + S) + 0j).astype(np.complex64)).to(DEV)
    traj = torch.from_numpy(rng.uniform(-math.pi, math.pi, size=(B, 3, K)).astype(np.float32))
    be = HipNUFFTAdapter(ndim=3)
    be.prepare((C,) + S, traj, None, maps, device=DEV)
    x = torch.ones((B, 1) + S, dtype=torch.complex64, device=DEV)
    y = be.A(x)
    assert tuple(y.shape) == (B, C, K)
    assert tuple(be.AH(y).shape) == (B, 1) + S
    assert be.k_per_frame() == K
    with pytest.raises(ValueError):
        be.A(torch.ones((B, 1, 16, 16), dtype=torch.complex64, device=DEV))
    with pytest.raises(ValueError):
        be.AH(torch.ones((B, C, K + 1), dtype=torch.complex64, device=DEV))
    assert be.diag_AHA_scalar() > 0


def test_nufft3_kooshball_centre_lists():
    """A kooshball puts every spoke through the k-space centre, so the centre grid points get long
    footprint lists (over a thousand entries here) that the adjoint splits into parts and sums in
    order: the adjoint still matches the oracle, and two runs are bitwise equal."""
    S, C, spokes, R = (16, 16, 16), 3, 240, 32
    i = np.arange(spokes) + 0.5
    phi = np.arccos(1.0 - 2.0 * i / spokes)
    theta = math.pi * (1.0 + 5.0 ** 0.5) * i
    dirs = np.stack([np.sin(phi) * np.cos(theta), np.sin(phi) * np.sin(theta), np.cos(phi)])
    rad = (np.arange(R) - R / 2) / (R / 2) * math.pi * 0.999
    traj = (dirs[:, :, None] * rad[None, None, :]).reshape(3, -1).astype(np.float32)
    rng = np.random.default_rng(80)
    maps = ((rng.standard_normal((C,) + S) + 1j * rng.standard_normal((C,) + S)) * 0.5).astype(np.complex64)
    be = HipNUFFTAdapter(ndim=3, traj_units="rad", kb_width=6, osf=2.0, dcf_mode="none")
    be.prepare((C,) + S, torch.from_numpy(traj), None, torch.from_numpy(maps).to(DEV), device=DEV)
    ora = KBNufft3(maps, traj, np.array([0, traj.shape[1]]), None, 6, 2.0)
    centre_entries = max(len(np.nonzero(f == 0)[0]) for f, _ in [ora._fp[0]])
    assert centre_entries > 64                    # the split path runs
    yr = (rng.standard_normal((1, C, traj.shape[1])) + 1j * rng.standard_normal((1, C, traj.shape[1]))).astype(
        np.complex64)
    yg = torch.from_numpy(yr).to(DEV)
    xa = be.AH(yg)
    xb = be.AH(yg)
    torch.cuda.synchronize()
    ea = rel_l2(xa.cpu().numpy().reshape(S), ora.adjoint(yr.astype(np.complex128).reshape(-1))[0])
    print(f"[3-D kooshball] centre grid point entries {centre_entries}, adjoint rel={ea:.2e}")
    assert ea <= 1e-5
    assert torch.equal(xa, xb)


def test_nufft3_through_the_reference_front_end():
    """NUFFT(maps (C, X, Y, Z), traj, dcf, axis, NUFFTConfig(ndim=3)) -- the reference's public entry
    (graspcg/nufft/api.py:85-89) -- in a user axis order equals the adapter's operator."""
    from mrirecon_amd.nufft.api import NUFFT, NUFFTConfig
    from mrirecon_amd.nufft.layout import AxisSpec
    S, C, K = (16, 16, 32), 2, 500
    be, ora, x, off = problem(S, C, [K], seed=81)
    traj = torch.from_numpy(be._traj_flat.reshape(1, 3, K).copy()).to(DEV)
    dcf = torch.from_numpy(be._dcf_flat.reshape(1, K).copy()).to(DEV)
    axis = AxisSpec(image=("B", "coil", "X", "Y", "Z"), kspace=("B", "K", "coil"), image_fft=("X", "Y", "Z"))
    op = NUFFT(be._maps, traj, dcf, axis=axis, config=NUFFTConfig(ndim=3, traj_units="rad"), device=DEV)
    xg = torch.from_numpy(x).to(DEV).reshape(1, 1, *S)
    yu = op.A(xg)                                             # (1, K, C) user order
    assert tuple(yu.shape) == (1, K, C)
    y1 = be.A(xg)
    assert torch.equal(yu.permute(0, 2, 1).contiguous(), y1)
    xb = op.AH(yu)
    torch.cuda.synchronize()
    assert tuple(xb.shape) == (1, 1) + S
    assert rel_l2(xb.cpu().numpy(), be.AH(y1).cpu().numpy()) <= 1e-6
