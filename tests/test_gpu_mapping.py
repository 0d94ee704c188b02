"""Temporal-basis kernels (csrc/tbasis.hip) and TV through them vs the oracle (oracle/mapping.py)."""
import numpy as np
import pytest
import torch

from oracle.mapping import (mapped_tv_add_diag, mapped_tv_energy_grad, tb_adjoint, tb_diag_push,
                            tb_diag_push_profile, tb_forward)
from oracle.nlcg import TV

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _c(rng, *shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30))


@pytest.mark.parametrize("T,K,shape", [(34, 4, (64, 64)), (5, 1, (3, 7)), (20, 16, (1000,)), (3, 7, (1, 9, 11)),
                                       (40, 6, (333,))])
def test_forward_adjoint_diag_match_the_oracle(T, K, shape):
    from mrirecon_amd.mapping import TemporalBasisOp
    rng = np.random.default_rng(T * 100 + K)
    U = _c(rng, T, K)
    V, W = _c(rng, K, *shape), _c(rng, T, *shape)
    ku = rng.random((T,) + shape).astype(np.float32)
    v = rng.random(T).astype(np.float32)
    D0 = rng.random((K,) + shape).astype(np.float32)
    print(f"[tbasis] T={T} K={K} shape={shape}", flush=True)
    op = TemporalBasisOp(torch.from_numpy(U))
    u = op.forward_apply(torch.from_numpy(V).to(DEV)).cpu().numpy()
    gV0 = _c(rng, K, *shape)
    gV = torch.from_numpy(gV0).to(DEV)
    op.adjoint_apply(torch.from_numpy(W).to(DEV), gV)
    gV_set = op.adjoint_apply(torch.from_numpy(W).to(DEV), torch.empty_like(gV), accumulate=False)
    D = torch.from_numpy(D0).to(DEV)
    op.diag_push(torch.from_numpy(ku).to(DEV), D)
    Dp = torch.from_numpy(D0).to(DEV)
    op.diag_push_profile(0, torch.from_numpy(v), Dp)
    U64, V64, W64 = U.astype(np.complex128), V.astype(np.complex128), W.astype(np.complex128)
    assert u.shape == (T,) + shape and _rel(u, tb_forward(U64, V64)) < 1e-6
    assert _rel(gV.cpu().numpy(), tb_adjoint(U64, W64, gV0.astype(np.complex128))) < 1e-6
    assert _rel(gV_set.cpu().numpy(), tb_adjoint(U64, W64)) < 1e-6
    assert _rel(D.cpu().numpy(), tb_diag_push(U64, ku.astype(np.float64), D0.astype(np.float64))) < 1e-6
    assert _rel(Dp.cpu().numpy(), tb_diag_push_profile(U64, v, D0.astype(np.float64))) < 1e-6
    # fixed-order sums: a repeat is bitwise equal
    u2 = op.forward_apply(torch.from_numpy(V).to(DEV)).cpu().numpy()
    assert np.array_equal(u, u2)
    # a profile axis other than frames is dropped (mapping.py:209-211)
    Dq = torch.from_numpy(D0).to(DEV)
    assert op.diag_push_profile(1, torch.ones(T), Dq) is None and torch.equal(Dq.cpu(), torch.from_numpy(D0))


def test_adjoint_identity_on_the_device():
    from mrirecon_amd.mapping import TemporalBasisOp
    rng = np.random.default_rng(5)
    T, K = 34, 8
    op = TemporalBasisOp(torch.from_numpy(_c(rng, T, K)))
    V, W = torch.from_numpy(_c(rng, K, 48, 48)).to(DEV), torch.from_numpy(_c(rng, T, 48, 48)).to(DEV)
    lhs = torch.vdot(W.reshape(-1).to(torch.complex128), op.forward_apply(V).reshape(-1).to(torch.complex128))
    gV = op.adjoint_apply(W, torch.empty_like(V), accumulate=False)
    rhs = torch.vdot(gV.reshape(-1).to(torch.complex128), V.reshape(-1).to(torch.complex128))
    assert abs(complex(lhs - rhs)) <= 1e-5 * abs(complex(lhs))


def test_reference_diag_push_values():
    """graspcg/tests/test_reg_diag_mapped.py:92-126 through MappedTV.add_diag."""
    from mrirecon_amd import TVND, TVParams
    from mrirecon_amd.mapping import MappedTV, TemporalBasisOp
    U = torch.tensor([[1.0, 0.0, 0.0], [0.0, 2.0, 0.0]])
    mt = MappedTV([TVND("tv", TVParams(weight=1e-3, eps=1e-3, axes="spatial", isotropic=True))], TemporalBasisOp(U))
    diag_V = torch.zeros((3, 1, 3, 2), dtype=torch.float32, device=DEV)
    mt.add_diag(diag_V)
    k = 2.0 * 1e-3 * 2.0
    d = diag_V.cpu()
    assert torch.allclose(d[0], torch.full((1, 3, 2), k), atol=1e-7, rtol=0)
    assert torch.allclose(d[1], torch.full((1, 3, 2), 4 * k), atol=1e-7, rtol=0)
    assert torch.all(d[2] == 0)


@pytest.mark.parametrize("image_diag", [False, True])
def test_mapped_tv_energy_gradient_and_diag_match_the_oracle(image_diag):
    from mrirecon_amd import TVND, TVParams
    from mrirecon_amd.mapping import MappedTV, TemporalBasisOp
    rng = np.random.default_rng(7)
    T, K, S = 20, 5, 40
    U = _c(rng, T, K) / np.float32(np.sqrt(K))
    V = _c(rng, K, S, S)
    terms = [TVND("t", TVParams(weight=0.02, eps=1e-3, axes="temporal")),
             TVND("s", TVParams(weight=0.005, eps=1e-3, axes="spatial", isotropic=False, axis_weights=(1.0, 0.5)))]
    mt = MappedTV(terms, TemporalBasisOp(torch.from_numpy(U)))
    g0 = _c(rng, K, S, S)
    gV = torch.from_numpy(g0).to(DEV)
    E = float(mt.energy_grad(torch.from_numpy(V).to(DEV), gV))
    regs = [TV(weight=0.02, axes=(0,), eps=1e-3), TV(weight=0.005, axes=(1, 2), eps=1e-3, isotropic=False,
                                                     axis_weights=(1.0, 0.5))]
    E_ref, g_ref = mapped_tv_energy_grad(U.astype(np.complex128), V.astype(np.complex128), regs,
                                         g0.astype(np.complex128))
    print(f"[mapped tv] E={E:.9g} oracle {E_ref:.9g}", flush=True)
    assert abs(E - E_ref) <= 1e-5 * abs(E_ref)
    assert _rel(gV.cpu().numpy(), g_ref) < 1e-5
    D = torch.zeros((K, S, S), dtype=torch.float32, device=DEV)
    mt.add_diag(D, image_diag=image_diag)
    D_ref = np.zeros((K, S, S))
    for r in regs:
        mapped_tv_add_diag(U.astype(np.complex128), r, D_ref, image_diag=image_diag)
    assert _rel(D.cpu().numpy(), D_ref) < 1e-6


def test_limits():
    from mrirecon_amd.mapping import TemporalBasisOp
    with pytest.raises(ValueError):
        TemporalBasisOp(torch.ones(4, 17))
    with pytest.raises(ValueError):
        TemporalBasisOp(torch.ones(0, 2))
    op = TemporalBasisOp(torch.ones(3, 2))
    with pytest.raises(ValueError):
        op.forward_apply(torch.zeros((3, 4), dtype=torch.complex64, device=DEV))      # K = 2 rows expected
    empty = op.forward_apply(torch.zeros((2, 0), dtype=torch.complex64, device=DEV))
    assert empty.shape == (3, 0)


def test_identity_basis_equals_plain_tv_bitwise():
    """U = I: u = V exactly (the other terms are exact zeros), so TV through the map must give the
    plain TV's energy and gradient bit for bit."""
    from mrirecon_amd import RegManager, TVND, TVParams
    from mrirecon_amd.mapping import MappedTV, TemporalBasisOp
    rng = np.random.default_rng(11)
    T, S = 12, 48
    V = torch.from_numpy(_c(rng, T, S, S)).to(DEV)
    terms = [TVND("t", TVParams(weight=0.02, eps=1e-3, axes="temporal")),
             TVND("s", TVParams(weight=0.004, eps=1e-3, axes="spatial"))]
    gV = torch.zeros_like(V)
    E = MappedTV(terms, TemporalBasisOp(torch.eye(T))).energy_grad(V, gV)
    g = torch.zeros_like(V)
    E0 = RegManager(terms).energy_and_grad_tensors(V, g, problems=1)[0]
    assert float(E) == float(E0) and torch.equal(gV, g)
