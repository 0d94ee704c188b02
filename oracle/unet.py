"""CPU oracle of the 2D+t U-Net warm start (torch autograd on the CPU).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline as the checker.  The product path never calls it.

No reference code exists for the learned warm start; the reference only keeps
a caller-filled initial point (init_mode="keep", graspcg/solvers/cg.py:48,142).
This restates the network the HIP kernels implement, in torch's own conv3d /
avg_pool3d / interpolate (NCDHW, D = frames), so that the GPU forward, the
three convolution passes and the full backward can be compared with autograd:

* taps: 3x3 spatial = conv3d kernel (1, 3, 3), padding (0, 1, 1); temporal =
  kernel (3, 1, 1), padding (1, 0, 0) (each volume is one batch entry, so the
  temporal padding never crosses volumes); 1x1 = kernel (1, 1, 1).
  Weight layout [tap][Cout][Cin] with tap t = 3 (dy+1) + (dx+1) (spatial) or
  df + 1 (temporal), i.e. conv3d weight[co, ci, kt, ky, kx] = W[t][co][ci]
  (cross-correlation: out[p] = sum W[t] x[p + off(t)]);
* graph (levels L, channels c_l = base 2^l):
    enc l:  a_l = relu(conv3x3(in_l)), e_l = relu(convT(a_l)),
            in_0 = packed input, in_l = avgpool2x2(e_{l-1});
    dec l (l = L-2 .. 0):  s_l = conv1x1(h_{l+1}) (h_{L-1} = e_{L-1}),
            u_l = concat(nearest_up2(s_l), e_l), b_l = relu(conv3x3(u_l)),
            h_l = relu(convT(b_l));
    out = conv1x1(h_0);  x_warm = x_adj + (out[0] + i out[1]) / scale.
* `round_bf16=True` rounds every stored activation and every stored gradient
  to bfloat16 (round to nearest even), the storage precision of the kernels.

Pinned by tests/test_oracle_unet.py: brute-force loops for each tap set, the
adjoint identities of pooling / upsampling, and finite-difference gradients.
"""
from __future__ import annotations

from typing import Dict, List, Tuple

import torch
import torch.nn.functional as Fn

TAP_1X1, TAP_3X3, TAP_T3 = 0, 1, 2
NTAP = {TAP_1X1: 1, TAP_3X3: 9, TAP_T3: 3}


def layer_specs(levels: int = 4, base: int = 32, in_ch: int = 8, out_ch: int = 8) -> List[Tuple[str, int, int, int, bool]]:
    """(name, tap kind, Cin, Cout, relu) in execution order of the forward pass."""
    ch = [base << l for l in range(levels)]
    specs = []
    cin = in_ch
    for l in range(levels):
        specs.append((f"enc{l}_s", TAP_3X3, cin, ch[l], True))
        specs.append((f"enc{l}_t", TAP_T3, ch[l], ch[l], True))
        cin = ch[l]
    for l in range(levels - 2, -1, -1):
        specs.append((f"dec{l}_up", TAP_1X1, ch[l + 1], ch[l], False))
        specs.append((f"dec{l}_s", TAP_3X3, 2 * ch[l], ch[l], True))
        specs.append((f"dec{l}_t", TAP_T3, ch[l], ch[l], True))
    specs.append(("out", TAP_1X1, ch[0], out_ch, False))
    return specs


class _RoundBF16(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return x.to(torch.bfloat16).to(x.dtype)

    @staticmethod
    def backward(ctx, g):
        return g.to(torch.bfloat16).to(g.dtype)


def bf16_round(x: torch.Tensor) -> torch.Tensor:
    return x.to(torch.bfloat16).to(x.dtype)


def conv_weight(W: torch.Tensor, kind: int) -> torch.Tensor:
    """[tap][Cout][Cin] -> conv3d weight (Cout, Cin, kt, ky, kx)."""
    T, co, ci = W.shape
    if kind == TAP_3X3:
        return W.permute(1, 2, 0).reshape(co, ci, 1, 3, 3)
    if kind == TAP_T3:
        return W.permute(1, 2, 0).reshape(co, ci, 3, 1, 1)
    return W.permute(1, 2, 0).reshape(co, ci, 1, 1, 1)


def conv(x: torch.Tensor, W: torch.Tensor, b, kind: int, relu: bool = False) -> torch.Tensor:
    """x (N, Cin, F, H, W) -> (N, Cout, F, H, W)."""
    pad = {TAP_3X3: (0, 1, 1), TAP_T3: (1, 0, 0), TAP_1X1: (0, 0, 0)}[kind]
    y = Fn.conv3d(x, conv_weight(W, kind), b, padding=pad)
    return torch.relu(y) if relu else y


def pool2(x: torch.Tensor) -> torch.Tensor:
    return Fn.avg_pool3d(x, (1, 2, 2))


def up2(x: torch.Tensor) -> torch.Tensor:
    return Fn.interpolate(x, scale_factor=(1, 2, 2), mode="nearest")


def to_ncdhw(x_cl: torch.Tensor, nvol: int, F: int) -> torch.Tensor:
    """(nvol*F, H, W, C) channels-last -> (nvol, C, F, H, W)."""
    nf, H, W, C = x_cl.shape
    return x_cl.reshape(nvol, F, H, W, C).permute(0, 4, 1, 2, 3)


def to_channels_last(x: torch.Tensor) -> torch.Tensor:
    """(nvol, C, F, H, W) -> (nvol*F, H, W, C)."""
    n, C, F, H, W = x.shape
    return x.permute(0, 2, 3, 4, 1).reshape(n * F, H, W, C)


def pack_input(x_adj: torch.Tensor, scale: float, in_ch: int = 8) -> torch.Tensor:
    """complex (nf, H, W) -> real (nf, H, W, in_ch) = (re, im, 0, ...) * scale."""
    out = torch.zeros(x_adj.shape + (in_ch,), dtype=torch.float64 if x_adj.dtype == torch.complex128 else torch.float32)
    out[..., 0] = x_adj.real * scale
    out[..., 1] = x_adj.imag * scale
    return out


def forward(x8: torch.Tensor, params: Dict[str, Tuple[torch.Tensor, torch.Tensor]], nvol: int, F: int,
            levels: int = 4, base: int = 32, round_bf16: bool = False) -> torch.Tensor:
    """Network output (nvol*F, H, W, out_ch) for the packed input x8 (nvol*F, H, W, in_ch)."""
    R = _RoundBF16.apply if round_bf16 else (lambda t: t)

    def L(name, x, kind, relu):
        W, b = params[name]
        return R(conv(x, W, b, kind, relu))

    x = R(to_ncdhw(x8, nvol, F))
    e = []
    inp = x
    for l in range(levels):
        a = L(f"enc{l}_s", inp, TAP_3X3, True)
        el = L(f"enc{l}_t", a, TAP_T3, True)
        e.append(el)
        if l < levels - 1:
            inp = R(pool2(el))
    h = e[-1]
    for l in range(levels - 2, -1, -1):
        s = L(f"dec{l}_up", h, TAP_1X1, False)
        u = torch.cat([up2(s), e[l]], dim=1)
        bl = L(f"dec{l}_s", u, TAP_3X3, True)
        h = L(f"dec{l}_t", bl, TAP_T3, True)
    out = L("out", h, TAP_1X1, False)
    return to_channels_last(out)


def mse_loss(out: torch.Tensor, x_adj: torch.Tensor, x_target: torch.Tensor, scale: float) -> torch.Tensor:
    """sum |(x_warm - x_target) scale|^2 / npix with x_warm = x_adj + (out0 + i out1) / scale."""
    er = (x_adj.real - x_target.real) * scale + out[..., 0]
    ei = (x_adj.imag - x_target.imag) * scale + out[..., 1]
    return (er * er + ei * ei).sum() / er.numel()


def warm_start(out: torch.Tensor, x_adj: torch.Tensor, scale: float) -> torch.Tensor:
    return x_adj + torch.complex(out[..., 0], out[..., 1]).to(x_adj.dtype) / scale
