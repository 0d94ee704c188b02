"""Learned warm start for NLCG: a 2D+t U-Net on hand-written BF16 MFMA kernels.

There is no reference code for the learned part.  The reference's only hook is
``init_mode="keep"``, which starts CG from a caller-filled ``x``
(graspcg/solvers/cg.py:48,142); this module fills it.  The warm start replaces
only the initial point: NLCG still runs to its own stop rule, and
:func:`select_warm_start` falls back to the adjoint start for every problem
whose objective at the network output is worse than at the adjoint image
(:func:`launch_warm_start` makes the same choice on the device, without a host
read between the network and the solver).

Network (``levels`` L, ``base`` channels c_l = base 2^l, BF16 activations
channels-last (volumes * frames, H, W, C), FP32 accumulation):

    enc l:  a_l = relu(conv3x3(in_l)),  e_l = relu(convT(a_l)),
            in_0 = (re, im, 0...) * scale,  in_l = avgpool2x2(e_{l-1})
    dec l:  s_l = conv1x1(h_{l+1}) (h_{L-1} = e_{L-1}),
            u_l = concat(nearest_up2(s_l), e_l),
            b_l = relu(conv3x3(u_l)),  h_l = relu(convT(b_l))          (l = L-2 .. 0)
    out  =  conv1x1(h_0);   x_warm = x_adj + (out_0 + i out_1) / scale

with conv3x3 spatial (zero padded), convT a 3-tap temporal convolution that
never crosses a volume, and scale = 1 / RMS(x_adj) per volume.  Training is
Adam on FP32 master weights with an MSE loss against plain-NLCG
reconstructions, every pass on the library's kernels (``mrh_unet_*``).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from pathlib import Path
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib

TAP_1X1, TAP_3X3, TAP_T3 = 0, 1, 2
NTAP = {TAP_1X1: 1, TAP_3X3: 9, TAP_T3: 3}


@dataclass
class ConvLayer:
    """One convolution: FP32 master weights [tap][Cout_pad][Cin] + bias, BF16 copies."""
    name: str
    kind: int
    cin: int
    cout: int
    relu: bool
    bufs: Dict[str, torch.Tensor] = field(default_factory=dict, repr=False)

    @property
    def ntap(self) -> int:
        return NTAP[self.kind]

    @property
    def cout_pad(self) -> int:
        return (self.cout + 31) // 32 * 32

    @property
    def n_weight(self) -> int:
        return self.ntap * self.cout_pad * self.cin

    def flops_per_pixel(self) -> int:
        return 2 * self.ntap * self.cin * self.cout


def layer_specs(levels: int, base: int, in_ch: int, out_ch: int) -> List[ConvLayer]:
    ch = [base << l for l in range(levels)]
    out: List[ConvLayer] = []
    cin = in_ch
    for l in range(levels):
        out.append(ConvLayer(f"enc{l}_s", TAP_3X3, cin, ch[l], True))
        out.append(ConvLayer(f"enc{l}_t", TAP_T3, ch[l], ch[l], True))
        cin = ch[l]
    for l in range(levels - 2, -1, -1):
        out.append(ConvLayer(f"dec{l}_up", TAP_1X1, ch[l + 1], ch[l], False))
        out.append(ConvLayer(f"dec{l}_s", TAP_3X3, 2 * ch[l], ch[l], True))
        out.append(ConvLayer(f"dec{l}_t", TAP_T3, ch[l], ch[l], True))
    out.append(ConvLayer("out", TAP_1X1, ch[0], out_ch, False))
    return out


class UNet2DT:
    """The warm-start network with its forward, backward and Adam passes on the device."""

    def __init__(self, levels: int = 4, base: int = 32, in_ch: int = 8, out_ch: int = 8,
                 device="cuda", seed: int = 0):
        if base % 32 or in_ch % 8 or out_ch % 8 or levels < 2:
            raise ValueError("need base % 32 == 0, in_ch % 8 == 0, out_ch % 8 == 0 and levels >= 2")
        self.levels, self.base, self.in_ch, self.out_ch = int(levels), int(base), int(in_ch), int(out_ch)
        self.ch = [self.base << l for l in range(self.levels)]
        self.device = torch.device(device)
        self.lib = _lib.load()
        self.layers = layer_specs(self.levels, self.base, self.in_ch, self.out_ch)
        self.layer = {L.name: L for L in self.layers}
        n = sum(L.n_weight + L.cout_pad for L in self.layers)
        f32 = dict(dtype=torch.float32, device=self.device)
        self.params = torch.zeros(n, **f32)
        self.grads = torch.zeros(n, **f32)
        self.m = torch.zeros(n, **f32)
        self.v = torch.zeros(n, **f32)
        off = 0
        for L in self.layers:
            nw, nb = L.n_weight, L.cout_pad
            for key, flat in (("w", self.params), ("gw", self.grads), ("m", self.m), ("v", self.v)):
                L.bufs[key] = flat[off:off + nw]
            for key, flat in (("b", self.params), ("gb", self.grads), ("mb", self.m), ("vb", self.v)):
                L.bufs[key] = flat[off + nw:off + nw + nb]
            off += nw + nb
            L.bufs["wf"] = torch.zeros(nw, dtype=torch.bfloat16, device=self.device)
            L.bufs["wbd"] = torch.zeros(L.ntap * L.cin * L.cout, dtype=torch.bfloat16, device=self.device)
        self.step = 0
        self._ws: Dict[tuple, Dict[str, torch.Tensor]] = {}
        self._red = torch.zeros(int(self.lib.mrh_reduce_workspace_len()), dtype=torch.float64, device=self.device)
        self.init_weights(seed)

    # ------------------------------------------------------------------ parameters
    def init_weights(self, seed: int) -> None:
        """He-normal convolutions (fan-in counts the 2 real input channels of the first
        layer), zero biases, zero output layer (the untrained network returns x_adj)."""
        rng = np.random.default_rng(seed)
        params = {}
        for L in self.layers:
            fan_in = L.ntap * (2 if L.name == "enc0_s" else L.cin)
            W = rng.standard_normal((L.ntap, L.cout, L.cin)) * math.sqrt(2.0 / fan_in)
            if L.name == "enc0_s":
                W[:, :, 2:] = 0.0
            if L.name == "out":
                W[:] = 0.0
            params[L.name] = (W.astype(np.float32), np.zeros(L.cout, np.float32))
        self.set_parameters(params)
        self.step = 0
        self.m.zero_()
        self.v.zero_()

    def parameters(self) -> Dict[str, Tuple[np.ndarray, np.ndarray]]:
        """name -> (W [tap][Cout][Cin], bias [Cout]) FP32 master copies on the host."""
        out = {}
        for L in self.layers:
            W = L.bufs["w"].reshape(L.ntap, L.cout_pad, L.cin)[:, :L.cout].cpu().numpy().copy()
            out[L.name] = (W, L.bufs["b"][:L.cout].cpu().numpy().copy())
        return out

    def set_parameters(self, params: Dict[str, Tuple[np.ndarray, np.ndarray]]) -> None:
        for L in self.layers:
            W, b = params[L.name]
            W = np.asarray(W, np.float32)
            if W.shape != (L.ntap, L.cout, L.cin):
                raise ValueError(f"{L.name}: weight shape {W.shape} != {(L.ntap, L.cout, L.cin)}")
            Wp = np.zeros((L.ntap, L.cout_pad, L.cin), np.float32)
            Wp[:, :L.cout] = W
            bp = np.zeros(L.cout_pad, np.float32)
            bp[:L.cout] = np.asarray(b, np.float32)
            L.bufs["w"].copy_(torch.from_numpy(Wp.reshape(-1)))
            L.bufs["b"].copy_(torch.from_numpy(bp))
        self._adam(lr=0.0)

    def save(self, path) -> None:
        arrays = {"meta": np.array([self.levels, self.base, self.in_ch, self.out_ch, self.step], np.int64)}
        for name, (W, b) in self.parameters().items():
            arrays[name + ".w"] = W
            arrays[name + ".b"] = b
        np.savez(path, **arrays)

    @classmethod
    def load(cls, path, device="cuda") -> "UNet2DT":
        with np.load(path, allow_pickle=False) as z:
            levels, base, in_ch, out_ch, step = (int(v) for v in z["meta"])
            net = cls(levels, base, in_ch, out_ch, device=device)
            net.set_parameters({L.name: (z[L.name + ".w"], z[L.name + ".b"]) for L in net.layers})
            net.step = step
        return net

    def flops_per_volume(self, F: int, H: int, W: int) -> int:
        """Dense multiply-add FLOPs of one forward pass over one volume (2 per MAC)."""
        total = 0
        for L in self.layers:
            lvl = 0
            if L.name.startswith("enc"):
                lvl = int(L.name[3])
            elif L.name.startswith("dec"):
                lvl = int(L.name[3]) + (1 if L.name.endswith("_up") else 0)
            total += L.flops_per_pixel() * F * (H >> lvl) * (W >> lvl)
        return total

    # ------------------------------------------------------------------ kernels
    def _st(self) -> int:
        return torch.cuda.current_stream(self.device).cuda_stream

    def _conv(self, L: ConvLayer, x, out, geo) -> None:
        nvol, F, H, W = geo
        _lib.check(self.lib.mrh_unet_conv_fwd(x.data_ptr(), L.bufs["wf"].data_ptr(), L.bufs["b"].data_ptr(),
                                              out.data_ptr(), L.kind, nvol, F, H, W, L.cin, L.cout, L.cout_pad,
                                              int(L.relu), self._st()))

    def _conv_bwd_data(self, L: ConvLayer, g, gx, geo) -> None:
        nvol, F, H, W = geo
        _lib.check(self.lib.mrh_unet_conv_fwd(g.data_ptr(), L.bufs["wbd"].data_ptr(), None, gx.data_ptr(),
                                              L.kind, nvol, F, H, W, L.cout, L.cin, L.cin, 0, self._st()))

    def _conv_bwd_weight(self, L: ConvLayer, g, x, geo) -> None:
        nvol, F, H, W = geo
        _lib.check(self.lib.mrh_unet_conv_bwd_weight(g.data_ptr(), x.data_ptr(), L.bufs["gw"].data_ptr(),
                                                     L.bufs["gb"].data_ptr(), L.kind, nvol, F, H, W, L.cin,
                                                     L.cout, L.cout_pad, self._st()))

    def _relu_bwd(self, g, y) -> None:
        _lib.check(self.lib.mrh_unet_relu_bwd(g.data_ptr(), y.data_ptr(), g.numel(), self._st()))

    def _adam(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        for L in self.layers:
            B = L.bufs
            wbd = None if L.name == "enc0_s" else B["wbd"].data_ptr()
            upd = lr > 0.0
            _lib.check(self.lib.mrh_unet_adam(
                B["w"].data_ptr(), B["gw"].data_ptr() if upd else None, B["m"].data_ptr() if upd else None,
                B["v"].data_ptr() if upd else None, B["wf"].data_ptr(), wbd, B["b"].data_ptr(),
                B["gb"].data_ptr() if upd else None, B["mb"].data_ptr() if upd else None,
                B["vb"].data_ptr() if upd else None, L.ntap, L.cout, L.cout_pad, L.cin,
                float(lr), beta1, beta2, eps, int(self.step), self._st()))

    # ------------------------------------------------------------------ buffers
    def _workspace(self, nvol: int, F: int, H: int, W: int, train: bool) -> Dict[str, torch.Tensor]:
        if H % (1 << (self.levels - 1)) or W % (1 << (self.levels - 1)):
            raise ValueError(f"image {H}x{W} must be divisible by {1 << (self.levels - 1)}")
        key = (nvol, F, H, W, train)
        ws = self._ws.get(key)
        if ws is not None:
            return ws
        for k in [k for k in self._ws if k[:4] != (nvol, F, H, W)]:
            del self._ws[k]
        nf = nvol * F
        bf = dict(dtype=torch.bfloat16, device=self.device)
        npx = lambda l: nf * (H >> l) * (W >> l)
        shapes = {"x8": (npx(0), self.in_ch), "out": (npx(0), self.out_ch)}
        Lv = self.levels
        for l in range(Lv):
            shapes[f"a{l}"] = (npx(l), self.ch[l])
            shapes[f"e{l}"] = (npx(l), self.ch[l])
            if l > 0:
                shapes[f"i{l}"] = (npx(l), self.ch[l - 1])
        for l in range(Lv - 1):
            shapes[f"s{l}"] = (npx(l + 1), self.ch[l])
            if train or not self._fused_upcat(l, H, W):
                shapes[f"u{l}"] = (npx(l), 2 * self.ch[l])
            shapes[f"b{l}"] = (npx(l), self.ch[l])
            shapes[f"h{l}"] = (npx(l), self.ch[l])
        ws = {k: torch.empty(s, **bf) for k, s in shapes.items()}
        ws[f"h{Lv - 1}"] = ws[f"e{Lv - 1}"]
        ws["x_warm"] = torch.empty((nf, H, W), dtype=torch.complex64, device=self.device)
        ws["rms"] = torch.zeros(nvol, dtype=torch.float64, device=self.device)
        if train:
            for k, s in shapes.items():
                if k != "x8":
                    ws["g_" + k] = torch.empty(s, **bf)
            for l in range(Lv - 1):
                ws[f"g_skip{l}"] = torch.empty((npx(l), self.ch[l]), **bf)
            ws[f"g_h{Lv - 1}"] = ws[f"g_e{Lv - 1}"]
            ws["loss"] = torch.zeros(1, dtype=torch.float64, device=self.device)
        self._ws[key] = ws
        return ws

    def volume_scales(self, x_adj: torch.Tensor, nvol: int, ws=None) -> List[float]:
        """1 / RMS |x_adj| of each volume (FP64 device reduction, one host read)."""
        nf = x_adj.shape[0]
        F = nf // nvol
        per = x_adj[0].numel() * F
        flat = x_adj.reshape(-1)
        rms = ws["rms"] if ws is not None else torch.zeros(nvol, dtype=torch.float64, device=self.device)
        for v in range(nvol):
            seg = flat[v * per:(v + 1) * per]
            _lib.check(self.lib.mrh_dot_real(seg.data_ptr(), seg.data_ptr(), per, self._red.data_ptr(),
                                             rms[v:v + 1].data_ptr(), self._st()))
        ss = rms.cpu().numpy()
        return [1.0 / math.sqrt(max(s / per, 1e-30)) for s in ss]

    # ------------------------------------------------------------------ passes
    def _fused_upcat(self, l: int, H: int, W: int) -> bool:
        """Inference reads the decoder concatenation of level l in place (no stored copy)."""
        return bool(self.lib.mrh_unet_upcat_supported(H >> l, W >> l, self.ch[l], self.ch[l]))

    def _level0_fused(self, H: int, W: int) -> bool:
        """Inference runs the full-resolution level as two fused kernels (mrh_unet_enc0_fused,
        mrh_unet_dec0_tail) when the shapes allow it."""
        return (self.ch[0] == 32 and self.in_ch == 8 and self.layer["out"].cout_pad == 32
                and bool(self.lib.mrh_unet_level0_fused_supported(H, W)))

    def _forward(self, ws, nvol: int, F: int, H: int, W: int, train: bool = True,
                 fused0: Optional[Tuple[torch.Tensor, torch.Tensor, Sequence[float]]] = None) -> None:
        """The network on ws["x8"] into ws["out"]; with fused0 = (x_adj, x_out, scales) (inference)
        the full-resolution level reads x_adj and writes x_out = x_adj + net / scale directly."""
        Lv, L = self.levels, self.layer
        geo = lambda l: (nvol, F, H >> l, W >> l)
        nf = nvol * F
        per = F * H * W
        inp = ws["x8"]
        for l in range(Lv):
            if l == 0 and fused0 is not None:
                x_adj, _, scales = fused0
                Ls, Lt = L["enc0_s"], L["enc0_t"]
                for v, sc in enumerate(scales):
                    _lib.check(self.lib.mrh_unet_enc0_fused(
                        x_adj.reshape(-1)[v * per:].data_ptr(), float(sc), Ls.bufs["wf"].data_ptr(),
                        Ls.bufs["b"].data_ptr(), Lt.bufs["wf"].data_ptr(), Lt.bufs["b"].data_ptr(),
                        ws["e0"][v * per:].data_ptr(), ws["i1"][v * per // 4:].data_ptr(), F, H, W, self._st()))
                inp = ws["i1"]
                continue
            self._conv(L[f"enc{l}_s"], inp, ws[f"a{l}"], geo(l))
            self._conv(L[f"enc{l}_t"], ws[f"a{l}"], ws[f"e{l}"], geo(l))
            if l < Lv - 1:
                _lib.check(self.lib.mrh_unet_pool2(ws[f"e{l}"].data_ptr(), ws[f"i{l + 1}"].data_ptr(), nf,
                                                   H >> l, W >> l, self.ch[l], self._st()))
                inp = ws[f"i{l + 1}"]
        for l in range(Lv - 2, -1, -1):
            self._conv(L[f"dec{l}_up"], ws[f"h{l + 1}"], ws[f"s{l}"], geo(l + 1))
            if not train and self._fused_upcat(l, H, W):
                Ls = L[f"dec{l}_s"]
                _lib.check(self.lib.mrh_unet_conv3x3_upcat(
                    ws[f"s{l}"].data_ptr(), ws[f"e{l}"].data_ptr(), Ls.bufs["wf"].data_ptr(),
                    Ls.bufs["b"].data_ptr(), ws[f"b{l}"].data_ptr(), nvol, F, H >> l, W >> l, self.ch[l],
                    self.ch[l], Ls.cout, Ls.cout_pad, int(Ls.relu), self._st()))
            else:
                _lib.check(self.lib.mrh_unet_up2_concat(ws[f"s{l}"].data_ptr(), ws[f"e{l}"].data_ptr(),
                                                        ws[f"u{l}"].data_ptr(), nf, H >> l, W >> l, self.ch[l],
                                                        self.ch[l], self._st()))
                self._conv(L[f"dec{l}_s"], ws[f"u{l}"], ws[f"b{l}"], geo(l))
            if l == 0 and fused0 is not None:
                x_adj, x_out, scales = fused0
                Lt, Lo = L["dec0_t"], L["out"]
                for v, sc in enumerate(scales):
                    _lib.check(self.lib.mrh_unet_dec0_tail(
                        ws["b0"][v * per:].data_ptr(), Lt.bufs["wf"].data_ptr(), Lt.bufs["b"].data_ptr(),
                        Lo.bufs["wf"].data_ptr(), Lo.bufs["b"].data_ptr(), x_adj.reshape(-1)[v * per:].data_ptr(),
                        float(1.0 / sc), x_out.reshape(-1)[v * per:].data_ptr(), F, H, W, self._st()))
                return
            self._conv(L[f"dec{l}_t"], ws[f"b{l}"], ws[f"h{l}"], geo(l))
        self._conv(L["out"], ws["h0"], ws["out"], geo(0))

    def _backward(self, ws, nvol: int, F: int, H: int, W: int) -> None:
        """Weight/bias gradients from ws['g_out'] (= dLoss/dout) into self.grads."""
        Lv, L = self.levels, self.layer
        geo = lambda l: (nvol, F, H >> l, W >> l)
        nf = nvol * F
        _lib.check(self.lib.mrh_unet_fill(self.grads.data_ptr(), self.grads.numel(), 0.0, self._st()))
        self._conv_bwd_weight(L["out"], ws["g_out"], ws["h0"], geo(0))
        self._conv_bwd_data(L["out"], ws["g_out"], ws["g_h0"], geo(0))
        for l in range(Lv - 1):
            self._relu_bwd(ws[f"g_h{l}"], ws[f"h{l}"])
            self._conv_bwd_weight(L[f"dec{l}_t"], ws[f"g_h{l}"], ws[f"b{l}"], geo(l))
            self._conv_bwd_data(L[f"dec{l}_t"], ws[f"g_h{l}"], ws[f"g_b{l}"], geo(l))
            self._relu_bwd(ws[f"g_b{l}"], ws[f"b{l}"])
            self._conv_bwd_weight(L[f"dec{l}_s"], ws[f"g_b{l}"], ws[f"u{l}"], geo(l))
            self._conv_bwd_data(L[f"dec{l}_s"], ws[f"g_b{l}"], ws[f"g_u{l}"], geo(l))
            _lib.check(self.lib.mrh_unet_up2_concat_bwd(ws[f"g_u{l}"].data_ptr(), ws[f"g_s{l}"].data_ptr(),
                                                        ws[f"g_skip{l}"].data_ptr(), nf, H >> l, W >> l,
                                                        self.ch[l], self.ch[l], self._st()))
            self._conv_bwd_weight(L[f"dec{l}_up"], ws[f"g_s{l}"], ws[f"h{l + 1}"], geo(l + 1))
            self._conv_bwd_data(L[f"dec{l}_up"], ws[f"g_s{l}"], ws[f"g_h{l + 1}"], geo(l + 1))
        for l in range(Lv - 1, -1, -1):
            if l < Lv - 1:
                _lib.check(self.lib.mrh_unet_pool2_bwd(ws[f"g_i{l + 1}"].data_ptr(), ws[f"g_skip{l}"].data_ptr(),
                                                       ws[f"g_e{l}"].data_ptr(), nf, H >> l, W >> l, self.ch[l],
                                                       self._st()))
            self._relu_bwd(ws[f"g_e{l}"], ws[f"e{l}"])
            self._conv_bwd_weight(L[f"enc{l}_t"], ws[f"g_e{l}"], ws[f"a{l}"], geo(l))
            self._conv_bwd_data(L[f"enc{l}_t"], ws[f"g_e{l}"], ws[f"g_a{l}"], geo(l))
            self._relu_bwd(ws[f"g_a{l}"], ws[f"a{l}"])
            inp = ws["x8"] if l == 0 else ws[f"i{l}"]
            self._conv_bwd_weight(L[f"enc{l}_s"], ws[f"g_a{l}"], inp, geo(l))
            if l > 0:
                self._conv_bwd_data(L[f"enc{l}_s"], ws[f"g_a{l}"], ws[f"g_i{l}"], geo(l))

    def _pack(self, ws, x_adj, scales, F) -> None:
        per = F * x_adj.shape[1] * x_adj.shape[2]
        for v, s in enumerate(scales):
            _lib.check(self.lib.mrh_unet_pack_input(x_adj.reshape(-1)[v * per:].data_ptr(),
                                                    ws["x8"][v * per:].data_ptr(), per, self.in_ch, float(s),
                                                    self._st()))

    def _unpack(self, ws, x_adj, out, scales, F) -> None:
        per = F * x_adj.shape[1] * x_adj.shape[2]
        for v, s in enumerate(scales):
            _lib.check(self.lib.mrh_unet_unpack_residual(x_adj.reshape(-1)[v * per:].data_ptr(),
                                                         ws["out"][v * per:].data_ptr(),
                                                         out.reshape(-1)[v * per:].data_ptr(), per, self.out_ch,
                                                         float(1.0 / s), self._st()))

    @staticmethod
    def _check_image(x: torch.Tensor, nvol: int) -> Tuple[int, int, int]:
        if x.dtype != torch.complex64 or x.dim() != 3 or not x.is_contiguous() or x.device.type != "cuda":
            raise ValueError("expected a contiguous complex64 CUDA tensor (volumes*frames, H, W)")
        nf, H, W = x.shape
        if nvol < 1 or nf % nvol:
            raise ValueError(f"{nf} frames do not split into {nvol} volumes")
        return nf // nvol, H, W

    @torch.no_grad()
    def infer(self, x_adj: torch.Tensor, nvol: int = 1, out: Optional[torch.Tensor] = None,
              scales: Optional[Sequence[float]] = None, fused: Optional[bool] = None) -> torch.Tensor:
        """x_warm = x_adj + U-Net(x_adj) for (nvol * F, H, W) complex64 input.  fused: run the
        full-resolution level as fused kernels (default: whenever the shapes allow it); False
        runs every layer separately, as the training forward does."""
        F, H, W = self._check_image(x_adj, nvol)
        ws = self._workspace(nvol, F, H, W, train=False)
        if scales is None:
            scales = self.volume_scales(x_adj, nvol, ws)
        out = ws["x_warm"] if out is None else out
        if fused is None:
            fused = self._level0_fused(H, W)
        if fused:
            self._forward(ws, nvol, F, H, W, train=False, fused0=(x_adj, out, list(scales)))
            return out
        self._pack(ws, x_adj, scales, F)
        self._forward(ws, nvol, F, H, W, train=False)
        self._unpack(ws, x_adj, out, scales, F)
        return out

    @torch.no_grad()
    def loss_and_grad(self, x_adj: torch.Tensor, x_target: torch.Tensor, nvol: int = 1,
                      scales: Optional[Sequence[float]] = None, objective: Optional["ObjectiveGrad"] = None,
                      y: Optional[torch.Tensor] = None, obj_weight: float = 0.0,
                      f_norm: Optional[Sequence[float]] = None) -> torch.Tensor:
        """Forward + backward of

            L = mean |(x_warm - x_target) scale|^2  +  obj_weight * sum_v f_v(x_warm) / f_norm[v]

        (f = the NLCG objective of k-space ``y``, evaluated by ``objective``; the second
        term is optional).  Gradients land in ``self.grads``; returns the MSE term as a
        device FP64 scalar (no host synchronisation; ``objective.value()`` reads f)."""
        F, H, W = self._check_image(x_adj, nvol)
        if x_target.shape != x_adj.shape or x_target.dtype != torch.complex64 or not x_target.is_contiguous():
            raise ValueError("x_target must match x_adj (contiguous complex64)")
        ws = self._workspace(nvol, F, H, W, train=True)
        if scales is None:
            scales = self.volume_scales(x_adj, nvol, ws)
        self._pack(ws, x_adj, scales, F)
        self._forward(ws, nvol, F, H, W)
        self._unpack(ws, x_adj, ws["x_warm"], scales, F)
        gx = None
        if objective is not None and obj_weight:
            if y is None or f_norm is None or len(f_norm) != nvol:
                raise ValueError("the objective term needs y and one f_norm per volume")
            gx = objective(ws["x_warm"], y)
        _lib.check(self.lib.mrh_unet_fill(ws["loss"].data_ptr(), 2, 0.0, self._st()))   # 2 floats = 1 double
        per = F * H * W
        npix = nvol * per
        for v, s in enumerate(scales):
            coef = float(obj_weight / (f_norm[v] * s)) if gx is not None else 0.0
            _lib.check(self.lib.mrh_unet_mse_grad(ws["x_warm"].reshape(-1)[v * per:].data_ptr(),
                                                  x_target.reshape(-1)[v * per:].data_ptr(),
                                                  None if gx is None else gx.reshape(-1)[v * per:].data_ptr(),
                                                  coef, ws["g_out"][v * per:].data_ptr(), ws["loss"].data_ptr(),
                                                  per, self.out_ch, float(s), float(1.0 / npix), self._st()))
        self._backward(ws, nvol, F, H, W)
        return ws["loss"]

    @torch.no_grad()
    def train_step(self, x_adj: torch.Tensor, x_target: torch.Tensor, lr: float, nvol: int = 1,
                   **objective_term) -> torch.Tensor:
        """One Adam step on :meth:`loss_and_grad` (keyword arguments select the objective term)."""
        loss = self.loss_and_grad(x_adj, x_target, nvol, **objective_term)
        self.step += 1
        self._adam(lr)
        return loss


# ---------------------------------------------------------------------- objective and fallback

class ObjectiveEval:
    """f_p(x) = 1/2 |A x - y|^2 + TV(x) per problem p on the device (FP64 sums),
    the objective of graspcg/ops/objective.py:184-291 at t = 0."""

    def __init__(self, plan, y: torch.Tensor, regm, problems: int, frames: int):
        from .regularization import RegManager
        self.plan, self.P, self.F = plan, int(problems), int(frames)
        self.lib = _lib.load()
        self.y = y.reshape(-1)
        self.Ax = torch.empty_like(self.y)
        self.r = torch.empty_like(self.y)
        dev = y.device
        self.work = torch.zeros(int(self.lib.mrh_reduce_workspace_len()), dtype=torch.float64, device=dev)
        self.e = torch.zeros(2 * self.P, dtype=torch.float64, device=dev)     # data terms, then TV terms
        self.e_data, self.e_tv = self.e[:self.P], self.e[self.P:]
        S0, S1 = plan.S
        if regm is not None and regm.graph() is not None:
            raise NotImplementedError("objective evaluation covers TV terms only (no graph-Laplacian term)")
        self.tv = (regm or RegManager([])).tv_desc(self.P, self.F, S0, S1)
        C = plan.C
        self.k_prob = [C * int(plan.k_offsets[p * self.F]) for p in range(self.P + 1)]

    def __call__(self, x: torch.Tensor) -> np.ndarray:
        self.launch(x)
        return self.read()

    def launch(self, x: torch.Tensor) -> None:
        """Enqueue the evaluation of f(x) on the current stream (no host synchronisation)."""
        st = torch.cuda.current_stream(x.device).cuda_stream
        self.plan.forward(x, self.Ax)
        for p in range(self.P):
            a, b = self.k_prob[p], self.k_prob[p + 1]
            _lib.check(self.lib.mrh_residual_energy(self.r[a:].data_ptr(), self.Ax[a:].data_ptr(),
                                                    self.Ax[a:].data_ptr(), self.y[a:].data_ptr(), 0.0, b - a,
                                                    self.work.data_ptr(), self.e_data[p:p + 1].data_ptr(), st))
        _lib.check(self.lib.mrh_tv_energy_grad(ctypes.byref(self.tv), x.data_ptr(), None, self.work.data_ptr(),
                                               self.e_tv.data_ptr(), st))

    def launch_adjoint_bound(self, x_adj: torch.Tensor) -> None:
        """Enqueue, for x_adj = A^H y, TV(x_adj) (into the TV terms) and the lower bound
        ``self.lb`` of f(x_adj) per problem without a NUFFT: Re<A A^H y, y> = |A^H y|^2, so by
        Cauchy-Schwarz 1/2 |A x_adj - y|^2 >= 1/2 (|y|^2 - |x_adj|^2)^2 / |y|^2."""
        st = torch.cuda.current_stream(x_adj.device).cuda_stream
        if not hasattr(self, "norms"):
            self.norms = torch.zeros(2 * self.P, dtype=torch.float64, device=self.y.device)   # |y_p|^2, |x_p|^2
        xf = x_adj.reshape(self.P, -1)
        n_img = xf.shape[1]
        for p in range(self.P):
            a, b = self.k_prob[p], self.k_prob[p + 1]
            _lib.check(self.lib.mrh_dot_real(self.y[a:].data_ptr(), self.y[a:].data_ptr(), b - a, self.work.data_ptr(),
                                             self.norms[p:p + 1].data_ptr(), st))
            _lib.check(self.lib.mrh_dot_real(xf[p].data_ptr(), xf[p].data_ptr(), n_img, self.work.data_ptr(),
                                             self.norms[self.P + p:self.P + p + 1].data_ptr(), st))
        _lib.check(self.lib.mrh_tv_energy_grad(ctypes.byref(self.tv), x_adj.data_ptr(), None, self.work.data_ptr(),
                                               self.e_tv.data_ptr(), st))
        yy, xx = self.norms[:self.P], self.norms[self.P:]
        self.lb = 0.5 * (yy - xx) ** 2 / torch.clamp(yy, min=1e-300) + self.e_tv

    def launch_data_masked(self, x: torch.Tensor, active: torch.Tensor) -> None:
        """Enqueue the data terms 1/2 |A x - y|^2 of the problems with ``active[p]`` != 0 (device
        int32[P]; the NUFFT skips the other problems' frames, whose data terms are then undefined)."""
        st = torch.cuda.current_stream(x.device).cuda_stream
        self.plan.forward_masked(x, self.Ax, active.data_ptr(), self.F)
        for p in range(self.P):
            a, b = self.k_prob[p], self.k_prob[p + 1]
            _lib.check(self.lib.mrh_residual_energy(self.r[a:].data_ptr(), self.Ax[a:].data_ptr(),
                                                    self.Ax[a:].data_ptr(), self.y[a:].data_ptr(), 0.0, b - a,
                                                    self.work.data_ptr(), self.e_data[p:p + 1].data_ptr(), st))

    def read(self) -> np.ndarray:
        """f per problem of the last launched evaluation (synchronises with its stream)."""
        e = self.e.cpu().numpy()
        return 0.5 * e[:self.P] + e[self.P:]

    def value_device(self) -> torch.Tensor:
        """f per problem as an FP64 device tensor (no host synchronisation).  Equal bit for
        bit to :meth:`read`: the product by 0.5 is exact, and the sum is one rounded add."""
        return self.e_data * 0.5 + self.e_tv


class ObjectiveGrad:
    """g = A^H (A x - y) + grad TV(x) and f = 1/2 |A x - y|^2 + TV(x) for one problem, on the
    device -- the gradient the NLCG engine computes (graspcg/ops/objective.py:184-291 at t = 0),
    used as the gradient of the objective term of the warm-start training loss."""

    def __init__(self, plan, regm, frames: int):
        from .regularization import RegManager
        self.plan, self.F = plan, int(frames)
        self.lib = _lib.load()
        dev = plan.maps.device
        S0, S1 = plan.S
        self.Ax = torch.empty(plan.n_k, dtype=torch.complex64, device=dev)
        self.r = torch.empty_like(self.Ax)
        self.g = torch.empty((plan.B, S0, S1), dtype=torch.complex64, device=dev)
        self.work = torch.zeros(int(self.lib.mrh_reduce_workspace_len()), dtype=torch.float64, device=dev)
        self.energy = torch.zeros(2, dtype=torch.float64, device=dev)
        if plan.B != self.F:
            raise ValueError("ObjectiveGrad handles one problem (plan frames == frames)")
        if regm is not None and regm.graph() is not None:
            raise NotImplementedError("objective gradients cover TV terms only (no graph-Laplacian term)")
        self.tv = (regm or RegManager([])).tv_desc(1, self.F, S0, S1)

    def __call__(self, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
        st = torch.cuda.current_stream(x.device).cuda_stream
        y = y.reshape(-1)
        self.plan.forward(x, self.Ax)
        _lib.check(self.lib.mrh_residual_energy(self.r.data_ptr(), self.Ax.data_ptr(), self.Ax.data_ptr(),
                                                y.data_ptr(), 0.0, y.numel(), self.work.data_ptr(),
                                                self.energy[0:1].data_ptr(), st))
        self.plan.adjoint(self.r, self.g)
        _lib.check(self.lib.mrh_tv_energy_grad(ctypes.byref(self.tv), x.data_ptr(), self.g.data_ptr(),
                                               self.work.data_ptr(), self.energy[1:2].data_ptr(), st))
        return self.g

    def value(self) -> float:
        e = self.energy.cpu().numpy()
        return float(0.5 * e[0] + e[1])


def select_warm_start(objective: ObjectiveEval, x_adj: torch.Tensor, x_warm: torch.Tensor,
                      out: Optional[torch.Tensor] = None) -> Tuple[torch.Tensor, np.ndarray, np.ndarray, np.ndarray]:
    """Per problem, keep x_warm unless f(x_warm) > f(x_adj) (then the adjoint start).

    Returns (x0, accepted[P] bool, f_warm[P], f_adj[P]).  x_warm is evaluated last, so
    ``objective.Ax`` holds A x0 when every problem accepted (``accepted.all()``)."""
    f_a = objective(x_adj)
    f_w = objective(x_warm)
    acc = f_w <= f_a
    x0 = x_warm if out is None else out.copy_(x_warm)
    if not acc.all():
        per = x_adj.shape[0] // objective.P
        for p in np.nonzero(~acc)[0]:
            x0[p * per:(p + 1) * per].copy_(x_adj[p * per:(p + 1) * per])
    return x0, acc, f_w, f_a


class WarmStartChoice:
    """The per-problem start chosen on the device by :func:`launch_warm_start`.

    ``x0`` is the start image and ``Ax0`` (the buffer ``objective.Ax``) holds A x0 for
    every problem, accepted or not.  :meth:`read` synchronises and returns
    (accepted[P] bool, f_warm[P], f_adj[P])."""

    def __init__(self, x0: torch.Tensor, Ax0: torch.Tensor, f_w: torch.Tensor, f_a: torch.Tensor):
        self.x0, self.Ax0, self._f_w, self._f_a = x0, Ax0, f_w, f_a

    def read(self) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        f = torch.stack((self._f_w, self._f_a)).cpu().numpy()       # one copy, one sync
        return f[0] <= f[1], f[0], f[1]


def launch_warm_start(net: UNet2DT, objective: ObjectiveEval, objective_adj: ObjectiveEval, x_adj: torch.Tensor,
                      nvol: int = 1, out: Optional[torch.Tensor] = None, side_stream=None,
                      bound: bool = True) -> WarmStartChoice:
    """Network + fallback check, the choice per problem made on the device, so nothing waits for
    the host after the network's input scales are read (the solver's first iteration queues right
    behind).

    ``objective`` and ``objective_adj`` are two evaluators of the same problem (separate
    buffers).  Problem p starts from x_warm unless f(x_warm) > f(x_adj); its slice of
    ``objective.Ax`` is then replaced by A x_adj from ``objective_adj``, so ``objective.Ax`` holds
    A x0 for every problem.

    ``bound`` (default): while the network runs, a side stream computes TV(x_adj) and the lower
    bound 1/2 (|y|^2 - |A^H y|^2)^2 / |y|^2 + TV(A^H y) <= f(x_adj) (no NUFFT).  Where
    f(x_warm) <= (1 - 1e-4) x bound the start is decided without f(x_adj); A x_adj and the exact
    f(x_adj) are computed (a masked forward NUFFT) only for the other problems, and f_adj reports
    the bound for the decided ones.  ``bound=False``: f(x_adj) in full on the side stream for
    every problem, overlapped with the network.  Both make the same choice."""
    main = torch.cuda.current_stream(x_adj.device)
    side = side_stream or torch.cuda.Stream(device=x_adj.device)
    ready = torch.cuda.Event()
    ready.record(main)
    with torch.cuda.stream(side):
        side.wait_event(ready)
        if bound:
            objective_adj.launch_adjoint_bound(x_adj)
        else:
            objective_adj.launch(x_adj)
    x_warm = net.infer(x_adj, nvol=nvol)
    main.wait_stream(side)
    objective.launch(x_warm)
    f_w = objective.value_device()
    if bound:
        lb = objective_adj.lb
        need = f_w > lb * (1.0 - 1e-4)                      # the bound does not decide these
        objective_adj.launch_data_masked(x_adj, need.to(torch.int32))
        f_a = torch.where(need, objective_adj.value_device(), lb)
    else:
        f_a = objective_adj.value_device()
    acc = f_w <= f_a                                       # device bool per problem
    x0 = x_warm if out is None else out
    P = objective.P
    sel = acc.view(P, *([1] * x_adj.dim()))                 # every problem holds the same frame count
    torch.where(sel, x_warm.view(P, -1, *x_adj.shape[1:]), x_adj.view(P, -1, *x_adj.shape[1:]),
                out=x0.view(P, -1, *x_adj.shape[1:]))
    for p in range(P):                                     # k-space segments are ragged
        a, b = objective.k_prob[p], objective.k_prob[p + 1]
        torch.where(acc[p], objective.Ax[a:b], objective_adj.Ax[a:b], out=objective.Ax[a:b])
    return WarmStartChoice(x0, objective.Ax, f_w, f_a)


def warm_start(net: UNet2DT, objective: ObjectiveEval, objective_adj: ObjectiveEval, x_adj: torch.Tensor,
               nvol: int = 1, out: Optional[torch.Tensor] = None, side_stream=None, bound: bool = True
               ) -> Tuple[torch.Tensor, np.ndarray, np.ndarray, np.ndarray]:
    """:func:`launch_warm_start` followed by the host read of its choice: returns
    (x0, accepted[P] bool, f_warm[P], f_adj[P]), the choice of
    ``select_warm_start(objective, x_adj, net.infer(...))`` (f_adj a lower bound where ``bound``
    decided without it); ``objective.Ax`` holds A x0."""
    ch = launch_warm_start(net, objective, objective_adj, x_adj, nvol=nvol, out=out, side_stream=side_stream,
                           bound=bound)
    acc, f_w, f_a = ch.read()
    return ch.x0, acc, f_w, f_a

