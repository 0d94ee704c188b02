"""U-Net warm-start kernels on the GPU against the CPU autograd oracle (oracle/unet.py).

Bars (BF16 storage, FP32 accumulation):
* conv forward / backward-data: the BF16 output equals the FP64 convolution of the
  same BF16 inputs to BF16 rounding (|err| <= 2^-8 |ref| + 1e-3 * RMS(ref) per element);
* backward-weight / bias: FP32 atomics of exact BF16 products, rel L2 <= 1e-5;
* whole network (forward, loss, every weight gradient) vs the oracle with BF16
  rounding of each stored activation and gradient: rel L2 <= 1e-2 per tensor.
"""
import numpy as np
import pytest
import torch

from mrirecon_amd import _lib, synth
from mrirecon_amd.warmstart import ObjectiveEval, UNet2DT, select_warm_start
from oracle import unet as ou

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def bf16_np(a):
    return torch.from_numpy(np.asarray(a, np.float32)).to(torch.bfloat16).to(torch.float64).numpy()


def to_dev_bf16(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(torch.bfloat16).to(DEV)


def st():
    return torch.cuda.current_stream(DEV).cuda_stream


def rel_l2(a, b):
    return float(np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel()) / max(np.linalg.norm(np.asarray(b).ravel()), 1e-300))


def oracle_conv(x, W, b, kind, nvol, F, relu=False):
    xt = ou.to_ncdhw(torch.from_numpy(x), nvol, F)
    return ou.to_channels_last(ou.conv(xt, torch.from_numpy(W), None if b is None else torch.from_numpy(b),
                                       kind, relu)).numpy()


def check_bf16_close(got, ref):
    err = np.abs(got - ref)
    tol = 2.0 ** -8 * np.abs(ref) + 1e-3 * np.sqrt(np.mean(ref ** 2))
    assert np.all(err <= tol), (float((err - tol).max()), float(np.sqrt(np.mean(ref ** 2))))


GEOM = (2, 5, 16, 24)   # volumes, frames, H, W: 3840 pixels (not a multiple of the 128-pixel tile)


@pytest.mark.parametrize("geom", [GEOM, (1, 3, 10, 18)])
@pytest.mark.parametrize("kind,cin,cout,relu", [
    (ou.TAP_3X3, 8, 32, True), (ou.TAP_3X3, 64, 32, False), (ou.TAP_T3, 32, 64, True),
    (ou.TAP_T3, 64, 128, False), (ou.TAP_1X1, 32, 8, False), (ou.TAP_1X1, 256, 128, True),
    (ou.TAP_3X3, 16, 64, True), (ou.TAP_3X3, 48, 96, False), (ou.TAP_3X3, 128, 64, True),
    (ou.TAP_T3, 32, 32, False)])
def test_conv_forward(kind, cin, cout, relu, geom):
    """Every kernel variant: register tiles (Cin = 8, 1x1, 32-cout temporal), LDS tiles 64 x 128
    (16-channel chunks), 64 x 256 (3x3, Cin >= 128), 32 x 128 (32-channel chunks) and 32 x 256;
    tiles that end inside a row and in the middle of the last tile (540 pixels)."""
    lib = _lib.load()
    nvol, F, H, W = geom
    npix = nvol * F * H * W
    rng = np.random.default_rng(cin + cout + kind)
    T = ou.NTAP[kind]
    x = bf16_np(rng.standard_normal((nvol * F, H, W, cin)))
    Wt = bf16_np(rng.standard_normal((T, cout, cin)) * np.sqrt(2.0 / (T * cin)))
    b = (rng.standard_normal(cout) * 0.1).astype(np.float32)
    cpad = (cout + 31) // 32 * 32
    Wp = np.zeros((T, cpad, cin))
    Wp[:, :cout] = Wt
    out = torch.full((npix, cout), 7.0, dtype=torch.bfloat16, device=DEV)
    bd = torch.from_numpy(b).to(DEV)
    xd, wd = to_dev_bf16(x), to_dev_bf16(Wp)     # keep every buffer alive while the kernel runs
    _lib.check(lib.mrh_unet_conv_fwd(xd.data_ptr(), wd.data_ptr(), bd.data_ptr(),
                                     out.data_ptr(), kind, nvol, F, H, W, cin, cout, cpad, int(relu), st()))
    ref = oracle_conv(x, Wt, b.astype(np.float64), kind, nvol, F, relu).reshape(npix, cout)
    check_bf16_close(out.float().cpu().numpy().astype(np.float64), ref)


@pytest.mark.parametrize("geom", [(1, 2, 8, 64), (2, 2, 16, 32), (1, 1, 4, 128), (2, 4, 8, 32)])
@pytest.mark.parametrize("kind", [ou.TAP_3X3, ou.TAP_T3])
@pytest.mark.parametrize("cin,cout,relu", [(32, 64, True), (64, 32, False), (128, 64, True), (64, 96, False),
                                           (32, 48, True), (256, 128, False)])
def test_conv_tiles(kind, cin, cout, relu, geom):
    """3x3 convolutions on 2-D tiles (4 x 64 or 8 x 32 pixels) and temporal ones on 256-pixel
    tiles (Cin % 32 == 0): halo rows and columns at every image border, first and last frames
    of each volume, 32- and 64-cout tiles, couts that end inside a tile."""
    lib = _lib.load()
    nvol, F, H, W = geom
    npix = nvol * F * H * W
    T = ou.NTAP[kind]
    rng = np.random.default_rng(cin * 7 + cout + W + kind)
    x = bf16_np(rng.standard_normal((nvol * F, H, W, cin)))
    Wt = bf16_np(rng.standard_normal((T, cout, cin)) * np.sqrt(2.0 / (T * cin)))
    b = (rng.standard_normal(cout) * 0.1).astype(np.float32)
    cpad = (cout + 31) // 32 * 32
    Wp = np.zeros((T, cpad, cin))
    Wp[:, :cout] = Wt
    out = torch.full((npix, cout), 7.0, dtype=torch.bfloat16, device=DEV)
    bd = torch.from_numpy(b).to(DEV)
    xd, wd = to_dev_bf16(x), to_dev_bf16(Wp)
    _lib.check(lib.mrh_unet_conv_fwd(xd.data_ptr(), wd.data_ptr(), bd.data_ptr(),
                                     out.data_ptr(), kind, nvol, F, H, W, cin, cout, cpad, int(relu), st()))
    ref = oracle_conv(x, Wt, b.astype(np.float64), kind, nvol, F, relu).reshape(npix, cout)
    check_bf16_close(out.float().cpu().numpy().astype(np.float64), ref)


@pytest.mark.parametrize("geom", [(1, 34, 16, 16), (2, 7, 32, 32), (1, 1, 16, 16), (2, 2, 64, 64)])
@pytest.mark.parametrize("cin,cout,relu,bias", [(64, 64, True, True), (64, 40, False, True), (128, 128, True, False),
                                                (128, 64, False, True), (256, 256, True, True), (256, 72, False, False)])
def test_temporal_layers_walking_frames(cin, cout, relu, bias, geom):
    """Temporal layers with Cin = 64, 128 or 256 (the network's levels 1-3): every wave walks a group
    of frames holding the frames before, at and after the current one in registers.  Groups of 5
    frames (34 frames: the last group shorter, and every group boundary inside the volume), one
    frame, two volumes, couts that end inside a cout block, with and without bias."""
    lib = _lib.load()
    nvol, F, H, W = geom
    npix = nvol * F * H * W
    rng = np.random.default_rng(cin * 11 + cout + F + H)
    x = bf16_np(rng.standard_normal((nvol * F, H, W, cin)))
    Wt = bf16_np(rng.standard_normal((3, cout, cin)) * np.sqrt(2.0 / (3 * cin)))
    b = (rng.standard_normal(cout) * 0.1).astype(np.float32) if bias else None
    cpad = (cout + 31) // 32 * 32
    Wp = np.zeros((3, cpad, cin))
    Wp[:, :cout] = Wt
    out = torch.full((npix, cout), 7.0, dtype=torch.bfloat16, device=DEV)
    xd, wd = to_dev_bf16(x), to_dev_bf16(Wp)
    bd = torch.from_numpy(b).to(DEV) if bias else None
    _lib.check(lib.mrh_unet_conv_fwd(xd.data_ptr(), wd.data_ptr(), bd.data_ptr() if bias else None,
                                     out.data_ptr(), ou.TAP_T3, nvol, F, H, W, cin, cout, cpad, int(relu), st()))
    ref = oracle_conv(x, Wt, None if b is None else b.astype(np.float64), ou.TAP_T3, nvol, F, relu).reshape(npix, cout)
    check_bf16_close(out.float().cpu().numpy().astype(np.float64), ref)


@pytest.mark.parametrize("kind,cin,cout", [(ou.TAP_3X3, 64, 32), (ou.TAP_T3, 32, 32), (ou.TAP_1X1, 256, 128),
                                           (ou.TAP_1X1, 32, 8), (ou.TAP_3X3, 32, 64)])
def test_conv_backward_data(kind, cin, cout):
    """dL/dx from the forward kernel with transposed, tap-flipped weights (as the network lays them out)."""
    lib = _lib.load()
    nvol, F, H, W = GEOM
    npix = nvol * F * H * W
    rng = np.random.default_rng(100 + cin + cout + kind)
    T = ou.NTAP[kind]
    x = rng.standard_normal((nvol * F, H, W, cin))
    Wt = bf16_np(rng.standard_normal((T, cout, cin)) * np.sqrt(2.0 / (T * cin)))
    g = bf16_np(rng.standard_normal((nvol * F, H, W, cout)))
    wbd = np.ascontiguousarray(Wt[::-1].transpose(0, 2, 1))        # [T-1-t][ci][co]
    gx = torch.empty((npix, cin), dtype=torch.bfloat16, device=DEV)
    gd, wd = to_dev_bf16(g), to_dev_bf16(wbd)
    _lib.check(lib.mrh_unet_conv_fwd(gd.data_ptr(), wd.data_ptr(), None, gx.data_ptr(),
                                     kind, nvol, F, H, W, cout, cin, cin, 0, st()))
    xt = torch.from_numpy(x).requires_grad_(True)
    y = ou.conv(ou.to_ncdhw(xt, nvol, F), torch.from_numpy(Wt), None, kind)
    (ou.to_channels_last(y) * torch.from_numpy(g)).sum().backward()
    check_bf16_close(gx.float().cpu().numpy().astype(np.float64), xt.grad.numpy().reshape(npix, cin))


@pytest.mark.parametrize("kind,cin,cout", [(ou.TAP_3X3, 8, 32), (ou.TAP_3X3, 64, 32), (ou.TAP_T3, 32, 64),
                                           (ou.TAP_1X1, 32, 8), (ou.TAP_T3, 256, 256)])
def test_conv_backward_weight_and_bias(kind, cin, cout):
    lib = _lib.load()
    nvol, F, H, W = GEOM
    npix = nvol * F * H * W
    rng = np.random.default_rng(200 + cin + cout + kind)
    T = ou.NTAP[kind]
    x = bf16_np(rng.standard_normal((nvol * F, H, W, cin)))
    g = bf16_np(rng.standard_normal((nvol * F, H, W, cout)))
    cpad = (cout + 31) // 32 * 32
    dw = torch.zeros(T * cpad * cin, dtype=torch.float32, device=DEV)
    db = torch.zeros(cpad, dtype=torch.float32, device=DEV)
    gd, xd = to_dev_bf16(g), to_dev_bf16(x)
    args = (gd.data_ptr(), xd.data_ptr(), dw.data_ptr(), db.data_ptr(), kind, nvol, F, H, W, cin, cout, cpad, st())
    _lib.check(lib.mrh_unet_conv_bwd_weight(*args))
    Wt = torch.zeros((T, cout, cin), dtype=torch.float64, requires_grad=True)
    bt = torch.zeros(cout, dtype=torch.float64, requires_grad=True)
    y = ou.conv(ou.to_ncdhw(torch.from_numpy(x), nvol, F), Wt, bt, kind)
    (ou.to_channels_last(y) * torch.from_numpy(g)).sum().backward()
    got = dw.cpu().numpy().reshape(T, cpad, cin)
    assert rel_l2(got[:, :cout], Wt.grad.numpy()) <= 1e-5
    assert np.all(got[:, cout:] == 0)
    assert rel_l2(db.cpu().numpy()[:cout], bt.grad.numpy()) <= 1e-5
    # the entry point accumulates (+=)
    _lib.check(lib.mrh_unet_conv_bwd_weight(*args))
    assert rel_l2(dw.cpu().numpy().reshape(T, cpad, cin)[:, :cout], 2 * Wt.grad.numpy()) <= 1e-5


def test_pool_upsample_concat_kernels():
    lib = _lib.load()
    nf, H, W, C1, C2 = 3, 8, 12, 32, 16
    rng = np.random.default_rng(5)
    e = bf16_np(rng.standard_normal((nf, H, W, C1)))
    p = torch.empty((nf, H // 2, W // 2, C1), dtype=torch.bfloat16, device=DEV)
    ed = to_dev_bf16(e)
    _lib.check(lib.mrh_unet_pool2(ed.data_ptr(), p.data_ptr(), nf, H, W, C1, st()))
    ref = e.reshape(nf, H // 2, 2, W // 2, 2, C1).mean(axis=(2, 4))
    check_bf16_close(p.float().cpu().numpy(), ref)
    # adjoint with an additive term
    gy = bf16_np(rng.standard_normal((nf, H // 2, W // 2, C1)))
    add = bf16_np(rng.standard_normal((nf, H, W, C1)))
    gx = torch.empty((nf, H, W, C1), dtype=torch.bfloat16, device=DEV)
    gyd, addd = to_dev_bf16(gy), to_dev_bf16(add)
    _lib.check(lib.mrh_unet_pool2_bwd(gyd.data_ptr(), addd.data_ptr(), gx.data_ptr(), nf, H, W, C1, st()))
    check_bf16_close(gx.float().cpu().numpy(), 0.25 * gy.repeat(2, axis=1).repeat(2, axis=2) + add)
    # up2 + concat and its adjoint
    s = bf16_np(rng.standard_normal((nf, H // 2, W // 2, C1)))
    sk = bf16_np(rng.standard_normal((nf, H, W, C2)))
    u = torch.empty((nf, H, W, C1 + C2), dtype=torch.bfloat16, device=DEV)
    sd, skd = to_dev_bf16(s), to_dev_bf16(sk)
    _lib.check(lib.mrh_unet_up2_concat(sd.data_ptr(), skd.data_ptr(), u.data_ptr(), nf, H, W, C1, C2, st()))
    ref_u = np.concatenate([s.repeat(2, axis=1).repeat(2, axis=2), sk], axis=-1)
    assert np.array_equal(u.float().cpu().numpy(), ref_u)
    gu = bf16_np(rng.standard_normal((nf, H, W, C1 + C2)))
    gs = torch.empty((nf, H // 2, W // 2, C1), dtype=torch.bfloat16, device=DEV)
    gk = torch.empty((nf, H, W, C2), dtype=torch.bfloat16, device=DEV)
    gud = to_dev_bf16(gu)
    _lib.check(lib.mrh_unet_up2_concat_bwd(gud.data_ptr(), gs.data_ptr(), gk.data_ptr(), nf, H, W, C1, C2, st()))
    check_bf16_close(gs.float().cpu().numpy(), gu[..., :C1].reshape(nf, H // 2, 2, W // 2, 2, C1).sum(axis=(2, 4)))
    assert np.array_equal(gk.float().cpu().numpy(), gu[..., C1:])


def _random_net(levels, seed):
    net = UNet2DT(levels=levels, base=32, device=DEV, seed=seed)
    rng = np.random.default_rng(seed + 1)
    params = net.parameters()
    for name, (W, b) in params.items():
        if name == "out":
            W[:] = rng.standard_normal(W.shape) * 0.2
        if name != "enc0_s":
            b[:] = rng.standard_normal(b.shape) * 0.05
    net.set_parameters(params)
    return net


def _oracle_params(net):
    out = {}
    for name, (W, b) in net.parameters().items():
        out[name] = (torch.tensor(bf16_np(W), dtype=torch.float64, requires_grad=True),
                     torch.tensor(b.astype(np.float64), requires_grad=True))
    return out


def _smooth_pair(nvol, F, H, W, seed):
    x_t = np.concatenate([synth.dynamic_phantom(H, F, seed=seed + v, n_ellipses=5) for v in range(nvol)])
    rng = np.random.default_rng(seed)
    noise = (rng.standard_normal(x_t.shape) + 1j * rng.standard_normal(x_t.shape)) * 0.2
    x_adj = (0.8 * x_t + noise).astype(np.complex64)
    return x_adj, x_t.astype(np.complex64)


def test_network_forward_loss_and_gradients_match_oracle():
    levels, nvol, F, H, W = 3, 2, 4, 16, 16
    net = _random_net(levels, seed=3)
    x_adj, x_t = _smooth_pair(nvol, F, H, W, seed=40)
    xa, xt = torch.from_numpy(x_adj).to(DEV), torch.from_numpy(x_t).to(DEV)
    s = 1.3
    loss = float(net.loss_and_grad(xa, xt, nvol=nvol, scales=[s, s]).item())
    xw = net._workspace(nvol, F, H, W, train=True)["x_warm"].cpu().numpy()
    params = _oracle_params(net)
    xa64, xt64 = torch.from_numpy(x_adj.astype(np.complex128)), torch.from_numpy(x_t.astype(np.complex128))
    out = ou.forward(ou.bf16_round(ou.pack_input(xa64, s)), params, nvol, F, levels, 32, round_bf16=True)
    ref_loss = ou.mse_loss(out, xa64, xt64, s)
    ref_loss.backward()
    ref_xw = ou.warm_start(out.detach(), xa64, s).numpy()
    errs = {"x_warm": rel_l2(xw - x_adj, ref_xw - x_adj), "loss": abs(loss - ref_loss.item()) / ref_loss.item()}
    grads = net.grads
    for L in net.layers:
        gw = L.bufs["gw"].reshape(L.ntap, L.cout_pad, L.cin)[:, :L.cout].cpu().numpy()
        gb = L.bufs["gb"][:L.cout].cpu().numpy()
        W, b = params[L.name]
        errs[L.name + ".w"] = rel_l2(gw, W.grad.numpy())
        errs[L.name + ".b"] = rel_l2(gb, b.grad.numpy())
    assert grads.abs().sum().item() > 0
    print("network errors:", {k: f"{v:.2e}" for k, v in errs.items()})
    bad = {k: v for k, v in errs.items() if not v <= 1e-2}
    assert not bad, bad


def test_training_reduces_loss_and_inference_is_deterministic(tmp_path):
    levels, nvol, F, H, W = 3, 1, 6, 32, 32
    net = UNet2DT(levels=levels, base=32, device=DEV, seed=9)
    x_adj, x_t = _smooth_pair(nvol, F, H, W, seed=50)
    xa, xt = torch.from_numpy(x_adj).to(DEV), torch.from_numpy(x_t).to(DEV)
    # untrained: zero output layer -> identity warm start
    assert torch.equal(net.infer(xa).clone(), xa)
    losses = [float(net.train_step(xa, xt, lr=2e-3).item()) for _ in range(60)]
    assert losses[-1] < 0.3 * losses[0], (losses[0], losses[-1])
    a = net.infer(xa).clone()
    b = net.infer(xa).clone()
    assert torch.equal(a, b)
    p = tmp_path / "w.npz"
    net.save(p)
    net2 = UNet2DT.load(p, device=DEV)
    assert torch.equal(net2.infer(xa).clone(), a)
    assert net2.step == net.step == 60


@pytest.mark.parametrize("geom,c", [((1, 2, 8, 64), 32), ((2, 2, 16, 32), 64), ((1, 1, 8, 128), 128)])
@pytest.mark.parametrize("cout", [32, 64])
def test_fused_upsample_concat_conv_equals_the_stored_concatenation(geom, c, cout):
    """The decoder's 3x3 convolution reading concat(up2(src), skip) in place gives the same bits as
    storing the concatenation first; unsupported shapes are refused."""
    lib = _lib.load()
    nvol, F, H, W = geom
    nf = nvol * F
    assert lib.mrh_unet_upcat_supported(H, W, c, c) == 1
    rng = np.random.default_rng(c + cout + W)
    src = to_dev_bf16(rng.standard_normal((nf, H // 2, W // 2, c)))
    skip = to_dev_bf16(rng.standard_normal((nf, H, W, c)))
    cpad = (cout + 31) // 32 * 32
    wd = to_dev_bf16(rng.standard_normal((9, cpad, 2 * c)) * np.sqrt(1.0 / (9 * c)))
    bd = torch.from_numpy((rng.standard_normal(cpad) * 0.1).astype(np.float32)).to(DEV)
    u = torch.empty((nf * H * W, 2 * c), dtype=torch.bfloat16, device=DEV)
    _lib.check(lib.mrh_unet_up2_concat(src.data_ptr(), skip.data_ptr(), u.data_ptr(), nf, H, W, c, c, st()))
    ref = torch.empty((nf * H * W, cout), dtype=torch.bfloat16, device=DEV)
    _lib.check(lib.mrh_unet_conv_fwd(u.data_ptr(), wd.data_ptr(), bd.data_ptr(), ref.data_ptr(), ou.TAP_3X3,
                                     nvol, F, H, W, 2 * c, cout, cpad, 1, st()))
    out = torch.full_like(ref, 7.0)
    _lib.check(lib.mrh_unet_conv3x3_upcat(src.data_ptr(), skip.data_ptr(), wd.data_ptr(), bd.data_ptr(),
                                          out.data_ptr(), nvol, F, H, W, c, c, cout, cpad, 1, st()))
    assert torch.equal(out, ref)
    assert lib.mrh_unet_upcat_supported(18, 18, c, c) == 0 and lib.mrh_unet_upcat_supported(H, W, 48, c) == 0
    rc = lib.mrh_unet_conv3x3_upcat(src.data_ptr(), skip.data_ptr(), wd.data_ptr(), bd.data_ptr(),
                                    out.data_ptr(), nvol, F, H, W, 48, c, cout, cpad, 1, st())
    with pytest.raises(NotImplementedError):
        _lib.check(rc)


def test_inference_reading_the_concatenation_in_place_equals_the_training_forward():
    levels, nvol, F, H, W = 3, 1, 3, 64, 64       # levels 0 and 1 fused, level 2 (16 x 16) stored
    net = UNet2DT(levels=levels, base=32, device=DEV, seed=5)
    rng = np.random.default_rng(6)
    for L in net.layer.values():          # random BF16 weights, a non-zero output layer included
        L.bufs["wf"].copy_((torch.randn_like(L.bufs["wf"].float()) * 0.05).to(torch.bfloat16))
    x_adj = (rng.standard_normal((nvol * F, H, W)) + 1j * rng.standard_normal((nvol * F, H, W))).astype(np.complex64)
    xa = torch.from_numpy(x_adj).to(DEV)
    a = net.infer(xa, scales=[1.0]).clone()
    ws = net._workspace(nvol, F, H, W, train=True)
    net._pack(ws, xa, [1.0], F)
    net._forward(ws, nvol, F, H, W, train=True)
    b = torch.empty_like(xa)
    net._unpack(ws, xa, b, [1.0], F)
    assert not torch.equal(a, xa)
    assert torch.equal(a, b)


@pytest.mark.parametrize("geom", [(2, 7, 32, 64), (1, 34, 64, 32), (1, 1, 8, 32), (1, 5, 256, 256)])
def test_fused_full_resolution_level_equals_the_separate_layers(geom):
    """Inference with the full-resolution level fused (enc0 3x3 + temporal + pool from the
    complex input; dec0 temporal + 1x1 output + residual into the complex output) gives the
    bits of the layer-by-layer forward: one frame per workgroup group and several frame
    groups per volume (256^2: two groups), two volumes with their own scales, a single frame."""
    nvol, F, H, W = geom
    lib = _lib.load()
    assert lib.mrh_unet_level0_fused_supported(H, W) == 1
    net = UNet2DT(levels=3, base=32, device=DEV, seed=11)
    rng = np.random.default_rng(F + H)
    for L in net.layer.values():
        L.bufs["wf"].copy_((torch.randn_like(L.bufs["wf"].float()) * 0.08).to(torch.bfloat16))
        L.bufs["b"].copy_(torch.randn_like(L.bufs["b"]) * 0.05)
    shape = (nvol * F, H, W)
    x_adj = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    xa = torch.from_numpy(x_adj).to(DEV)
    scales = [0.7, 1.3][:nvol]
    a = net.infer(xa, nvol=nvol, scales=scales, fused=True).clone()
    b = net.infer(xa, nvol=nvol, scales=scales, fused=False).clone()
    assert not torch.equal(a, xa)
    assert torch.equal(a, b)


def test_fused_full_resolution_level_refuses_other_shapes():
    lib = _lib.load()
    assert lib.mrh_unet_level0_fused_supported(16, 48) == 0 and lib.mrh_unet_level0_fused_supported(12, 32) == 0
    net = UNet2DT(levels=3, base=32, device=DEV)
    assert not net._level0_fused(16, 48)
    x = torch.zeros((2, 16, 48), dtype=torch.complex64, device=DEV)
    e = torch.empty((2 * 16 * 48, 32), dtype=torch.bfloat16, device=DEV)
    Ls, Lt = net.layer["enc0_s"], net.layer["enc0_t"]
    rc = lib.mrh_unet_enc0_fused(x.data_ptr(), 1.0, Ls.bufs["wf"].data_ptr(), Ls.bufs["b"].data_ptr(),
                                 Lt.bufs["wf"].data_ptr(), Lt.bufs["b"].data_ptr(), e.data_ptr(), e.data_ptr(),
                                 2, 16, 48, st())
    with pytest.raises(NotImplementedError):
        _lib.check(rc)
    assert torch.equal(net.infer(x), x)   # the separate layers still run (zero output layer: identity)


def test_shape_errors():
    net = UNet2DT(levels=3, base=32, device=DEV)
    with pytest.raises(ValueError):
        net.infer(torch.zeros((4, 18, 18), dtype=torch.complex64, device=DEV))   # 18 not divisible by 4
    with pytest.raises(ValueError):
        net.infer(torch.zeros((5, 16, 16), dtype=torch.complex64, device=DEV), nvol=2)
    with pytest.raises(ValueError):
        UNet2DT(levels=3, base=16, device=DEV)
    lib = _lib.load()
    with pytest.raises(ValueError):
        _lib.check(lib.mrh_unet_conv_fwd(1, 1, None, 1, 3, 1, 1, 8, 8, 8, 32, 32, 0, st()))
    with pytest.raises(ValueError):
        _lib.check(lib.mrh_unet_conv_fwd(1, 1, None, 1, 1, 1, 1, 8, 8, 12, 32, 32, 0, st()))


def test_fallback_rejects_a_worse_warm_start():
    from mrirecon_amd import RegManager, TVND, TVParams
    from mrirecon_amd.nufft.backends.hip import HipNUFFTAdapter
    P, F, S, C = 2, 3, 64, 4
    traj, dcf, off = synth.radial_frames([13] * (P * F), 128)
    x_true = np.concatenate([synth.dynamic_phantom(S, F, seed=60 + p, n_ellipses=5) for p in range(P)])
    maps = synth.gaussian_maps(C, S)
    be = HipNUFFTAdapter(ndim=2, traj_units="rad", kb_width=6, osf=2.0)
    be.prepare((C, S, S), torch.from_numpy(traj), torch.from_numpy(dcf),
               torch.from_numpy(maps.astype(np.complex64)).to(DEV), device=DEV, k_offsets=off)
    plan = be.plan
    xt = torch.from_numpy(x_true).to(DEV)
    y = torch.empty(plan.n_k, dtype=torch.complex64, device=DEV)
    plan.forward(xt, y)
    x_adj = torch.empty_like(xt)
    plan.adjoint(y, x_adj)
    regm = RegManager([TVND("t", TVParams(weight=0.02, axes="temporal")),
                       TVND("s", TVParams(weight=0.002, axes="spatial"))])
    obj = ObjectiveEval(plan, y, regm, problems=P, frames=F)
    # problem 0 gets the truth (better than the adjoint), problem 1 a blown-up image (worse)
    x_warm = x_adj.clone()
    x_warm[:F] = xt[:F]
    x_warm[F:] = 5.0 * x_adj[F:]
    x0, acc, f_w, f_a = select_warm_start(obj, x_adj, x_warm)
    assert acc.tolist() == [True, False]
    assert f_w[0] < f_a[0] and f_w[1] > f_a[1]
    assert torch.equal(x0[:F], xt[:F]) and torch.equal(x0[F:], x_adj[F:])
    # objective value = 1/2 |A x - y|^2 + TV: zero data term at the truth
    e = regm.energy_and_grad_tensors(xt[:F].contiguous(), None).cpu().numpy()
    assert abs(f_w[0] - e[0]) <= 1e-6 * abs(e[0])


def test_objective_gradient_and_loss_term():
    """ObjectiveGrad returns the engine's gradient A^H (A x - y) + grad TV and f; the loss
    kernel adds coef * (Re g, Im g) to the MSE gradient."""
    from mrirecon_amd import NLCGEngine, RegManager, TVND, TVParams
    from mrirecon_amd.nufft.backends.hip import HipNUFFTAdapter
    from mrirecon_amd.warmstart import ObjectiveGrad
    F, S, C = 4, 64, 4
    traj, dcf, off = synth.radial_frames([13] * F, 128)
    x_true = synth.dynamic_phantom(S, F, seed=70, n_ellipses=5)
    be = HipNUFFTAdapter(ndim=2, traj_units="rad", kb_width=6, osf=2.0)
    be.prepare((C, S, S), torch.from_numpy(traj), torch.from_numpy(dcf),
               torch.from_numpy(synth.gaussian_maps(C, S).astype(np.complex64)).to(DEV), device=DEV, k_offsets=off)
    plan = be.plan
    xt = torch.from_numpy(x_true).to(DEV)
    y = torch.empty(plan.n_k, dtype=torch.complex64, device=DEV)
    plan.forward(xt, y)
    y += 0.01 * y.abs().max() * torch.randn(y.shape, dtype=torch.complex64, device=DEV, generator=torch.Generator(DEV).manual_seed(1))
    x = torch.empty_like(xt)
    plan.adjoint(y, x)
    regm = RegManager([TVND("t", TVParams(weight=0.02, axes="temporal")), TVND("s", TVParams(weight=0.002, axes="spatial"))])
    og = ObjectiveGrad(plan, regm, F)
    g = og(x, y).clone()
    eng = NLCGEngine(plan, y, problems=1, frames=F, regm=regm, direction="prplus")
    eng.prime(x)
    torch.cuda.synchronize()
    assert rel_l2(g.cpu().numpy(), eng.g.cpu().numpy()) <= 1e-6
    f = og.value()
    obj = ObjectiveEval(plan, y, regm, problems=1, frames=F)
    assert abs(f - obj(x)[0]) <= 1e-9 * abs(f)
    # loss kernel with the extra gradient term
    lib = _lib.load()
    xt_img = torch.zeros_like(x)
    gnet = torch.empty((x.numel(), 8), dtype=torch.bfloat16, device=DEV)
    loss = torch.zeros(1, dtype=torch.float64, device=DEV)
    s, w, coef = 0.7, 1.0 / x.numel(), 0.3
    _lib.check(lib.mrh_unet_mse_grad(x.data_ptr(), xt_img.data_ptr(), g.data_ptr(), coef, gnet.data_ptr(),
                                     loss.data_ptr(), x.numel(), 8, s, w, st()))
    xn, gn = x.cpu().numpy().reshape(-1), g.cpu().numpy().reshape(-1)
    ref = np.stack([2 * w * xn.real * s + coef * gn.real, 2 * w * xn.imag * s + coef * gn.imag], axis=1)
    got = gnet.float().cpu().numpy()
    check_bf16_close(got[:, :2].astype(np.float64), ref.astype(np.float64))
    assert np.all(got[:, 2:] == 0)
    assert abs(loss.item() - w * s * s * np.sum(np.abs(xn) ** 2)) <= 1e-5 * loss.item()


def test_trained_warm_start_gain_on_held_out_grasp_volumes():
    """What the trained network gains: on held-out GRASP volumes (seeds outside every training
    set) the warm start is accepted, starts far below the adjoint's objective, and NLCG from it
    reaches the plain run's final objective in fewer iterations than the plain run took."""
    from mrirecon_amd.train_warmstart import GraspProblem, _phantom, run_plain, run_warm
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    net = UNet2DT.load(os.path.join(root, "mrirecon_amd", "weights", "unet_grasp.npz"), device=DEV)
    pb = GraspProblem()
    for seed in (640_001, 640_002):
        pb.measure(_phantom(seed), noise_seed=seed + 7)
        p = run_plain(pb)
        w = run_warm(pb, net, target=p["f_final"] * (1 + 1e-4))
        print(f"[gain] seed {seed}: plain {p['iterations']} it f={p['f_final']:.6g}; warm start f={w['f_start']:.6g} "
              f"(adjoint {w['f_adjoint']:.6g}), reached the plain objective in {w['iterations']} it ({w['stop']})")
        assert w["accepted"] and w["f_start"] < 0.2 * w["f_adjoint"]
        assert w["stop"] == "target"
        assert w["iterations"] < p["iterations"]


@pytest.mark.parametrize("bound", [False, True])
def test_overlapped_warm_start_equals_sequential_selection(bound):
    """warm_start (network on the main stream, f(A^H y) -- or, with ``bound``, its NUFFT-free
    lower bound -- on a side stream, the choice made on the device) chooses what
    select_warm_start chooses for the same network output, and leaves A x0 in the objective's
    Ax buffer for rejected and accepted starts alike.  With the bound, f_adj is exact where the
    bound does not decide and otherwise a lower bound of it that f_warm stays under."""
    from mrirecon_amd import RegManager, TVND, TVParams
    from mrirecon_amd.nufft.backends.hip import HipNUFFTAdapter
    from mrirecon_amd.warmstart import launch_warm_start, warm_start
    P, F, S, C = 2, 4, 64, 4
    traj, dcf, off = synth.radial_frames([13] * (P * F), 128)
    x_true = np.concatenate([synth.dynamic_phantom(S, F, seed=90 + p, n_ellipses=5) for p in range(P)])
    be = HipNUFFTAdapter(ndim=2, traj_units="rad", kb_width=6)
    be.prepare((C, S, S), torch.from_numpy(traj), torch.from_numpy(dcf),
               torch.from_numpy(synth.gaussian_maps(C, S).astype(np.complex64)).to(DEV), device=DEV, k_offsets=off)
    plan = be.plan
    y = torch.empty(plan.n_k, dtype=torch.complex64, device=DEV)
    plan.forward(torch.from_numpy(x_true).to(DEV), y)
    x_adj = torch.empty((P * F, S, S), dtype=torch.complex64, device=DEV)
    plan.adjoint(y, x_adj)
    regm = RegManager([TVND("t", TVParams(weight=0.02, axes="temporal"))])
    net = _random_net(3, seed=11)
    o1, o2 = ObjectiveEval(plan, y, regm, P, F), ObjectiveEval(plan, y, regm, P, F)
    x0, acc, f_w, f_a = warm_start(net, o1, o2, x_adj, nvol=P, out=torch.empty_like(x_adj), bound=bound)
    x0 = x0.clone()
    xw = net.infer(x_adj, nvol=P).clone()
    r0, racc, rf_w, rf_a = select_warm_start(ObjectiveEval(plan, y, regm, P, F), x_adj, xw, out=torch.empty_like(x_adj))
    assert torch.equal(x0, r0) and acc.tolist() == racc.tolist()
    assert np.allclose(f_w, rf_w, rtol=1e-12)
    if bound:
        exact = np.isclose(f_a, rf_a, rtol=1e-12)
        assert np.all(exact | ((f_a <= rf_a) & (f_w <= f_a * (1 - 1e-4)))), (f_w, f_a, rf_a)
    else:
        assert np.allclose(f_a, rf_a, rtol=1e-12)
    Ax = torch.empty_like(y)
    plan.forward(x0, Ax)
    assert torch.equal(o1.Ax, Ax), "objective.Ax must hold A x0 whichever start each problem took"
    # second call on the same evaluators, with a network whose output layer is zero:
    # x_warm == x_adj, so f_warm == f_adj and every problem keeps the warm start
    params = net.parameters()
    params["out"][0][:] = 0.0
    params["out"][1][:] = 0.0
    net.set_parameters(params)
    ch = launch_warm_start(net, o1, o2, x_adj, nvol=P, out=torch.empty_like(x_adj), bound=bound)
    acc2, f_w2, f_a2 = ch.read()
    assert acc2.all() and np.array_equal(f_w2, f_a2) and torch.equal(ch.x0, x_adj)
    plan.forward(ch.x0, Ax)
    assert ch.Ax0 is o1.Ax and torch.equal(o1.Ax, Ax)

