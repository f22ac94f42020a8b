"""Denoiser kernels (BF16 MFMA, FP32 accumulation) against plain PyTorch FP32
references of the same ops on the same bf16-rounded inputs, and the whole
U-Net (forward, gradients, a few training steps) against the PyTorch
restatement of its graph."""
import ctypes as C

import numpy as np
import pytest
import torch
import torch.nn.functional as F

import offlinerenderer_amd as wr
from offlinerenderer_amd import denoiser, scenes
from tests import gpu_util as G
from tests import unet_ref as U

pytestmark = pytest.mark.gpu


def _dev_bits(a):
    return torch.from_numpy(U.bf16_bits(a)).cuda()


def _host_bits(t):
    torch.cuda.synchronize()
    return U.from_bits(t.cpu().numpy())


def _rand(rng, shape, scale=1.0):
    return U.bf16_round(rng.normal(size=shape).astype(np.float32) * scale)


@pytest.mark.parametrize("n,h,w,cin,cout,relu,xpitch,ypitch", [
    (2, 16, 32, 32, 32, 1, 32, 32),
    (1, 12, 20, 96, 64, 0, 128, 96),     # partial tiles, pitched input and output slices
    (1, 24, 16, 224, 96, 1, 224, 96),
    (3, 8, 8, 64, 128, 1, 64, 128),
    (2, 264, 528, 32, 32, 1, 64, 96),    # 32 input channels: several tiles per block, channel slices
    (1, 60, 100, 32, 64, 0, 32, 64),     # 32 input channels, 64 output channels, partial tiles
])
def test_conv_fwd_matches_torch(gpu, n, h, w, cin, cout, relu, xpitch, ypitch):
    rng = np.random.default_rng(cin * 7 + cout)
    x = _rand(rng, (n, h, w, xpitch))
    wt = _rand(rng, (9, cout, cin), 1.0 / np.sqrt(9 * cin))
    b = rng.normal(size=cout).astype(np.float32) * 0.1
    xd, wd = _dev_bits(x), _dev_bits(wt)
    bd = torch.from_numpy(b).cuda()
    yd = torch.zeros((n, h, w, ypitch), dtype=torch.int16, device="cuda")
    G.check(wr.lib().wrk_conv3x3_fwd(G.vp(xd), xpitch, cin, G.vp(wd), G.vp(bd), G.vp(yd), ypitch, cout, relu, n, h, w,
                                     G.stream()))
    got = _host_bits(yd)[..., :cout]
    ref = F.conv2d(torch.from_numpy(x[..., :cin]).permute(0, 3, 1, 2), U.torch_weight(wt.ravel(), cin, cout),
                   torch.from_numpy(b), padding=1)
    if relu:
        ref = F.relu(ref)
    ref = ref.permute(0, 2, 3, 1).numpy()
    # output rounded to bf16 (2^-8 relative); accumulation order differs
    err = np.abs(got - ref)
    assert np.all(err <= np.abs(ref) * 2.0 ** -8 + 1e-4 * np.abs(ref).max() + 1e-6), err.max()
    if ypitch > cout:
        assert np.all(_host_bits(yd)[..., cout:] == 0)


@pytest.mark.parametrize("cin,cout,used", [(32, 32, 3), (96, 32, 20), (64, 64, 16)])
def test_conv_fwd_first_channels(gpu, cin, cout, used):
    """wrk_conv3x3_fwd_first computes output channels [0, ceil16(used)) of a
    convolution with cout output channels, as wrk_conv3x3_fwd does, and
    leaves the rest of y untouched."""
    n, h, w = 2, 40, 72
    rng = np.random.default_rng(cin + cout + used)
    x = _rand(rng, (n, h, w, cin))
    wt = _rand(rng, (9, cout, cin), 1.0 / np.sqrt(9 * cin))
    b = rng.normal(size=cout).astype(np.float32) * 0.1
    xd, wd, bd = _dev_bits(x), _dev_bits(wt), torch.from_numpy(b).cuda()
    full = torch.zeros((n, h, w, cout), dtype=torch.int16, device="cuda")
    G.check(wr.lib().wrk_conv3x3_fwd(G.vp(xd), cin, cin, G.vp(wd), G.vp(bd), G.vp(full), cout, cout, 1, n, h, w,
                                     G.stream()))
    part = torch.full((n, h, w, cout), 0x7F7F, dtype=torch.int16, device="cuda")
    G.check(wr.lib().wrk_conv3x3_fwd_first(G.vp(xd), cin, cin, G.vp(wd), G.vp(bd), G.vp(part), cout, cout, used, 1,
                                           n, h, w, G.stream()))
    torch.cuda.synchronize()
    c = min(cout, (used + 15) // 16 * 16)
    f, p_ = full.cpu().numpy(), part.cpu().numpy()
    assert np.array_equal(p_[..., :c], f[..., :c])
    assert np.all(p_[..., c:] == 0x7F7F)


@pytest.mark.parametrize("n,h,w,cup,cskip,cout,uppitch", [(2, 36, 52, 64, 32, 32, 64), (1, 24, 40, 96, 64, 64, 128),
                                                           (1, 20, 36, 128, 96, 96, 128)])
def test_conv_fwd_up_equals_upsample_then_conv(gpu, n, h, w, cup, cskip, cout, uppitch):
    """wrk_conv3x3_fwd_up reads the upsampled channels of a decoder input
    from the half-resolution tensor: bit for bit wrk_up2_fwd into the
    concatenation followed by wrk_conv3x3_fwd (partial tiles, pitches)."""
    cin = cup + cskip
    rng = np.random.default_rng(cin + cout)
    low = _rand(rng, (n, h // 2, w // 2, uppitch))
    cat = _rand(rng, (n, h, w, cin))
    wt = _rand(rng, (9, cout, cin), 1.0 / np.sqrt(9 * cin))
    b = rng.normal(size=cout).astype(np.float32) * 0.1
    lowd, wd, bd = _dev_bits(low), _dev_bits(wt), torch.from_numpy(b).cuda()
    ref_in = _dev_bits(cat)
    G.check(wr.lib().wrk_up2_fwd(G.vp(lowd), uppitch, cup, G.vp(ref_in), cin, n, h, w, G.stream()))
    ref = torch.zeros((n, h, w, cout), dtype=torch.int16, device="cuda")
    G.check(wr.lib().wrk_conv3x3_fwd(G.vp(ref_in), cin, cin, G.vp(wd), G.vp(bd), G.vp(ref), cout, cout, 1, n, h, w,
                                     G.stream()))
    # the fused form never reads the first cup channels of its input
    junk = cat.copy()
    junk[..., :cup] = np.float32(1e4)
    got_in = _dev_bits(junk)
    got = torch.zeros((n, h, w, cout), dtype=torch.int16, device="cuda")
    G.check(wr.lib().wrk_conv3x3_fwd_up(G.vp(got_in), cin, cin, G.vp(lowd), uppitch, cup, G.vp(wd), G.vp(bd),
                                        G.vp(got), cout, cout, 1, n, h, w, G.stream()))
    torch.cuda.synchronize()
    assert np.array_equal(got.cpu().numpy(), ref.cpu().numpy())
    assert np.count_nonzero(ref.cpu().numpy()) > ref.numel() // 4


@pytest.mark.parametrize("n,h,w,cin,cout,ypitch,ppitch", [(2, 36, 52, 32, 32, 96, 32), (1, 24, 40, 64, 64, 160, 64),
                                                          (1, 20, 36, 96, 96, 224, 96), (1, 18, 34, 32, 64, 64, 64)])
def test_conv_fwd_pool_equals_conv_then_pool(gpu, n, h, w, cin, cout, ypitch, ppitch):
    """wrk_conv3x3_fwd_pool writes the convolution and its 2x2 average pool
    from one epilogue: bit for bit wrk_conv3x3_fwd then wrk_pool2_fwd (both
    kernels: 32 and more input channels; partial tiles; pitched slices)."""
    rng = np.random.default_rng(cin + cout + h)
    x = _rand(rng, (n, h, w, cin))
    wt = _rand(rng, (9, cout, cin), 1.0 / np.sqrt(9 * cin))
    b = rng.normal(size=cout).astype(np.float32) * 0.1
    xd, wd, bd = _dev_bits(x), _dev_bits(wt), torch.from_numpy(b).cuda()
    y_ref = torch.zeros((n, h, w, ypitch), dtype=torch.int16, device="cuda")
    p_ref = torch.zeros((n, h // 2, w // 2, ppitch), dtype=torch.int16, device="cuda")
    G.check(wr.lib().wrk_conv3x3_fwd(G.vp(xd), cin, cin, G.vp(wd), G.vp(bd), G.vp(y_ref), ypitch, cout, 1, n, h, w,
                                     G.stream()))
    G.check(wr.lib().wrk_pool2_fwd(G.vp(y_ref), ypitch, cout, G.vp(p_ref), ppitch, n, h, w, G.stream()))
    y = torch.zeros_like(y_ref)
    pooled = torch.zeros_like(p_ref)
    G.check(wr.lib().wrk_conv3x3_fwd_pool(G.vp(xd), cin, cin, G.vp(wd), G.vp(bd), G.vp(y), ypitch, cout, 1,
                                          G.vp(pooled), ppitch, n, h, w, G.stream()))
    torch.cuda.synchronize()
    assert np.array_equal(y.cpu().numpy(), y_ref.cpu().numpy())
    assert np.array_equal(pooled.cpu().numpy(), p_ref.cpu().numpy())
    assert np.count_nonzero(p_ref.cpu().numpy()) > p_ref.numel() // 8


@pytest.mark.parametrize("n,h,w", [(2, 36, 52), (1, 1080 // 8, 1920 // 8)])
def test_conv_fwd_unpack_equals_conv_then_unpack(gpu, n, h, w):
    """wrk_conv3x3_fwd_unpack: the residual head's output convolution and
    wrk_unet_unpack from one epilogue, bit for bit the two steps."""
    cin, cout = 32, 32
    rng = np.random.default_rng(h * w)
    x = _rand(rng, (n, h, w, cin))
    wt = _rand(rng, (9, cout, cin), 0.3 / np.sqrt(9 * cin))
    b = rng.normal(size=cout).astype(np.float32) * 0.1
    base = _rand(rng, (n, h, w, 32), 0.5)
    xd, wd, bd, based = _dev_bits(x), _dev_bits(wt), torch.from_numpy(b).cuda(), _dev_bits(base)
    pred = torch.zeros((n, h, w, cout), dtype=torch.int16, device="cuda")
    G.check(wr.lib().wrk_conv3x3_fwd_first(G.vp(xd), cin, cin, G.vp(wd), G.vp(bd), G.vp(pred), cout, cout, 3, 0, n, h,
                                           w, G.stream()))
    ref = torch.zeros((n * h * w, 3), dtype=torch.float32, device="cuda")
    G.check(wr.lib().wrk_unet_unpack(G.vp(pred), cout, G.vp(based), 32, G.vp(ref), n * h * w, G.stream()))
    got = torch.full((n * h * w, 3), -1.0, dtype=torch.float32, device="cuda")
    G.check(wr.lib().wrk_conv3x3_fwd_unpack(G.vp(xd), cin, cin, G.vp(wd), G.vp(bd), cout, G.vp(based), 32, G.vp(got),
                                            n, h, w, G.stream()))
    torch.cuda.synchronize()
    r, g = ref.cpu().numpy(), got.cpu().numpy()
    assert np.array_equal(g.view(np.uint32), r.view(np.uint32))
    assert np.count_nonzero(r) > r.size // 4


@pytest.mark.parametrize("n,h,w,cin,cout", [(2, 16, 32, 32, 32), (1, 12, 40, 96, 64), (1, 8, 16, 224, 96)])
def test_conv_gradients_match_torch(gpu, n, h, w, cin, cout):
    rng = np.random.default_rng(cin + 3 * cout)
    x = _rand(rng, (n, h, w, cin))
    dz = _rand(rng, (n, h, w, cout), 0.1)
    w32 = (rng.normal(size=(9, cout, cin)) / np.sqrt(9 * cin)).astype(np.float32)
    xd, dzd = _dev_bits(x), _dev_bits(dz)
    dw = torch.zeros(9 * cout * cin, dtype=torch.float32, device="cuda")
    db = torch.zeros(cout, dtype=torch.float32, device="cuda")
    nbytes = wr.lib().wrk_conv3x3_wgrad_scratch_bytes(cin, cout, n, h, w)
    scratch = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    G.check(wr.lib().wrk_conv3x3_wgrad(G.vp(xd), cin, cin, G.vp(dzd), cout, cout, n, h, w, G.vp(dw), G.vp(db),
                                       G.vp(scratch), nbytes, G.stream()))
    xt = torch.from_numpy(x).permute(0, 3, 1, 2)
    dzt = torch.from_numpy(dz).permute(0, 3, 1, 2)
    ref_w = torch.nn.grad.conv2d_weight(xt, (cout, cin, 3, 3), dzt, padding=1)  # [co][ci][3][3]
    ref_w = ref_w.permute(2, 3, 0, 1).reshape(-1).numpy()
    got_w = dw.cpu().numpy()
    assert np.abs(got_w - ref_w).max() <= 1e-4 * np.abs(ref_w).max()
    assert np.allclose(db.cpu().numpy(), dzt.sum(dim=(0, 2, 3)).numpy(), rtol=1e-4, atol=1e-3)
    # data gradient: the forward kernel with the tap-flipped transposed weights
    w32d = torch.from_numpy(w32.ravel()).cuda()
    wb = torch.zeros(9 * cout * cin, dtype=torch.int16, device="cuda")
    wtb = torch.zeros(9 * cout * cin, dtype=torch.int16, device="cuda")
    G.check(wr.lib().wrk_weights_bf16(G.vp(w32d), cout, cin, G.vp(wb), G.vp(wtb), G.stream()))
    dx = torch.zeros((n, h, w, cin), dtype=torch.int16, device="cuda")
    G.check(wr.lib().wrk_conv3x3_fwd(G.vp(dzd), cout, cout, G.vp(wtb), None, G.vp(dx), cin, cin, 0, n, h, w,
                                     G.stream()))
    wq = U.bf16_round(w32)
    gw = _host_bits(wb).reshape(9, cout, cin)
    bad = np.argwhere(gw != wq)
    assert bad.size == 0, (len(bad), [(float(w32[tuple(i)]), float(gw[tuple(i)]), float(wq[tuple(i)])) for i in bad[:4]])
    ref_x = torch.nn.grad.conv2d_input((n, cin, h, w), U.torch_weight(wq.ravel(), cin, cout), dzt, padding=1)
    ref_x = ref_x.permute(0, 2, 3, 1).numpy()
    got_x = _host_bits(dx)
    assert np.all(np.abs(got_x - ref_x) <= np.abs(ref_x) * 2.0 ** -8 + 1e-4 * np.abs(ref_x).max() + 1e-6)


def test_pool_upsample_relu(gpu):
    rng = np.random.default_rng(5)
    n, h, w, c = 2, 8, 12, 32
    x = _rand(rng, (n, h, w, 48))                      # pitched: 32 of 48 channels
    xd = _dev_bits(x)
    y = torch.zeros((n, h // 2, w // 2, c), dtype=torch.int16, device="cuda")
    G.check(wr.lib().wrk_pool2_fwd(G.vp(xd), 48, c, G.vp(y), c, n, h, w, G.stream()))
    ref = x[..., :c].reshape(n, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4))
    assert np.allclose(_host_bits(y), U.bf16_round(ref), rtol=2 ** -8, atol=0)
    # pool backward + skip
    dy = _rand(rng, (n, h // 2, w // 2, c))
    skip = _rand(rng, (n, h, w, c))
    dx = torch.zeros((n, h, w, c), dtype=torch.int16, device="cuda")
    # device copies stay referenced until the kernel has run
    dyd, skipd = _dev_bits(dy), _dev_bits(skip)
    G.check(wr.lib().wrk_pool2_bwd(G.vp(dyd), c, c, G.vp(skipd), c, G.vp(dx), c, n, h, w, G.stream()))
    ref = np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2) * 0.25 + skip
    gd, rd = _host_bits(dx), U.bf16_round(ref)
    bad = np.argwhere(gd != rd)
    assert bad.size == 0, (len(bad), [(float(gd[tuple(i)]), float(rd[tuple(i)]), float(ref[tuple(i)])) for i in bad[:4]])
    # upsample forward into a pitched slice, backward sums 2x2 blocks
    up = torch.zeros((n, h, w, 64), dtype=torch.int16, device="cuda")
    small = _rand(rng, (n, h // 2, w // 2, c))
    smalld = _dev_bits(small)
    G.check(wr.lib().wrk_up2_fwd(G.vp(smalld), c, c, G.vp(up), 64, n, h, w, G.stream()))
    assert np.array_equal(_host_bits(up)[..., :c], np.repeat(np.repeat(small, 2, axis=1), 2, axis=2))
    dbig = _rand(rng, (n, h, w, c))
    dsm = torch.zeros((n, h // 2, w // 2, c), dtype=torch.int16, device="cuda")
    dbigd = _dev_bits(dbig)
    G.check(wr.lib().wrk_up2_bwd(G.vp(dbigd), c, c, G.vp(dsm), c, n, h, w, G.stream()))
    ref = dbig.reshape(n, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4))
    assert np.allclose(_host_bits(dsm), U.bf16_round(ref), rtol=2 ** -8, atol=1e-6)
    # relu backward
    g = _rand(rng, (n, h, w, c))
    gd = _dev_bits(g)
    G.check(wr.lib().wrk_relu_bwd(G.vp(gd), c, G.vp(skipd), c, c, n * h * w, G.stream()))
    assert np.array_equal(_host_bits(gd), np.where(skip > 0, g, 0.0).astype(np.float32))


def _images(rng, n, h, w):
    color = np.abs(rng.normal(size=(n, h, w, 3))).astype(np.float32) * 2.0
    albedo = rng.uniform(size=(n, h, w, 3)).astype(np.float32)
    nrm = rng.normal(size=(n, h, w, 3)).astype(np.float32)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    return color, albedo, nrm


def test_unet_forward_matches_torch(gpu):
    rng = np.random.default_rng(11)
    n, h, w = 2, 32, 48
    color, albedo, nrm = _images(rng, n, h, w)
    net = denoiser.UNet(seed=3)
    params = net.get_params()
    # biases are zero after init; give them values so the test covers them
    for _, bs in denoiser.layer_slices():
        params[bs] = rng.normal(size=bs.stop - bs.start).astype(np.float32) * 0.05
    net.set_params(params)
    out = torch.zeros((n, h, w, 3), dtype=torch.float32, device="cuda")
    net.forward(*(torch.from_numpy(a).cuda() for a in (color, albedo, nrm)), out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    x = U.pack_input(color, albedo, nrm)
    pred = (U.forward(params, x)[:, 0:3] + x[:, 0:3]).permute(0, 2, 3, 1).numpy()
    ref = np.maximum(np.expm1(pred), 0.0)
    # 15 layers of bf16 activations: compare in log space
    err = np.abs(np.log1p(got) - np.log1p(ref))
    assert err.mean() < 2e-3 and err.max() < 5e-2, (err.mean(), err.max())


def test_unet_sample_count_channel(gpu):
    """wrk_unet_set_sample_counts feeds log2(spp) / 16 of each pixel's image
    into input channel 9; the output matches the reference with that input
    and differs from the output without it."""
    rng = np.random.default_rng(21)
    n, h, w = 2, 16, 16
    color, albedo, nrm = _images(rng, n, h, w)
    net = denoiser.UNet(seed=5)
    # the last layer starts at zero (identity network); random weights make
    # the output depend on every input channel
    params = (rng.normal(size=net.num_params) * 0.05).astype(np.float32)
    net.set_params(params)
    dev = [torch.from_numpy(a).cuda() for a in (color, albedo, nrm)]
    out0 = torch.zeros((n, h, w, 3), dtype=torch.float32, device="cuda")
    net.forward(*dev, out0)
    spp = np.array([4, 1024], np.uint32)
    net.set_sample_counts(spp)
    out = torch.zeros_like(out0)
    net.forward(*dev, out)
    torch.cuda.synchronize()
    x = U.pack_input(color, albedo, nrm, spp=spp)
    pred = (U.forward(params, x)[:, 0:3] + x[:, 0:3]).permute(0, 2, 3, 1).numpy()
    ref = np.maximum(np.expm1(pred), 0.0)
    got = out.cpu().numpy()
    err = np.abs(np.log1p(got) - np.log1p(ref))
    assert err.mean() < 2e-3 and err.max() < 5e-2, (err.mean(), err.max())
    assert np.abs(got - out0.cpu().numpy()).max() > 1e-3
    # crops take the feature of their source image
    net.set_sample_counts([16])
    origins = torch.tensor([[0, 0, 0]], dtype=torch.int32, device="cuda").view(-1)
    crop = torch.zeros((1, 8, 8, 3), dtype=torch.float32, device="cuda")
    net.forward(*dev, crop, origins=origins, crop=(1, 8, 8))
    torch.cuda.synchronize()
    assert np.isfinite(crop.cpu().numpy()).all()


def test_unet_noise_channel(gpu):
    """wrk_unet_set_noise packs log(1 + noise / (1 + y)) of each source pixel
    into input channel 10 (crops included); weights whose channel-10 slice is
    zero, as after training without it, give the same output with or
    without the layer."""
    rng = np.random.default_rng(23)
    n, h, w = 2, 16, 24
    color, albedo, nrm = _images(rng, n, h, w)
    noise = (np.abs(rng.normal(size=(n, h, w))) * 0.7).astype(np.float32)
    net = denoiser.UNet(seed=7)
    params = (rng.normal(size=net.num_params) * 0.05).astype(np.float32)
    net.set_params(params)
    dev = [torch.from_numpy(a).cuda() for a in (color, albedo, nrm)]
    out0 = torch.zeros((n, h, w, 3), dtype=torch.float32, device="cuda")
    net.forward(*dev, out0)
    net.set_noise(torch.from_numpy(noise).cuda())
    out = torch.zeros_like(out0)
    net.forward(*dev, out)
    torch.cuda.synchronize()
    x = U.pack_input(color, albedo, nrm, noise=noise)
    pred = (U.forward(params, x)[:, 0:3] + x[:, 0:3]).permute(0, 2, 3, 1).numpy()
    ref = np.maximum(np.expm1(pred), 0.0)
    got = out.cpu().numpy()
    err = np.abs(np.log1p(got) - np.log1p(ref))
    assert err.mean() < 2e-3 and err.max() < 5e-2, (err.mean(), err.max())
    assert np.abs(got - out0.cpu().numpy()).max() > 1e-3
    # crops read the noise of their own source pixels
    origins = torch.tensor([[1, 8, 4]], dtype=torch.int32, device="cuda").view(-1)
    crop = torch.zeros((1, 8, 8, 3), dtype=torch.float32, device="cuda")
    net.forward(*dev, crop, origins=origins, crop=(1, 8, 8))
    torch.cuda.synchronize()
    assert np.isfinite(crop.cpu().numpy()).all()
    # a zero channel-10 slice in the first convolution: the layer has no effect
    ws, _ = denoiser.layer_slices()[0]
    w0 = params[ws].reshape(9, 32, 32)  # [tap][cout][cin]
    w0[:, :, 10] = 0.0
    params[ws] = w0.reshape(-1)
    net.set_params(params)
    a, b = torch.zeros_like(out0), torch.zeros_like(out0)
    net.forward(*dev, a)
    net.set_noise(None)
    net.forward(*dev, b)
    torch.cuda.synchronize()
    assert torch.equal(a, b)


def test_noise_layer_matches_sample_moments(gpu):
    """renderWithGuides(noise=True): the standard error of each pixel's mean
    y = (r + g + b) / 3, from the per-path results summed in sample order."""
    d = scenes.cornell(24, 16, 5)
    s = wr.Scene()
    d.build(s)
    s.finalize()
    r = wr.Renderer()
    r.setWavefrontConfig(wr.WavefrontConfig(iterationCap=0, maxPathsInFlight=1 << 16))
    p = wr.RenderParams(width=24, height=16, spp=5, maxBounces=d.maxBounces, russianRouletteDepth=d.rrDepth)
    col, alb, nrm, nse = r.renderWithGuides(s, wr.Camera(**d.camera), p, noise=True)
    assert np.array_equal(col, r.render(s, wr.Camera(**d.camera), p))
    npix = 24 * 16
    L = r.renderPaths(s, wr.Camera(**d.camera), p, 0, npix * 5).reshape(5, npix, 3).astype(np.float32)
    s1 = np.zeros(npix, np.float32)
    s2 = np.zeros(npix, np.float32)
    for k in range(5):  # sample order, FP32 as on the device
        y = (L[k, :, 0] + L[k, :, 1] + L[k, :, 2]) / np.float32(3.0)
        s1 = s1 + y
        s2 = s2 + y * y
    n = np.float32(5.0)
    var = np.maximum(s2 - s1 * s1 / n, np.float32(0.0)) / (n - np.float32(1.0))
    ref = np.sqrt(var / n).reshape(16, 24)
    assert np.allclose(nse, ref, rtol=2e-6, atol=0.0), np.abs(nse - ref).max()
    assert np.count_nonzero(nse) > npix // 4


def test_unet_gradients_and_training(gpu):
    rng = np.random.default_rng(12)
    n, h, w = 2, 32, 32
    color, albedo, nrm = _images(rng, n, h, w)
    target = U.bf16_round(color * 0.5 + 0.1)
    net = denoiser.UNet(seed=4)
    params = net.get_params()
    dev = [torch.from_numpy(a).cuda() for a in (color, albedo, nrm, target)]
    loss0 = net.train_step(*dev, lr=0.0)
    grads = net.get_grads()
    ref_loss, ref_grads = U.loss_and_grads(params, U.pack_input(color, albedo, nrm), target)
    assert abs(loss0 - ref_loss) <= 2e-2 * ref_loss
    for i, (ws, bs) in enumerate(denoiser.layer_slices()):
        g, r = grads[ws], ref_grads[ws]
        cos = float(np.dot(g, r) / (np.linalg.norm(g) * np.linalg.norm(r) + 1e-30))
        assert cos > 0.98, (i, cos)
        rel = np.linalg.norm(g - r) / (np.linalg.norm(r) + 1e-30)
        assert rel < 0.2, (i, rel)
    # a few Adam steps on the same batch lower the loss
    losses = [net.train_step(*dev, lr=2e-3) for _ in range(40)]
    assert losses[-1] < 0.5 * losses[0], losses[::8]
    # crops: origins select sub-images of the source layers
    origins = torch.tensor([[0, 8, 0], [1, 0, 16]], dtype=torch.int32, device="cuda").view(-1)
    out = torch.zeros((2, 16, 24, 3), dtype=torch.float32, device="cuda")
    net.forward(*dev[:3], out, origins=origins, crop=(2, 16, 24))
    full = torch.zeros((n, h, w, 3), dtype=torch.float32, device="cuda")
    net.forward(*dev[:3], full)
    torch.cuda.synchronize()
    assert np.isfinite(out.cpu().numpy()).all() and np.isfinite(full.cpu().numpy()).all()


def test_renderer_denoises_with_shipped_weights(gpu, tmp_path):
    from offlinerenderer_amd import scenes
    d = scenes.large_scene(64, 48, 4, num_spheres=200, num_emitters=16)
    s = wr.Scene()
    d.build(s)
    s.finalize()
    cam = wr.Camera(**d.camera)
    r = wr.Renderer()
    r.setWavefrontConfig(wr.WavefrontConfig(iterationCap=0, maxPathsInFlight=1 << 20))
    p = wr.RenderParams(width=64, height=48, spp=4, maxBounces=8)
    plain = r.render(s, cam, p)
    p.denoiser.enabled = True
    den = r.render(s, cam, p)
    assert np.isfinite(den).all() and den.min() >= 0.0
    assert not np.array_equal(den, plain)
    # the network the renderer runs is the shipped one
    net = denoiser.UNet()
    net.load(str(wr.LIB_PATH.parent.parent / "models" / "denoiser_unet.f32"))
    col, alb, nrm, nse = r.renderWithGuides(s, cam, wr.RenderParams(width=64, height=48, spp=4, maxBounces=8),
                                            noise=True)
    dev = [torch.from_numpy(a).cuda().unsqueeze(0) for a in (col, alb, nrm)]
    out = torch.empty_like(dev[0])
    net.set_sample_counts([4])  # the renderer tells the network its spp
    nd = torch.from_numpy(nse).cuda().unsqueeze(0)
    net.set_noise(nd)  # and each pixel's noise level
    net.forward(*dev, out)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy()[0], den)
    # a file of the wrong size is refused
    bad = tmp_path / "bad.f32"
    np.zeros(10, np.float32).tofile(bad)
    with pytest.raises(RuntimeError):
        r.loadDenoiser(bad)


def test_training_checkpoint_resume_is_bit_exact(gpu, tmp_path):
    """Six training steps in one run equal three steps, a checkpoint, and
    three more steps in a fresh network restored from it (weights, Adam
    moments and step, crop generator), bit for bit."""
    from offlinerenderer_amd import denoise_train as dt

    rng = np.random.default_rng(21)
    n, res = 3, 64
    color, albedo, nrm = _images(rng, n, res, res)
    cuda = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()
    data = {"color": cuda(color), "albedo": cuda(albedo), "normal": cuda(nrm),
            "target": cuda(color * 0.5 + 0.1), "target2": cuda(color * 0.5 + 0.12),
            "spp": np.array([4, 16, 64], dtype=np.uint32)}
    kw = dict(steps=6, batch=2, crop=32, lr=1e-3, seed=5, loss="relmse_split", log=lambda *a: None)
    whole = denoiser.UNet(seed=4)
    dt.train(whole, data, **kw)
    ck = tmp_path / "ck.npz"
    first = denoiser.UNet(seed=4)
    dt.train(first, data, checkpoint=str(ck), until=3, **kw)
    resumed = denoiser.UNet(seed=99)  # different start: everything must come from the file
    dt.train(resumed, data, resume=str(ck), **kw)
    a, b = whole.get_params(), resumed.get_params()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), np.abs(a - b).max()
    assert not np.array_equal(a, first.get_params())  # the last three steps did change the weights


def test_unet_forward_1080p_matches_torch(gpu):
    """The shipped network on a 1920x1080 4-spp render with guides (the
    denoiser workload's shape) against the PyTorch FP32 restatement of the
    graph run on the CPU with the same bf16 rounding between layers. 15
    layers of bf16 activations whose accumulation order differs: compared in
    log space, mean |error| < 2e-3, 99.9% of values within 2e-2."""
    from offlinerenderer_amd import scenes

    d = scenes.large_scene(1920, 1080, 4, num_spheres=200, num_emitters=16)
    s = G.product_scene(d)
    r = wr.Renderer()
    r.setWavefrontConfig(wr.WavefrontConfig(iterationCap=0, maxPathsInFlight=1 << 24))
    col, alb, nrm = r.renderWithGuides(s, wr.Camera(**d.camera), wr.RenderParams(width=1920, height=1080, spp=4))
    assert np.count_nonzero(col) > col.size // 4
    net = denoiser.UNet()
    net.load(str(wr.LIB_PATH.parent.parent / "models" / "denoiser_unet.f32"))
    net.set_sample_counts([4])
    dev = [torch.from_numpy(a).cuda().unsqueeze(0) for a in (col, alb, nrm)]
    out = torch.empty_like(dev[0])
    net.forward(*dev, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()[0]
    x = U.pack_input(col[None], alb[None], nrm[None], spp=[4])
    with torch.no_grad():
        pred = (U.forward(net.get_params(), x)[:, 0:3] + x[:, 0:3]).permute(0, 2, 3, 1).numpy()[0]
    ref = np.maximum(np.expm1(pred), 0.0)
    err = np.abs(np.log1p(got) - np.log1p(ref))
    print(f"1080p forward: mean {err.mean():.2e}, p99.9 {np.quantile(err, 0.999):.2e}, max {err.max():.2e}")
    assert err.mean() < 2e-3 and np.quantile(err, 0.999) < 2e-2, (err.mean(), np.quantile(err, 0.999))


@pytest.mark.parametrize("h,w", [(36, 52), (41, 64), (64, 43)])
def test_renderer_denoises_sizes_not_divisible_by_8(gpu, h, w):
    """denoiser.cpp:77-195 accepts any size: the renderer pads the guide
    layers to multiples of 8 by repeating the last row and column, runs the
    network and returns the top-left crop; equal to the network run on the
    edge-padded layers, bit for bit."""
    from offlinerenderer_amd import scenes

    d = scenes.large_scene(w, h, 4, num_spheres=200, num_emitters=16)
    s = G.product_scene(d)
    cam = wr.Camera(**d.camera)
    r = wr.Renderer()
    r.setWavefrontConfig(wr.WavefrontConfig(iterationCap=0, maxPathsInFlight=1 << 20))
    p = wr.RenderParams(width=w, height=h, spp=4, maxBounces=8)
    col, alb, nrm, nse = r.renderWithGuides(s, cam, p, noise=True)
    p.denoiser.enabled = True
    den = r.render(s, cam, p)
    assert den.shape == (h, w, 3) and np.isfinite(den).all()
    ph, pw = (h + 7) // 8 * 8, (w + 7) // 8 * 8
    pad = lambda a: np.pad(a, ((0, ph - h), (0, pw - w), (0, 0)), mode="edge")  # noqa: E731
    net = denoiser.UNet()
    net.load(str(wr.LIB_PATH.parent.parent / "models" / "denoiser_unet.f32"))
    net.set_sample_counts([4])
    # the noise layer is edge-padded like the others: the network reads every
    # layer at the same (clamped) source position
    nd = torch.from_numpy(np.ascontiguousarray(np.pad(nse, ((0, ph - h), (0, pw - w)), mode="edge"))).cuda()
    net.set_noise(nd.unsqueeze(0))
    dev = [torch.from_numpy(np.ascontiguousarray(pad(a))).cuda().unsqueeze(0) for a in (col, alb, nrm)]
    out = torch.empty_like(dev[0])
    net.forward(*dev, out)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy()[0][:h, :w], den)
    assert not np.array_equal(den, col)


def _kpn_ref(z, color, origins, h, w):
    """PyTorch FP64 restatement of the kernel-prediction output: z (n, K + 3,
    h, w) = filter logits and per-channel offsets, color (imgs, H, W, 3);
    crop k reads image origins[k][0] at (x0, y0) = origins[k][1:],
    neighbours clamped to the image. Returns u = sum_k softmax(logits)_k
    log(1 + max(c_k, 0)) + offset, (n, 3, h, w); the output is exp(u) - 1."""
    R = denoiser.KPN_RADIUS
    side = 2 * R + 1
    K = side * side
    imgs, H, W, _ = color.shape
    out = []
    for k, (img, x0, y0) in enumerate(origins):
        ys = torch.arange(h) + y0
        xs = torch.arange(w) + x0
        wk = torch.softmax(z[k, :K], dim=0)  # (K, h, w)
        acc = torch.zeros((3, h, w), dtype=torch.float64)
        for dy in range(-R, R + 1):
            yy = (ys + dy).clamp(0, H - 1)
            for dx in range(-R, R + 1):
                xx = (xs + dx).clamp(0, W - 1)
                c = torch.log1p(color[img][yy][:, xx].clamp(min=0)).permute(2, 0, 1)  # (3, h, w)
                acc = acc + wk[(dy + R) * side + dx + R] * c
        out.append(acc + z[k, K:K + 3])
    return torch.stack(out)


@pytest.mark.parametrize("mode", [0, 1, 3])
def test_kpn_apply_and_loss_match_torch(gpu, mode):
    rng = np.random.default_rng(31 + mode)
    imgs, H, W = 2, 20, 28
    n, h, w, pitch = 3, 8, 16, 64
    K = denoiser.KPN_TAPS
    assert int(wr.lib().wrk_kpn_taps()) == K
    color = (rng.gamma(0.7, 0.6, size=(imgs, H, W, 3)) - 0.02).astype(np.float32)  # a few negatives, clamped
    target = rng.gamma(0.7, 0.6, size=(imgs, H, W, 3)).astype(np.float32)
    target2 = rng.gamma(0.7, 0.6, size=(imgs, H, W, 3)).astype(np.float32)
    z = _rand(rng, (n, h, w, pitch), 1.5)
    z[..., K:K + 3] = U.bf16_round(z[..., K:K + 3] * 0.1)  # offsets
    # crops at the image edges, so the 7x7 neighbourhoods reach past them
    origins = [(0, 0, 0), (1, W - w, H - h), (1, 5, 7)]
    zt = torch.from_numpy(z[..., :K + 3]).permute(0, 3, 1, 2).double().requires_grad_(True)
    u = _kpn_ref(zt, torch.from_numpy(color).double(), origins, h, w)
    y = torch.expm1(u)
    o = torch.tensor(origins, dtype=torch.int32, device="cuda").view(-1)
    out = torch.zeros((n, h, w, 3), dtype=torch.float32, device="cuda")
    zb = _dev_bits(z)
    cd = torch.from_numpy(color).cuda()
    rc = wr.lib().wrk_kpn_apply(C.c_void_p(zb.data_ptr()), pitch, C.c_void_p(cd.data_ptr()), H, W,
                                C.c_void_p(o.data_ptr()), C.c_void_p(out.data_ptr()), n, h, w, None)
    assert rc == 0
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    want = y.detach().clamp(min=0).permute(0, 2, 3, 1).numpy()
    np.testing.assert_allclose(got, want, rtol=3e-5, atol=3e-6)
    # loss and gradient against autograd of the same expression
    t = torch.from_numpy(target).double()
    t2 = torch.from_numpy(target2).double()
    tc = torch.stack([t[img, y0:y0 + h, x0:x0 + w] for img, x0, y0 in origins]).permute(0, 3, 1, 2)
    t2c = torch.stack([t2[img, y0:y0 + h, x0:x0 + w] for img, x0, y0 in origins]).permute(0, 3, 1, 2)
    if mode == 0:
        per = (u - torch.log1p(tc)) ** 2
    else:
        den = (t2c if mode == 3 else tc) ** 2 + 0.01
        per = (y - tc) ** 2 / den
    per.mean().backward()
    dz = torch.zeros((n, h, w, pitch), dtype=torch.int16, device="cuda")
    part = torch.zeros(((n * h * w + 255) // 256,), dtype=torch.float32, device="cuda")
    td, t2d = torch.from_numpy(target).cuda(), torch.from_numpy(target2).cuda()
    rc = wr.lib().wrk_kpn_loss(C.c_void_p(zb.data_ptr()), pitch, C.c_void_p(cd.data_ptr()),
                               C.c_void_p(td.data_ptr()), H, W, C.c_void_p(o.data_ptr()),
                               C.c_void_p(t2d.data_ptr()), C.c_void_p(dz.data_ptr()), n, h, w, mode,
                               C.c_void_p(part.data_ptr()), None)
    assert rc == 0
    torch.cuda.synchronize()
    loss = float(part.sum()) / (3 * n * h * w)
    want_loss = float(per.detach().mean())
    assert abs(loss - want_loss) <= 2e-5 * abs(want_loss) + 1e-9
    g = U.from_bits(dz.cpu().numpy())
    want_g = zt.grad.permute(0, 2, 3, 1).numpy()
    # gradients are stored in bf16: within one bf16 rounding of the FP64 value
    np.testing.assert_allclose(g[..., :K + 3], want_g, rtol=2 ** -8, atol=1e-3 * np.abs(want_g).max())
    assert not g[..., K + 3:].any()


def test_kpn_network_trains_and_loads_by_size(gpu, tmp_path):
    rng = np.random.default_rng(41)
    n, h, w = 2, 32, 32
    color, albedo, nrm = _images(rng, n, h, w)
    target = U.bf16_round(color * 0.5 + 0.1)
    net = denoiser.UNet(seed=5, head="kernel")
    assert net.num_params == int(wr.lib().wrk_unet_params_for_head(1)) != int(wr.lib().wrk_unet_params_for_head(0))
    dev = [torch.from_numpy(a).cuda() for a in (color, albedo, nrm, target)]
    losses = [net.train_step(*dev, lr=2e-3, loss="relmse") for _ in range(40)]
    assert losses[-1] < 0.7 * losses[0], losses[::8]
    # the forward equals the filter of the network's own logits
    out = torch.zeros((n, h, w, 3), dtype=torch.float32, device="cuda")
    net.forward(*dev[:3], out)
    torch.cuda.synchronize()
    assert np.isfinite(out.cpu().numpy()).all()
    # Renderer.loadDenoiser picks the head from the file's parameter count
    path = tmp_path / "kpn.f32"
    net.save(str(path))
    d = scenes.cornell(24, 16, 4)
    s = wr.Scene()
    d.build(s)
    s.finalize()
    r = wr.Renderer()
    r.loadDenoiser(str(path))
    p = wr.RenderParams(width=24, height=16, spp=4, maxBounces=8, russianRouletteDepth=3.0)
    p.denoiser.enabled = True
    img = r.render(s, wr.Camera(**d.camera), p)
    p.denoiser.enabled = False  # the noisy layers the network takes
    col, alb, nr, nse = r.renderWithGuides(s, wr.Camera(**d.camera), p, noise=True)
    net.set_sample_counts([4])
    nd = torch.from_numpy(nse).cuda().unsqueeze(0)
    net.set_noise(nd)
    ref = torch.zeros((1, 16, 24, 3), dtype=torch.float32, device="cuda")
    net.forward(*[torch.from_numpy(a).cuda().unsqueeze(0) for a in (col, alb, nr)], ref)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(img, ref[0].cpu().numpy())
