"""Plain PyTorch FP32 (CPU) restatement of the U-Net graph of
offlinerenderer_amd/csrc/unet.cpp, used as the numerics reference of the
denoiser kernels. Activations are rounded to bfloat16 between layers as the
device stores them."""
import numpy as np
import torch
import torch.nn.functional as F

from offlinerenderer_amd import denoiser


def bf16_round(a: np.ndarray) -> np.ndarray:
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(torch.bfloat16).to(torch.float32).numpy()


def bf16_bits(a: np.ndarray) -> np.ndarray:
    """float32 array -> bfloat16 bit patterns (int16, round to nearest even)."""
    t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(torch.bfloat16)
    return t.view(torch.int16).numpy()


def from_bits(b: np.ndarray) -> np.ndarray:
    return torch.from_numpy(np.ascontiguousarray(b)).view(torch.bfloat16).to(torch.float32).numpy()


def torch_weight(wflat: np.ndarray, cin: int, cout: int) -> torch.Tensor:
    """[9][cout][cin] -> torch [cout][cin][3][3]"""
    return torch.from_numpy(wflat.reshape(3, 3, cout, cin)).permute(2, 3, 0, 1).contiguous()


def pack_input(color, albedo, normal, spp=None, noise=None):
    """(n, h, w, 3) each -> (n, 32, h, w) network input (bf16-rounded);
    spp (per image) sets channel 9 to log2(spp) / 16; noise (n, h, w), the
    standard error of each pixel's mean y = (r + g + b) / 3, sets channel 10
    to log(1 + noise / (1 + y))."""
    x = np.zeros(color.shape[:3] + (32,), np.float32)
    x[..., 0:3] = np.log1p(np.maximum(color, 0.0))
    x[..., 3:6] = albedo
    x[..., 6:9] = normal
    if spp is not None:
        x[..., 9] = (np.log2(np.asarray(spp, np.float32)) / np.float32(16.0))[:, None, None]
    if noise is not None:
        y = color.astype(np.float64).sum(-1) / 3.0
        x[..., 10] = np.log1p(np.maximum(noise, 0.0) / (1.0 + np.maximum(y, 0.0)))
    return torch.from_numpy(bf16_round(x)).permute(0, 3, 1, 2)


def _q(t):
    return t.to(torch.bfloat16).to(torch.float32)


def forward(params: np.ndarray, x: torch.Tensor, quantize=True):
    """x: (n, 32, h, w) -> network output (n, 32, h, w); channels 0..2 are the
    correction added to the input's log(1 + colour)."""
    q = _q if quantize else (lambda t: t)
    sl = denoiser.layer_slices()
    W = []
    for (ws, bs), (cin, cout, _) in zip(sl, denoiser.LAYERS):
        w = torch_weight(params[ws], cin, cout)
        if quantize:
            w = _q(w)
        W.append((w, torch.from_numpy(params[bs].copy())))

    def conv(i, t, relu=True):
        y = F.conv2d(t, W[i][0], W[i][1], padding=1)
        return q(F.relu(y) if relu else y)

    a0 = conv(0, x)
    e0 = conv(1, a0)
    p1 = q(F.avg_pool2d(e0, 2))
    a1 = conv(2, p1)
    e1 = conv(3, a1)
    p2 = q(F.avg_pool2d(e1, 2))
    a2 = conv(4, p2)
    e2 = conv(5, a2)
    p3 = q(F.avg_pool2d(e2, 2))
    a3 = conv(6, p3)
    b3 = conv(7, a3)
    t2 = conv(8, torch.cat([F.interpolate(b3, scale_factor=2, mode="nearest"), e2], 1))
    d2 = conv(9, t2)
    t1 = conv(10, torch.cat([F.interpolate(d2, scale_factor=2, mode="nearest"), e1], 1))
    d1 = conv(11, t1)
    t0 = conv(12, torch.cat([F.interpolate(d1, scale_factor=2, mode="nearest"), e0], 1))
    d0 = conv(13, t0)
    return conv(14, d0, relu=False)


def loss_and_grads(params: np.ndarray, x: torch.Tensor, target: np.ndarray):
    """Log-space MSE of the unquantized FP32 network and its gradient."""
    p = torch.from_numpy(params.copy()).requires_grad_(True)
    sl = denoiser.layer_slices()
    W = [(p[ws].reshape(3, 3, cout, cin).permute(2, 3, 0, 1), p[bs])
         for (ws, bs), (cin, cout, _) in zip(sl, denoiser.LAYERS)]

    def conv(i, t, relu=True):
        y = F.conv2d(t, W[i][0], W[i][1], padding=1)
        return F.relu(y) if relu else y

    a0 = conv(0, x); e0 = conv(1, a0); p1 = F.avg_pool2d(e0, 2)
    a1 = conv(2, p1); e1 = conv(3, a1); p2 = F.avg_pool2d(e1, 2)
    a2 = conv(4, p2); e2 = conv(5, a2); p3 = F.avg_pool2d(e2, 2)
    a3 = conv(6, p3); b3 = conv(7, a3)
    d2 = conv(9, conv(8, torch.cat([F.interpolate(b3, scale_factor=2, mode="nearest"), e2], 1)))
    d1 = conv(11, conv(10, torch.cat([F.interpolate(d2, scale_factor=2, mode="nearest"), e1], 1)))
    d0 = conv(13, conv(12, torch.cat([F.interpolate(d1, scale_factor=2, mode="nearest"), e0], 1)))
    pred = conv(14, d0, relu=False)[:, 0:3] + x[:, 0:3]
    tgt = torch.from_numpy(np.log1p(np.maximum(target, 0.0))).permute(0, 3, 1, 2)
    loss = ((pred - tgt) ** 2).mean()
    loss.backward()
    return float(loss.detach()), p.grad.numpy()
