"""Learned denoiser: the U-Net of csrc/unet.cpp (BF16 MFMA kernels in
csrc/denoise.hip) over the C ABI of include/wrk_unet.h.

Stands in for libWR's OptiX AI denoiser (libWR/src/internal/denoiser.cpp),
which takes three of its layers: noisy colour, first-hit albedo and
normal (Renderer.renderWithGuides). The network also takes each image's
sample count and each pixel's noise level (the standard error of its mean,
renderWithGuides(noise=True)). Image data stays on the device (torch CUDA
tensors as buffers); the parameters are one FP32 vector.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib

# 15 convolutions: (cin, cout, level); weights [9][cout][cin] then bias [cout].
# Output heads: 0 = channels 0..2 of a 32-channel output correct the input's
# log(1 + colour); 1 = the first KPN_TAPS channels of a 64-channel output are
# a 7x7 filter per pixel over the noisy colour (kernel prediction).
LAYERS = [(32, 32, 0), (32, 32, 0), (32, 64, 1), (64, 64, 1), (64, 96, 2), (96, 96, 2), (96, 128, 3), (128, 128, 3),
          (224, 96, 2), (96, 96, 2), (160, 64, 1), (64, 64, 1), (96, 32, 0), (32, 32, 0), (32, 32, 0)]
KPN_RADIUS = 3
KPN_TAPS = (2 * KPN_RADIUS + 1) ** 2
HEADS = {"residual": 0, "kernel": 1}


def layers(head: int = 0):
    return LAYERS[:-1] + [(32, 64 if head else 32, 0)]


def layer_slices(head: int = 0):
    """(weight slice, bias slice) of each convolution in the parameter vector."""
    out, off = [], 0
    for cin, cout, _ in layers(head):
        w = slice(off, off + 9 * cin * cout)
        off += 9 * cin * cout
        b = slice(off, off + cout)
        off += cout
        out.append((w, b))
    return out


def flops_per_pixel(head: int = 0) -> float:
    """Forward FLOPs the network executes per full-resolution pixel (2 per
    multiply-add). The residual head's output convolution computes its first
    16 output channels (wrk_conv3x3_fwd_first), of which 3 are read."""
    ls = layers(head)
    total = 0.0
    for k, (cin, cout, lvl) in enumerate(ls):
        if head == 0 and k == len(ls) - 1:
            cout = 16
        total += 2.0 * 9 * cin * cout / (4 ** lvl)
    return total


def _err():
    return lib().wrk_unet_last_error().decode()


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


class UNet:
    def __init__(self, seed: int = 1, head: int | str = 0):
        self.head = HEADS[head] if isinstance(head, str) else int(head)
        self._h = lib().wrk_unet_create_head(self.head)
        if not self._h:
            raise RuntimeError(f"wrk_unet_create_head: {_err()}")
        self.num_params = int(lib().wrk_unet_num_params(self._h))
        self.init(seed)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().wrk_unet_destroy(h)
            self._h = None

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what}: {_err()}")

    def init(self, seed: int):
        self._check(lib().wrk_unet_init(self._h, int(seed)), "init")

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.num_params, np.float32)
        self._check(lib().wrk_unet_get_params(self._h, out.ctypes.data_as(C.c_void_p)), "get_params")
        return out

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.num_params
        self._check(lib().wrk_unet_set_params(self._h, p.ctypes.data_as(C.c_void_p)), "set_params")

    def get_grads(self) -> np.ndarray:
        out = np.zeros(self.num_params, np.float32)
        self._check(lib().wrk_unet_get_grads(self._h, out.ctypes.data_as(C.c_void_p)), "get_grads")
        return out

    def save(self, path):
        self.get_params().tofile(path)

    def save_checkpoint(self, path, **extra):
        """Parameters, Adam moments and step, and any extra arrays (the
        training loop adds its crop generator's state and next step) in one
        .npz of plain arrays."""
        m = np.zeros(self.num_params, np.float32)
        v = np.zeros(self.num_params, np.float32)
        step = C.c_uint32(0)
        self._check(lib().wrk_unet_get_state(self._h, m.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p),
                                             C.byref(step)), "get_state")
        with open(path, "wb") as f:
            np.savez(f, params=self.get_params(), adam_m=m, adam_v=v, adam_step=np.uint32(step.value), **extra)

    def load_checkpoint(self, path) -> dict:
        """Restores what save_checkpoint wrote; returns the extra arrays."""
        with np.load(path, allow_pickle=False) as z:
            self.set_params(z["params"])
            m = np.ascontiguousarray(z["adam_m"], dtype=np.float32)
            v = np.ascontiguousarray(z["adam_v"], dtype=np.float32)
            assert m.size == self.num_params and v.size == self.num_params
            self._check(lib().wrk_unet_set_state(self._h, m.ctypes.data_as(C.c_void_p),
                                                 v.ctypes.data_as(C.c_void_p), int(z["adam_step"])), "set_state")
            return {k: z[k] for k in z.files if k not in ("params", "adam_m", "adam_v", "adam_step")}

    def load(self, path):
        self.set_params(np.fromfile(path, dtype=np.float32))

    def set_sample_counts(self, spp):
        """Samples per pixel of each source image (input channel 9 =
        log2(spp) / 16); forward and training calls use them from now on."""
        a = np.ascontiguousarray(np.atleast_1d(spp), dtype=np.uint32)
        self._check(lib().wrk_unet_set_sample_counts(self._h, a.ctypes.data_as(C.c_void_p), a.size),
                    "set_sample_counts")

    def set_noise(self, noise):
        """Per-pixel noise layer of the source images (CUDA float32 (imgs, H,
        W): the standard error of each pixel's mean (r + g + b) / 3, as
        Renderer.renderWithGuides(noise=True) returns it), packed into input
        channel 10 by later forward and training calls; None clears it. The
        tensor is kept referenced while set."""
        self._noise = noise
        self._check(lib().wrk_unet_set_noise(self._h, _ptr(noise) if noise is not None else None), "set_noise")

    def forward(self, color, albedo, normal, out, origins=None, crop=None, stream=None):
        """color/albedo/normal: CUDA float32 (imgs, H, W, 3); out: (n, h, w, 3).
        crop = (n, h, w) with origins (CUDA uint32, 3 per crop) for crops."""
        n, h, w = crop if crop else (color.shape[0], color.shape[1], color.shape[2])
        self._check(lib().wrk_unet_forward(self._h, _ptr(color), _ptr(albedo), _ptr(normal), color.shape[1],
                                           color.shape[2], _ptr(origins), n, h, w, _ptr(out), stream), "forward")

    def train_step(self, color, albedo, normal, target, lr, origins=None, crop=None, stream=None,
                   loss: str = "log", target2=None) -> float:
        """One Adam step; loss "log" = MSE of log(1 + L), "relmse" = the
        evaluation metric (y - t)^2 / (t^2 + 0.01), "relmse_pred" = the same
        with the prediction's (constant) square in the denominator, unbiased
        under noisy targets, "relmse_split" = numerator against target and
        denominator from target2 (disjoint samples), unbiased as well."""
        n, h, w = crop if crop else (color.shape[0], color.shape[1], color.shape[2])
        mode = {"log": 0, "relmse": 1, "relmse_pred": 2, "relmse_split": 3}[loss]
        out = C.c_float(0.0)
        self._check(lib().wrk_unet_train_step(self._h, _ptr(color), _ptr(albedo), _ptr(normal), _ptr(target),
                                              _ptr(target2), color.shape[1], color.shape[2], _ptr(origins), n, h, w,
                                              C.c_float(lr), mode, C.byref(out), stream), "train_step")
        return float(out.value)
