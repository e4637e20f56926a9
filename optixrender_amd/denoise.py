"""Kernel-predicting denoiser (KPCNN) for the path tracer.

Replaces the OptiX AI denoiser (core/src/Renderer.cpp:660-785), whose network
is closed, with a small network of our own:

  features (10 channels per pixel): log(1 + colour) (3), albedo (3),
      normal (3), depth (1) -- the running-mean AOVs of the render; the depth
      AOV is already d / (1 + d) of the first-hit distance d (1 for a miss).
  5 layers of 3x3 convolutions: 10 -> 32 -> 32 -> 32 -> 32 -> 25, ReLU after
      the first four.
  output: softmax over the 25 logits gives a 5x5 kernel per pixel, applied
      to the linear colour of the 5x5 neighbourhood (edge pixels clamp).

Inference runs in the HIP library (BF16 operands on MFMA, FP32 accumulate;
oxr_denoiser_load, then oxr_render with denoise = 1 or oxr_denoise_host).  Training uses PyTorch on the GPU with
the tracer's own renders as data: low-spp inputs against high-spp targets of
seeded sphere-field scenes.  Weights are stored in denoiser_kpcnn.npz next to
this file (float32 masters; the library converts them to BF16).
"""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np

WEIGHTS = Path(__file__).resolve().parent / "denoiser_kpcnn.npz"
CHANNELS = [10, 32, 32, 32, 32, 25]
KERNEL = 5


def features(color: np.ndarray, albedo: np.ndarray, normal: np.ndarray, depth: np.ndarray) -> np.ndarray:
    """(h, w, 10) float32 input features from the render AOVs: colour,
    albedo, normal (h, w, 4) and the depth AOV (h, w)."""
    c = np.log1p(np.maximum(color[..., :3], 0.0))
    d = np.asarray(depth, np.float32).reshape(depth.shape[0], depth.shape[1], 1)
    return np.concatenate([c, albedo[..., :3], normal[..., :3], d], -1).astype(np.float32)


def parameter_count() -> int:
    return sum(9 * a * b + b for a, b in zip(CHANNELS[:-1], CHANNELS[1:]))


def flops_per_pixel() -> int:
    """Multiply-adds x 2 of the convolutions plus the 5x5x3 kernel application."""
    conv = 2 * 9 * sum(a * b for a, b in zip(CHANNELS[:-1], CHANNELS[1:]))
    return conv + 2 * KERNEL * KERNEL * 3


RAW_WEIGHTS = WEIGHTS.with_suffix(".f32")


def save_weights(layers, path: Path = WEIGHTS, meta: dict | None = None):
    """layers: list of (w[co, ci, 3, 3], b[co]) float arrays.  Writes the
    .npz and, beside it, the flat float32 file (.f32) the C++ facade loads."""
    path = Path(path)
    arrs = {}
    for i, (w, b) in enumerate(layers):
        arrs[f"w{i}"] = np.asarray(w, np.float32)
        arrs[f"b{i}"] = np.asarray(b, np.float32)
    for k, v in (meta or {}).items():
        arrs[f"meta_{k}"] = np.asarray(v)
    np.savez(path, **arrs)
    flat_weights(layers).astype("<f4").tofile(path.with_suffix(".f32"))


def load_weights(path: Path = WEIGHTS):
    with np.load(path, allow_pickle=False) as z:
        n = len(CHANNELS) - 1
        return [(z[f"w{i}"], z[f"b{i}"]) for i in range(n)]


def flat_weights(layers) -> np.ndarray:
    """Weights in the order the C ABI takes them: per layer w[co][ci][ky][kx] then b[co]."""
    parts = []
    for (w, b), ci, co in zip(layers, CHANNELS[:-1], CHANNELS[1:]):
        assert w.shape == (co, ci, 3, 3) and b.shape == (co,)
        parts += [np.asarray(w, np.float32).reshape(-1), np.asarray(b, np.float32).reshape(-1)]
    return np.concatenate(parts)


# ---------------------------------------------------------------- torch side
def torch_model():
    import torch
    from torch import nn

    class KPCNN(nn.Module):
        def __init__(self):
            super().__init__()
            self.convs = nn.ModuleList(
                [nn.Conv2d(a, b, 3, padding=1) for a, b in zip(CHANNELS[:-1], CHANNELS[1:])])

        def forward(self, feats, color):
            """feats (n, 10, h, w), color (n, 3, h, w) linear -> denoised (n, 3, h, w)."""
            x = feats
            for i, cv in enumerate(self.convs):
                x = cv(x)
                if i < len(self.convs) - 1:
                    x = torch.relu(x)
            k = torch.softmax(x, dim=1)                                   # (n, 25, h, w)
            pad = KERNEL // 2
            cp = torch.nn.functional.pad(color, (pad, pad, pad, pad), mode="replicate")
            out = torch.zeros_like(color)
            for dy in range(KERNEL):
                for dx in range(KERNEL):
                    t = dy * KERNEL + dx
                    out = out + k[:, t:t + 1] * cp[:, :, dy:dy + color.shape[2], dx:dx + color.shape[3]]
            return out

    return KPCNN()


def torch_layers(model):
    return [(cv.weight.detach().float().cpu().numpy(), cv.bias.detach().float().cpu().numpy()) for cv in model.convs]


def reference_denoise(layers, color, albedo, normal, depth, bf16: bool = True):
    """FP32 torch reference of the device denoiser on one image ((h, w, 4) AOVs).

    bf16=True rounds weights and every layer's input activations to BF16 as
    the device kernel does (accumulation stays FP32)."""
    import torch

    def q(t):
        return t.to(torch.bfloat16).to(torch.float32) if bf16 else t

    f = torch.from_numpy(features(color, albedo, normal, depth)).permute(2, 0, 1)[None]
    c = torch.from_numpy(np.ascontiguousarray(color[..., :3])).permute(2, 0, 1)[None].float()
    x = f
    for i, (w, b) in enumerate(layers):
        x = torch.nn.functional.conv2d(q(x), q(torch.from_numpy(w)), torch.from_numpy(b), padding=1)
        if i < len(layers) - 1:
            x = torch.relu(x)
    k = torch.softmax(x, dim=1)
    pad = KERNEL // 2
    cp = torch.nn.functional.pad(c, (pad, pad, pad, pad), mode="replicate")
    out = torch.zeros_like(c)
    for dy in range(KERNEL):
        for dx in range(KERNEL):
            t = dy * KERNEL + dx
            out = out + k[:, t:t + 1] * cp[:, :, dy:dy + c.shape[2], dx:dx + c.shape[3]]
    return out[0].permute(1, 2, 0).numpy()


def rmse(a: np.ndarray, b: np.ndarray) -> float:
    return float(np.sqrt(np.mean((np.asarray(a, np.float64)[..., :3] - np.asarray(b, np.float64)[..., :3]) ** 2)))


def log_rmse(a, b) -> float:
    return rmse(np.log1p(np.maximum(a[..., :3], 0)), np.log1p(np.maximum(b[..., :3], 0)))
