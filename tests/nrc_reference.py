"""Float reference of the learned radiance cache's network, for the GPU
tests: the hash-grid encoding in float64 (numpy) and the MLP forward and
backward in float32 (torch autograd) with the kernels' FP16 rounding points
(weights, network input, hidden activations and back-propagated gradients)
reproduced, so that tolerances only absorb FP32 summation order."""

from __future__ import annotations

import math

import numpy as np
import torch

LAYER_SHAPES = ((64, 48), (64, 64), (64, 64), (64, 64), (16, 64))


def level_grid(levels: int, log2_table: int, base: float, finest: float):
    """(resolution, dense) per level, as the library derives them."""
    b = math.exp((math.log(finest) - math.log(base)) / (levels - 1)) if levels > 1 else 1.0
    res, dense = [], []
    for level in range(levels):
        r = max(1, int(math.floor(base * b ** level)))
        res.append(r)
        dense.append((r + 1) ** 3 <= (1 << log2_table))
    return res, dense


def hash_index(x, y, z, entries):
    x = x.astype(np.uint64)
    y = y.astype(np.uint64)
    z = z.astype(np.uint64)
    h = (x ^ ((y * np.uint64(2654435761)) & np.uint64(0xFFFFFFFF)) ^
         ((z * np.uint64(805459861)) & np.uint64(0xFFFFFFFF))) & np.uint64(0xFFFFFFFF)
    return (h & np.uint64(entries - 1)).astype(np.int64)


def encode(pos, nrm, alb, table, bbox_min, bbox_max, base, finest, log2_table):
    """float64[n, 48] network input (hash features from the FP16 table,
    normal, albedo, zeros) before the final FP16 rounding."""
    levels, entries = table.shape[0], table.shape[1]
    res, dense = level_grid(levels, log2_table, base, finest)
    tab = table.astype(np.float16).astype(np.float64)
    u = np.clip((pos.astype(np.float64) - bbox_min) / (bbox_max - bbox_min), 0.0, 1.0)
    out = np.zeros((len(pos), 48))
    for level in range(levels):
        r = res[level]
        x = u * r
        fl = np.minimum(np.floor(x), r - 1)
        f = x - fl
        base_i = fl.astype(np.int64)
        feat = np.zeros((len(pos), 2))
        for k in range(8):
            d = np.array([k & 1, (k >> 1) & 1, (k >> 2) & 1])
            c = base_i + d
            if dense[level]:
                idx = c[:, 0] + (r + 1) * (c[:, 1] + (r + 1) * c[:, 2])
            else:
                idx = hash_index(c[:, 0], c[:, 1], c[:, 2], entries)
            w = np.prod(np.where(d == 1, f, 1.0 - f), axis=1)
            feat += w[:, None] * tab[level, idx]
        out[:, 2 * level:2 * level + 2] = feat
    out[:, 32:35] = nrm
    out[:, 35:38] = alb
    return out


def _round16(x: torch.Tensor) -> torch.Tensor:
    """FP16 rounding in the forward pass, identity in the backward pass."""
    return x + (x.half().float() - x).detach()


class _GradRound16(torch.autograd.Function):
    """Identity forward; rounds the incoming gradient to FP16."""

    @staticmethod
    def forward(ctx, x):
        return x.view_as(x)

    @staticmethod
    def backward(ctx, g):
        return g.half().float()


def mlp(inputs: torch.Tensor, weights: np.ndarray):
    """Forward with the kernels' rounding; returns (outputs[n, 16], input
    leaf tensor, per-layer weight leaves)."""
    x_leaf = inputs.detach().float().half().float().requires_grad_(True)
    ws, off = [], 0
    for o, i in LAYER_SHAPES:
        w = torch.from_numpy(weights[off:off + o * i].reshape(o, i).astype(np.float16).astype(np.float32))
        ws.append(w.to(inputs.device).requires_grad_(True))
        off += o * i
    x = x_leaf
    for k, w in enumerate(ws):
        y = x @ w.t()
        if k < len(ws) - 1:
            x = _round16(torch.relu(_GradRound16.apply(y)))
        else:
            x = _GradRound16.apply(y)
    return x, x_leaf, ws


def relative_l2_sum(pred: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
    """Sum over records and RGB of (p - t)^2 / (sg(p)^2 + 0.01)."""
    p = pred[:, :3]
    return ((p - target) ** 2 / (p.detach() ** 2 + 0.01)).sum()


def table_gradient(pos, input_grad, bbox_min, bbox_max, levels, entries, base, finest, log2_table):
    """float64[levels, entries, 2]: each sample's gradient with respect to its
    level features spread over the 8 cell corners with the trilinear
    weights of ``encode`` and summed exactly (float64)."""
    res, dense = level_grid(levels, log2_table, base, finest)
    u = np.clip((pos.astype(np.float64) - bbox_min) / (bbox_max - bbox_min), 0.0, 1.0)
    g = input_grad.astype(np.float64)
    out = np.zeros((levels, entries, 2))
    for level in range(levels):
        r = res[level]
        x = u * r
        fl = np.minimum(np.floor(x), r - 1)
        f = x - fl
        base_i = fl.astype(np.int64)
        for k in range(8):
            d = np.array([k & 1, (k >> 1) & 1, (k >> 2) & 1])
            c = base_i + d
            if dense[level]:
                idx = c[:, 0] + (r + 1) * (c[:, 1] + (r + 1) * c[:, 2])
            else:
                idx = hash_index(c[:, 0], c[:, 1], c[:, 2], entries)
            w = np.prod(np.where(d == 1, f, 1.0 - f), axis=1)
            np.add.at(out[level], idx, w[:, None] * g[:, 2 * level:2 * level + 2])
    return out
