"""CPU restatement of the MVC surrogate network (test infrastructure only).

numpy float64 forward and backward of the network that cindergs_amd/csrc/
k_mlp.hip trains (no counterpart in the reference): Fourier features
[sin 2pi Bx, cos 2pi Bx] -> 4 x (Linear 256 + ReLU) -> Linear n_out -> affine
normalisation w = z + (1 - sum z)/n_out, loss = mean over points and outputs of
(w - target)^2.  With bf16=True the weights, the input features and the hidden
activations are rounded to bfloat16 (round to nearest even) where the GPU
rounds them, the arithmetic staying float64; the backward pass is exact for the
network so evaluated.  Parameters use the library's flat layout.
"""
from __future__ import annotations

import numpy as np


def bf16(a) -> np.ndarray:
    """Round to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layers(params, rows, cols, offsets):
    out = []
    for r, c, o in zip(rows, cols, offsets):
        blk = np.asarray(params[o:o + r * (c + 32)], np.float64).reshape(r, c + 32)
        out.append((blk[:, :c], blk[:, c]))
    return out


def forward(params, fourier, rows, cols, offsets, x, use_bf16=True):
    """Activations [A0, H1..H4] and Z [n x rows_4]."""
    L = layers(params, rows, cols, offsets)
    x = np.asarray(x, np.float32)
    B = np.asarray(fourier, np.float32)
    s = np.float32(6.283185307179586) * (x @ B.T).astype(np.float32)
    a0 = np.concatenate([np.sin(s.astype(np.float64)), np.cos(s.astype(np.float64))], 1)
    r = bf16 if use_bf16 else (lambda v: np.asarray(v, np.float64))
    acts = [r(a0)]
    h = acts[0]
    for l, (W, b) in enumerate(L):
        z = h @ r(W).T + b
        if l < len(L) - 1:
            h = r(np.maximum(z, 0.0))
            acts.append(h)
        else:
            return acts, z


def loss_and_grad(params, fourier, rows, cols, offsets, x, targets, n_out, use_bf16=True):
    """(loss, flat gradient) of the mean squared error of the normalised weights."""
    L = layers(params, rows, cols, offsets)
    r = bf16 if use_bf16 else (lambda v: np.asarray(v, np.float64))
    acts, z = forward(params, fourier, rows, cols, offsets, x, use_bf16)
    P = x.shape[0]
    zz = z[:, :n_out]
    w = zz + (1.0 - zz.sum(1, keepdims=True)) / n_out
    e = w - np.asarray(targets, np.float64)
    loss = float((e ** 2).sum() / (P * n_out))
    g = 2.0 * e / (P * n_out)
    dz = np.zeros_like(z)
    dz[:, :n_out] = g - g.sum(1, keepdims=True) / n_out
    grad = np.zeros(len(params), np.float64)
    G = dz
    for l in range(len(L) - 1, -1, -1):
        W, _ = L[l]
        h = acts[l]
        o, rr, c = offsets[l], rows[l], cols[l]
        blk = np.zeros((rr, c + 32))
        blk[:, :c] = G.T @ h
        blk[:, c] = G.sum(0)
        grad[o:o + rr * (c + 32)] = blk.reshape(-1)
        if l > 0:
            G = (G @ r(W)) * (acts[l] > 0)
    return loss, grad


def weights(params, fourier, rows, cols, offsets, x, n_out, use_bf16=True):
    _, z = forward(params, fourier, rows, cols, offsets, x, use_bf16)
    zz = z[:, :n_out]
    return zz + (1.0 - zz.sum(1, keepdims=True)) / n_out


def input_derivatives(params, fourier, rows, cols, offsets, x, n_out):
    """dw/dx [n x n_out x 3] by forward-mode propagation of the three input
    tangents with the GPU's bf16 roundings: the Fourier tangents, each masked
    hidden tangent, and the ReLU masks taken from the bf16 primal activations."""
    L = layers(params, rows, cols, offsets)
    acts, _ = forward(params, fourier, rows, cols, offsets, x, use_bf16=True)
    x = np.asarray(x, np.float32)
    B = np.asarray(fourier, np.float32)
    s = (np.float32(6.283185307179586) * (x @ B.T).astype(np.float32)).astype(np.float64)
    out = np.zeros((x.shape[0], n_out, 3))
    for a in range(3):
        ba = 2.0 * np.pi * B[:, a].astype(np.float64)
        t = bf16(np.concatenate([ba * np.cos(s), -ba * np.sin(s)], 1))
        for l, (W, _) in enumerate(L):
            z = t @ bf16(W).T
            if l < len(L) - 1:
                t = bf16(z * (acts[l + 1] > 0))
            else:
                zz = z[:, :n_out]
                out[:, :, a] = zz - zz.sum(1, keepdims=True) / n_out
    return out
