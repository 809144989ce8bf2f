"""CPU FP64 reference of the surrogate network — TEST INFRASTRUCTURE ONLY.

Restates, in numpy, the network liblyap.so trains on the GPU (lyap_mlp_*):
(p1, p2) -> [sin, cos](2 pi B u) -> 3 x GELU(256) -> [mu1, mu2, mu3, s, e],
loss = 0.5*((y1-mu1)^2 exp(-s) + s) + 0.5*((y2-mu2)^2 + (y3-mu3)^2)   (valid points)
     + softplus(e) - bad*e                                              (all points),
averaged over the batch.  Parameters are the flat FP32 vector the library
exposes (per layer: W[out][in] then b[out]; head padded to 16 rows).

With bf16=True the values the GPU stores in BF16 (GEMM inputs: features,
activations, head/hidden gradients, weights) are rounded the same way, so the
comparison isolates accumulation order; with bf16=False it is the exact FP64
network (finite-difference checks).
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

OUT, OUT_PAD = 5, 16


def bf16(a):
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params, dims):
    out, off = [], 0
    for l in range(len(dims) - 1):
        o, i = dims[l + 1], dims[l]
        W = params[off:off + o * i].reshape(o, i).astype(np.float64)
        off += o * i
        b = params[off:off + o].astype(np.float64)
        off += o
        out.append((W, b))
    return out


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2 * math.pi)


def forward_backward(params, fourier, p, lam, status, ranges, y_mean, y_std, hidden=256,
                     bf16_store=True):
    """Returns (loss, flat gradient, head outputs Z[n,16])."""
    rnd = bf16 if bf16_store else (lambda a: np.asarray(a, dtype=np.float64))
    F = fourier.reshape(-1, 2).astype(np.float64)
    nf = F.shape[0]
    dims = [2 * nf, hidden, hidden, hidden, OUT_PAD]
    layers = unpack(params, dims)
    lo1, hi1, lo2, hi2 = ranges
    u = np.stack([(p[:, 0] - lo1) / (hi1 - lo1), (p[:, 1] - lo2) / (hi2 - lo2)], axis=1)
    f = 2 * math.pi * (u @ F.T)
    X = rnd(np.concatenate([np.sin(f), np.cos(f)], axis=1))
    acts, zs = [X], []
    for l, (W, b) in enumerate(layers):
        Z = acts[-1] @ rnd(W).T + b
        zs.append(Z)
        if l + 1 < len(layers):
            acts.append(rnd(gelu(Z)))
    Z = zs[-1]
    n = p.shape[0]
    bad = (status != 0).astype(np.float64)
    y = (lam - np.asarray(y_mean)) / np.asarray(y_std)
    e = Z[:, 4]
    loss = np.sum(np.logaddexp(0.0, e) - bad * e)
    g = np.zeros_like(Z)
    g[:, 4] = 1.0 / (1.0 + np.exp(-e)) - bad
    ok = bad == 0
    s = Z[ok, 3]
    d1, d2, d3 = Z[ok, 0] - y[ok, 0], Z[ok, 1] - y[ok, 1], Z[ok, 2] - y[ok, 2]
    loss += np.sum(0.5 * (d1 * d1 * np.exp(-s) + s) + 0.5 * (d2 * d2 + d3 * d3))
    g[ok, 0] = d1 * np.exp(-s)
    g[ok, 1] = d2
    g[ok, 2] = d3
    g[ok, 3] = 0.5 * (1.0 - d1 * d1 * np.exp(-s))
    loss /= n
    g /= n
    grads = [None] * len(layers)
    dz = g
    for l in range(len(layers) - 1, -1, -1):
        W, b = layers[l]
        dzs = rnd(dz)
        db = dz.sum(axis=0) if l == len(layers) - 1 else dzs.sum(axis=0)
        grads[l] = (dzs.T @ acts[l], db)
        if l == 0:
            break
        dh = dzs @ rnd(W)
        dz = dh * gelu_grad(zs[l - 1])
    flat = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
    return loss, flat, Z
