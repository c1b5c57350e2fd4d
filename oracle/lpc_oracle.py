"""CPU restatement of the learned CG preconditioner network.

TEST INFRASTRUCTURE: only tests/ use this module, as the checker of the GPU
forward/backward kernels.  The product runs the network only on the GPU.

The network (no counterpart in the reference; it stands in for the kinetic
smoother precond_hprsdft, cgpc_hprsdft_module.f90:133-361):

    s   = rms(g)                       per band
    x0  = g / s,  x1 = v               (v: the cell's normalised Vloc)
    a1  = relu(conv(W1, [x0, x1]) + b1)    2 -> 16 channels
    a2  = relu(conv(W2, a1) + b2)          16 -> 16
    y   = conv(W3, a2) + b3                16 -> 1
    P g = s * y

conv is a 3x3x3 cross-correlation with periodic wrap on the grid
(j = i1 + n1 i2 + n1 n2 i3, x fastest), W[co, ci, d3, d2, d1] with
d = offset + 1.  Fields are arrays f[c, i3, i2, i1] (C order), so a flat
field f.reshape(C, -1) uses the grid index j.

The GPU stores x0, x1, a1 and a2 and the weights in bfloat16 and
accumulates in FP32; `forward(..., bf16_storage=True)` reproduces those
roundings.
"""
from __future__ import annotations

import numpy as np

CH = 16


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), a, out)


def conv(W, x):
    """out[co] = sum_{ci, d} W[co, ci, d3, d2, d1] x[ci](i + d - 1), periodic."""
    cout, cin = W.shape[:2]
    out = np.zeros((cout,) + x.shape[1:], dtype=np.float64)
    for d3 in range(3):
        for d2 in range(3):
            for d1 in range(3):
                xs = np.roll(x, shift=(1 - d3, 1 - d2, 1 - d1), axis=(1, 2, 3))
                out += np.einsum("oc,c...->o...", W[:, :, d3, d2, d1], xs)
    return out


def conv_t(W, dy):
    """Adjoint of conv in its input: dx[ci] = sum_{co, d} W[co, ci, d] dy[co](i - d + 1)."""
    cout, cin = W.shape[:2]
    dx = np.zeros((cin,) + dy.shape[1:], dtype=np.float64)
    for d3 in range(3):
        for d2 in range(3):
            for d1 in range(3):
                ys = np.roll(dy, shift=(d3 - 1, d2 - 1, d1 - 1), axis=(1, 2, 3))
                dx += np.einsum("oc,o...->c...", W[:, :, d3, d2, d1], ys)
    return dx


def conv_dw(x, dy):
    """dW[co, ci, d] = sum_i dy[co](i) x[ci](i + d - 1)."""
    cout, cin = dy.shape[0], x.shape[0]
    dW = np.zeros((cout, cin, 3, 3, 3))
    for d3 in range(3):
        for d2 in range(3):
            for d1 in range(3):
                xs = np.roll(x, shift=(1 - d3, 1 - d2, 1 - d1), axis=(1, 2, 3))
                dW[:, :, d3, d2, d1] = dy.reshape(cout, -1) @ xs.reshape(cin, -1).T
    return dW


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = {
        "W1": rng.standard_normal((CH, 2, 3, 3, 3)) * np.sqrt(2.0 / (2 * 27)),
        "b1": np.zeros(CH),
        "W2": rng.standard_normal((CH, CH, 3, 3, 3)) * np.sqrt(2.0 / (CH * 27)),
        "b2": np.zeros(CH),
        "W3": rng.standard_normal((1, CH, 3, 3, 3)) * 1e-2,
        "b3": np.zeros(1),
    }
    return p


def forward(p, g3, v3, bf16_storage=False):
    """g3, v3: (n3, n2, n1).  Returns P g and the cache for backward."""
    q = (lambda a: bf16(a).astype(np.float64)) if bf16_storage else (lambda a: np.asarray(a, np.float64))
    W1, W2, W3 = q(p["W1"]), q(p["W2"]), q(p["W3"])
    s = np.sqrt(np.mean(g3.astype(np.float64) ** 2))
    x = np.stack([q(g3 / s), q(v3)])
    z1 = conv(W1, x) + p["b1"][:, None, None, None]
    a1 = q(np.maximum(z1, 0.0))
    z2 = conv(W2, a1) + p["b2"][:, None, None, None]
    a2 = q(np.maximum(z2, 0.0))
    y = conv(W3, a2)[0] + p["b3"][0]
    out = s * y
    cache = dict(s=s, x=x, z1=z1, a1=a1, z2=z2, a2=a2, y=y, W1=W1, W2=W2, W3=W3)
    return out, cache


def loss_and_grads(p, g3, v3, t3, bf16_storage=False, kappa=1.0):
    """L = mean((P g - kappa t)^2) / s^2 = mean((y - kappa t/s)^2) and its
    gradients."""
    out, c = forward(p, g3, v3, bf16_storage)
    s = c["s"]
    r = c["y"] - kappa * t3 / s
    npts = r.size
    L = float(np.mean(r * r))
    dy = (2.0 / npts) * r
    grads = {"b3": np.array([dy.sum()]), "W3": conv_dw(c["a2"], dy[None])}
    da2 = conv_t(c["W3"], dy[None])
    dz2 = da2 * (c["z2"] > 0)
    grads["b2"] = dz2.reshape(CH, -1).sum(1)
    grads["W2"] = conv_dw(c["a1"], dz2)
    da1 = conv_t(c["W2"], dz2)
    dz1 = da1 * (c["z1"] > 0)
    grads["b1"] = dz1.reshape(CH, -1).sum(1)
    grads["W1"] = conv_dw(c["x"], dz1)
    return L, grads
