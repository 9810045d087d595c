"""numpy restatement of the learned-stride MLP -- TEST INFRASTRUCTURE ONLY.

The checker for the BF16 MFMA kernels of nbodysimulator.jl_amd/csrc/nbx_learn.hip; only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.  The network
has no counterpart in the reference (it accelerates the velocity-Verlet path of
src/nbody_simulation_result.jl:480-487); its definition is the one in include/nbx.h:

    x  = bf16((feats - mean) * inv_scale)            (FP32 arithmetic, then BF16)
    h1 = bf16(SiLU(W1 x + b1)), h2 = bf16(SiLU(W2 h1 + b2)), h3 = bf16(SiLU(W3 h2 + b3))
    y  = W4 h3 + b4                                  (weights BF16, sums exact here)
    loss = mean over n x 6 of (y - targets / out_scale)^2

With rounding=True every value is rounded to BF16 where the kernels round it (inputs,
weights, activations, the output gradient 2 e / (6 n) and the pre-activation gradients);
with rounding=False the same network is evaluated in FP64 throughout (the form that
tests/test_learn_oracle.py pins against PyTorch autograd).
"""
from __future__ import annotations

import numpy as np

NIN, NH, NOUT = 35, 128, 6
NPARAM = 38406
NNORM = 76
OFFSETS = {"W1": 0, "b1": 4480, "W2": 4608, "b2": 20992, "W3": 21120, "b3": 37504,
           "W4": 37632, "b4": 38400}
SHAPES = {"W1": (NH, NIN), "b1": (NH,), "W2": (NH, NH), "b2": (NH,), "W3": (NH, NH),
          "b3": (NH,), "W4": (NOUT, NH), "b4": (NOUT,)}


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out = {}
    for k, off in OFFSETS.items():
        size = int(np.prod(SHAPES[k]))
        out[k] = p[off:off + size].reshape(SHAPES[k])
    return out


def pack(d):
    p = np.zeros(NPARAM)
    for k, off in OFFSETS.items():
        v = np.asarray(d[k], dtype=np.float64).ravel()
        p[off:off + v.size] = v
    return p


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _ident(x):
    return np.asarray(x, dtype=np.float64)


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * sigmoid(z)


def dsilu(z):
    s = sigmoid(z)
    return s * (1.0 + z * (1.0 - s))


def normalize_inputs(feats, norm, rounding=True):
    mean = np.asarray(norm[:NIN])
    inv = np.asarray(norm[NIN:2 * NIN])
    if rounding:
        x = (np.asarray(feats, np.float32) - mean.astype(np.float32)) * inv.astype(np.float32)
        return bf16(x)
    return (np.asarray(feats, np.float64) - mean) * inv


def forward(params, norm, feats, rounding=True):
    """Outputs (n x 6, normalized units) and the intermediate values."""
    q = bf16 if rounding else _ident
    P = unpack(np.asarray(params, np.float32).astype(np.float64) if rounding else params)
    x = normalize_inputs(feats, norm, rounding)
    W = [q(P["W1"]), q(P["W2"]), q(P["W3"]), q(P["W4"])]
    z1 = x @ W[0].T + P["b1"]
    h1 = q(silu(z1))
    z2 = h1 @ W[1].T + P["b2"]
    h2 = q(silu(z2))
    z3 = h2 @ W[2].T + P["b3"]
    h3 = q(silu(z3))
    y = h3 @ W[3].T + P["b4"]
    return y, dict(x=x, z1=z1, h1=h1, z2=z2, h2=h2, z3=z3, h3=h3, W=W)


def loss_and_gradient(params, norm, feats, targets, rounding=True):
    """mean((y - targets / out_scale)^2) over n x 6 and its gradient (flat layout)."""
    q = bf16 if rounding else _ident
    y, c = forward(params, norm, feats, rounding)
    n = y.shape[0]
    osc = np.asarray(norm[2 * NIN:2 * NIN + NOUT])
    if rounding:
        t = (np.asarray(targets, np.float32) / osc.astype(np.float32)).astype(np.float64)
        e = y - t
        dY = q(np.float32(2.0 / (6.0 * n)) * e.astype(np.float32))
    else:
        t = np.asarray(targets, np.float64) / osc
        e = y - t
        dY = 2.0 * e / (6.0 * n)
    loss = float(np.mean(e * e))
    W = c["W"]
    g = {}
    g["W4"] = dY.T @ c["h3"]
    g["b4"] = dY.sum(0)
    dZ3 = q((dY @ W[3]) * dsilu(c["z3"]))
    g["W3"] = dZ3.T @ c["h2"]
    g["b3"] = dZ3.sum(0)
    dZ2 = q((dZ3 @ W[2]) * dsilu(c["z2"]))
    g["W2"] = dZ2.T @ c["h1"]
    g["b2"] = dZ2.sum(0)
    dZ1 = q((dZ2 @ W[1]) * dsilu(c["z1"]))
    g["W1"] = dZ1.T @ c["x"]
    g["b1"] = dZ1.sum(0)
    return loss, pack(g)


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, bias-corrected); returns (p, m, v)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def fit_normalization(feats, targets):
    """Feature mean and inverse standard deviation (floored at 1e-6), per-component RMS of
    the targets (floored at 1e-12)."""
    f = np.asarray(feats, np.float64)
    t = np.asarray(targets, np.float64)
    mean = f.mean(0)
    sd = np.sqrt(np.maximum(0.0, (f * f).mean(0) - mean * mean))
    rms = np.sqrt((t * t).mean(0))
    return np.concatenate([mean, 1.0 / np.maximum(sd, 1e-6), np.maximum(rms, 1e-12)])


def rotate_out(y, frames, out_scale):
    """Lab-frame corrections (dx, dv) from normalized outputs and frames (e1, e2)."""
    e1 = frames[:, 0:3]
    e2 = frames[:, 3:6]
    e3 = np.cross(e1, e2)
    d = y * out_scale
    dx = d[:, 0:1] * e1 + d[:, 1:2] * e2 + d[:, 2:3] * e3
    dv = d[:, 3:4] * e1 + d[:, 4:5] * e2 + d[:, 5:6] * e3
    return dx, dv
