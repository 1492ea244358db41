"""NumPy reference of the learned warm-start network (test infrastructure only).

The network is this project's own addition (the reference has no learned
part): 14 -> 64 -> 64 -> 2 with ReLU, parameters packed as
W1[64][14], b1[64], W2[64][64], b2[64], W3[2][64], b3[2] (5250 floats) and a
32-float normalisation (input mean[14], input scale[14], output mean[2],
output scale[2]).

`forward` and `gradients` evaluate in float64 with a rounding function applied
at exactly the points where the GPU kernels round to bf16 (matrix-core operands:
normalised inputs, weights, hidden activations, output errors and the
back-propagated pre-activation gradients, which also feed the bias sums).  With
`rnd=bf16` they mirror the GPU arithmetic up to accumulation order; with the
identity they are the exact mathematics, which `numeric_gradient` pins by
central differences.  Only tests/ import this module.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID, N_OUT = 14, 64, 2
OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3 = 0, 896, 960, 5056, 5120, 5248
N_PARAMS = 5250
N_NORM = 32


def bf16(x):
    """Round float32 values to the nearest bfloat16 (ties to even), returned as float64."""
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def exact(x):
    return np.asarray(x, np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(N_HID, N_IN), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(N_HID, N_HID),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3].reshape(N_OUT, N_HID), p[OFF_B3:N_PARAMS])


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3])


def _f32(x):
    """fp32 rounding of values the GPU computes in fp32 before a bf16 rounding."""
    return np.asarray(x, np.float64).astype(np.float32).astype(np.float64)


def forward(params, norm, x, rnd=bf16):
    """Returns a dict of the intermediate values and the predicted impulses `out` (n x 2)."""
    W1, b1, W2, b2, W3, b3 = unpack(params)
    norm = np.asarray(norm, np.float64)
    f32 = _f32 if rnd is bf16 else exact
    xn = rnd(f32((np.asarray(x, np.float64) - norm[0:14]) * norm[14:28]))
    z1 = xn @ rnd(W1).T + b1
    h1 = rnd(f32(np.maximum(z1, 0.0)))
    z2 = h1 @ rnd(W2).T + b2
    h2 = rnd(f32(np.maximum(z2, 0.0)))
    y = h2 @ rnd(W3).T + b3
    out = y * norm[30:32] + norm[28:30]
    return {"xn": xn, "z1": z1, "h1": h1, "z2": z2, "h2": h2, "y": y, "out": out}


def loss_and_error(params, norm, x, t, rnd=bf16):
    F = forward(params, norm, x, rnd)
    norm = np.asarray(norm, np.float64)
    tn = (np.asarray(t, np.float64) - norm[28:30]) / norm[30:32]
    e = F["y"] - tn
    return float(np.sum(e * e) / len(x)), e, F


def gradients(params, norm, x, t, rnd=bf16):
    """Gradient of 1/n sum (y - t_normalised)^2 with respect to the 5250 parameters, and the loss."""
    W1, b1, W2, b2, W3, b3 = unpack(params)
    f32 = _f32 if rnd is bf16 else exact
    loss, e, F = loss_and_error(params, norm, x, t, rnd)
    n = len(x)
    dy = rnd(f32(2.0 * e / n))
    dz2 = rnd(f32((dy @ rnd(W3)) * (F["z2"] > 0)))
    dz1 = rnd(f32((dz2 @ rnd(W2)) * (F["z1"] > 0)))
    gW3 = dy.T @ F["h2"]
    gW2 = dz2.T @ F["h1"]
    gW1 = dz1.T @ F["xn"]
    return pack(gW1, dz1.sum(0), gW2, dz2.sum(0), gW3, dy.sum(0)), loss


def numeric_gradient(params, norm, x, t, idx, h=1e-6):
    """Central differences of the exact loss for the parameter indices idx."""
    p = np.asarray(params, np.float64).copy()
    g = np.zeros(len(idx))
    for k, i in enumerate(idx):
        old = p[i]
        p[i] = old + h
        lp, _, _ = loss_and_error(p, norm, x, t, exact)
        p[i] = old - h
        lm, _, _ = loss_and_error(p, norm, x, t, exact)
        p[i] = old
        g[k] = (lp - lm) / (2 * h)
    return g


def he_init(seed):
    rng = np.random.default_rng(seed)
    W1 = rng.uniform(-1, 1, (N_HID, N_IN)) * np.sqrt(6.0 / N_IN)
    W2 = rng.uniform(-1, 1, (N_HID, N_HID)) * np.sqrt(6.0 / N_HID)
    W3 = rng.uniform(-1, 1, (N_OUT, N_HID)) * np.sqrt(1.0 / N_HID)
    b1 = rng.uniform(-0.1, 0.1, N_HID)
    b2 = rng.uniform(-0.1, 0.1, N_HID)
    b3 = rng.uniform(-0.1, 0.1, N_OUT)
    return pack(W1, b1, W2, b2, W3, b3).astype(np.float32)
