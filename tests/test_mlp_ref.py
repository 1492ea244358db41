"""CPU checks that pin the NumPy reference of the warm-start network."""
from __future__ import annotations

import numpy as np

from oracle import mlp_ref


def test_bf16_rounding_is_round_to_nearest_even():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)  # bf16 spacing at 1.0
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -8 + 2.0 ** -20, -2.5, 0.0, 3.0e-39],
                    np.float32)
    got = mlp_ref.bf16(vals)
    # exact tie rounds to even (down to 1.0), 3 half-ulps ties up to 1 + 2 ulp, just above a tie rounds up
    assert got[0] == one and got[1] == one
    assert got[2] == np.float32(one + 2 * ulp)
    assert got[3] == np.float32(one + ulp)
    assert got[4] == -2.5 and got[5] == 0.0
    assert abs(got[6] - 3.0e-39) <= 2.0 ** -134  # half the bf16 subnormal spacing
    r = np.random.default_rng(0).normal(0, 10, 10000).astype(np.float32)
    assert np.all(np.abs(mlp_ref.bf16(r) - r) <= np.abs(r) * 2.0 ** -8)


def test_exact_gradient_matches_central_differences():
    rng = np.random.default_rng(1)
    p = mlp_ref.he_init(2).astype(np.float64)
    norm = np.zeros(mlp_ref.N_NORM)
    norm[0:14] = rng.normal(0, 1, 14)
    norm[14:28] = rng.uniform(0.5, 2, 14)
    norm[28:30] = (0.3, -0.1)
    norm[30:32] = (1.5, 0.7)
    x = rng.normal(0, 1, (24, 14))
    t = rng.normal(0, 1, (24, 2))
    g, _ = mlp_ref.gradients(p, norm, x, t, mlp_ref.exact)
    idx = np.concatenate([rng.choice(np.arange(a, b), 12, replace=False) for a, b in [
        (mlp_ref.OFF_W1, mlp_ref.OFF_B1), (mlp_ref.OFF_B1, mlp_ref.OFF_W2), (mlp_ref.OFF_W2, mlp_ref.OFF_B2),
        (mlp_ref.OFF_B2, mlp_ref.OFF_W3), (mlp_ref.OFF_W3, mlp_ref.OFF_B3)]] + [np.arange(mlp_ref.OFF_B3, mlp_ref.N_PARAMS)])
    num = mlp_ref.numeric_gradient(p, norm, x, t, idx)
    assert np.allclose(g[idx], num, rtol=1e-5, atol=1e-8), np.abs(g[idx] - num).max()


def test_forward_layer_shapes_and_relu():
    p = mlp_ref.he_init(3)
    norm = np.zeros(mlp_ref.N_NORM)
    norm[14:28] = 1.0
    norm[30:32] = 1.0
    x = np.zeros((5, 14))
    F = mlp_ref.forward(p, norm, x, mlp_ref.exact)
    W1, b1, W2, b2, W3, b3 = mlp_ref.unpack(p)
    # zero input: the output is the bias path only
    h1 = np.maximum(b1, 0)
    h2 = np.maximum(W2 @ h1 + b2, 0)
    assert np.allclose(F["out"], W3 @ h2 + b3)
    assert F["h1"].shape == (5, 64) and F["out"].shape == (5, 2)
