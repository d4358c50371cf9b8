"""The FP64 MLP oracle's backprop against central finite differences (CPU)."""

import numpy as np

from oracle import mlp_oracle as M


def _params(seed=0):
    rng = np.random.default_rng(seed)
    n = 2 * 64 + 64 + 3 * (64 * 64 + 64) + 64 + 1
    return rng.normal(0, 0.2, n)


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(1)
    x = rng.uniform(-50, 50, 96)
    ly = rng.uniform(-4, 2, 96)
    t = rng.normal(0, 0.3, 96)
    p = _params()
    _, g = M.loss_and_grad(p, x, ly, t, quantize=False)
    n = x.size

    def mse(pp):
        out = M.forward(pp, x, ly, quantize=False)[0]
        return float(np.mean((out - t) ** 2))

    idx = rng.choice(p.size, 60, replace=False)
    idx = np.concatenate([idx, [0, 1, 128, 192, 4288, 12672, p.size - 1]])
    for i in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mse(pp) - mse(pm)) / (2 * h)
        assert abs(g[i] - fd) <= 1e-6 * max(1.0, abs(fd)), (i, g[i], fd)


def test_bf16_rounding_is_round_to_nearest_even():
    a = np.array([1.0, 1.00390625, 1.005859375, 1.0078125 + 2 ** -9, -3.0e-5], np.float32)
    b = M.bf16(a)
    assert b[0] == 1.0 and b[1] == 1.0          # halfway 1 + 2^-8: ties to even (1.0)
    assert b[2] == 1.0078125                      # above halfway rounds up
    assert abs(b[4] - (-3.0e-5)) / 3.0e-5 < 2 ** -8
