"""CPU checks of the learned-guess network checker (oracle/mlp_oracle.py):
its backward pass against central finite differences of its own loss in
float64, and the bfloat16 rounding against hand-worked values."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(__file__), '..', 'oracle'))
import mlp_oracle as M  # noqa: E402


def _setup(seed=3, n=7):
    rng = np.random.default_rng(seed)
    p = rng.normal(size=M.NPARAM) * 0.3
    x = rng.normal(size=(n, M.NF)) * 2 + 1
    t = rng.normal(size=(n, 3))
    mean, inv_std = rng.normal(size=M.NF) * 0.1, 1.0 / (0.5 + rng.random(M.NF))
    return p, x, t, mean, inv_std


def test_backward_matches_finite_differences():
    p, x, t, mean, inv_std = _setup()
    scale = 1.7
    loss, g = M.loss_and_grad(x, t, p, mean, inv_std, scale, bf16_ops=False)
    rng = np.random.default_rng(0)
    blocks = [(M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
              (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.NPARAM)]
    h = 1e-6
    for lo, hi in blocks:  # a sample of parameters from every block
        for k in rng.choice(np.arange(lo, hi), size=min(6, hi - lo), replace=False):
            pp, pm = p.copy(), p.copy()
            pp[k] += h
            pm[k] -= h
            fd = (M.loss_and_grad(x, t, pp, mean, inv_std, scale, False)[0] -
                  M.loss_and_grad(x, t, pm, mean, inv_std, scale, False)[0]) / (2 * h)
            assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (lo, k, fd, g[k])
    # the forward pass of loss_and_grad is the same network as forward()
    y = M.forward(x, p, mean, inv_std, bf16_ops=False)
    assert abs(np.sum((y - t / scale) ** 2) / len(x) - loss) < 1e-12


def test_bf16_rounding():
    vals = np.array([1.0, 1.00390625, 1.005859375, 1.0 + 2 ** -8 + 2 ** -9, -3.14159, 0.0])
    r = M.bf16(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # halfway (1 + 2^-8): ties to even -> 1
    assert r[2] == 1.0078125                # above halfway -> up
    assert r[3] == 1.0078125                # 1 + 3*2^-9 -> up
    assert abs(r[4] - (-3.140625)) < 1e-12  # -3.14159 -> -3.140625
    assert r[5] == 0.0
    # idempotent
    assert np.array_equal(M.bf16(r), r)
