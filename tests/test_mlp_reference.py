"""Pins the CPU FP64 reference of the step-proposal network: its analytic
gradient equals central finite differences of its loss (without the BF16
rounding, where the loss is piecewise smooth), and the BF16 rounding helper
rounds to nearest-even."""
import numpy as np

from tests import mlp_reference as R


def test_bf16_rounding_helper():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -8, 1.0 + 3 * 2 ** -9, -3.14159, 1e-30], dtype=np.float32)
    r = R.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie, even mantissa kept
    assert r[2] == 1.0 + 2 ** -6            # tie, rounds to even
    assert r[3] == 1.0 + 2 ** -7
    assert abs(r[4] + 3.140625) < 1e-12
    assert r[5] != 0.0


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(scale=0.3, size=R.N_PARAMS)
    x = rng.normal(size=(12, 8))
    t = rng.normal(size=(12, 2))
    args = (np.zeros(8), np.ones(8), np.zeros(2), np.ones(2))
    loss, g, _ = R.loss_and_grad(p, x, t, *args, bf16_round=False)
    idx = rng.choice(R.N_PARAMS, 60, replace=False)
    for i in idx:
        e = np.zeros(R.N_PARAMS)
        e[i] = 1e-6
        lp, _, _ = R.loss_and_grad(p + e, x, t, *args, bf16_round=False)
        lm, _, _ = R.loss_and_grad(p - e, x, t, *args, bf16_round=False)
        fd = (lp - lm) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])
