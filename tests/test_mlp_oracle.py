"""The CPU oracle of the surrogate network: its backward pass against central finite
differences of its own loss (plain float64 network, so the difference quotient is accurate),
and its bfloat16 rounding against the bit-level definition."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import mlp_ref  # noqa: E402


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    dims = [7, 11, 9, 5]
    n_params = sum(a * b + b for a, b in zip(dims[:-1], dims[1:]))
    p = rng.normal(0.0, 0.5, n_params)
    x = rng.normal(0.0, 1.0, (6, dims[0]))
    t = rng.normal(0.0, 1.0, (6, dims[-1]))
    g, loss = mlp_ref.gradients(p, dims, x, t, bf16_products=False)
    h = 1e-6
    for i in rng.choice(n_params, 40, replace=False):
        q = p.copy()
        q[i] += h
        lp = mlp_ref.gradients(q, dims, x, t, bf16_products=False)[1]
        q[i] -= 2 * h
        lm = mlp_ref.gradients(q, dims, x, t, bf16_products=False)[1]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_rounding_is_round_to_nearest_even():
    # 1 + 2^-8 lies halfway between bf16 neighbours 1 and 1 + 2^-7: ties go to the even mantissa
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -9, -2.5, 3.0e38], np.float32)
    got = mlp_ref.bf16(x)
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0 + 2.0 ** -6 and got[3] == 1.0
    assert got[4] == -2.5
    assert abs(got[5] - 3.0e38) / 3.0e38 < 2.0 ** -8
