"""The CPU restatement of the surrogate network, pinned by finite differences
(FP64 path) and by the BF16 rounding rule on known values."""
from __future__ import annotations

import numpy as np

from oracle import mlp_oracle as mo


def test_bf16_rounding_rule():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5, 3.0e38, 1e-40], dtype=np.float64)
    got = mo.bf16(x)
    # 1 + 2^-8 is exact in BF16 (7 mantissa bits: step 2^-7 at 1) -> ties: 1 + 2^-8 rounds to even 1.0
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0 and got[3] == 1.0 + 2 ** -7
    assert got[4] == -2.5
    assert np.isfinite(got[5]) and abs(got[5] / 3.0e38 - 1) < 2 ** -8
    assert got[6] == 2.0 ** -133  # BF16 subnormal step nearest to 1e-40


def test_fp64_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    dims = (7, 9, 4)
    n_in, nh, n_out = dims
    n_params = n_in * nh + nh + nh * nh + nh + nh * n_out + n_out
    p = 0.5 * rng.standard_normal(n_params)
    X = rng.standard_normal((6, n_in))
    T = rng.standard_normal((6, n_out))
    loss, g = mo.loss_and_grad(p, X, T, dims, bf=False)
    eps = 1e-6
    idx = rng.choice(n_params, 60, replace=False)
    for i in idx:
        e = np.zeros(n_params)
        e[i] = eps
        lp, _ = mo.loss_and_grad(p + e, X, T, dims, bf=False)
        lm, _ = mo.loss_and_grad(p - e, X, T, dims, bf=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(fd)), (i, fd, g[i])
