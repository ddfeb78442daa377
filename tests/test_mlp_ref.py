"""The FP64 MLP reference (the checker of the device training kernels) is
pinned by central finite differences of its loss, and its BF16 rounding by
known bit patterns."""
import numpy as np

from oracle import mlp_ref


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.00390625, 1.005859375, 3.0e-39, -2.5], np.float32)
    r = mlp_ref.bf16_round(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to even (1.0); 1.005859375 = 1 + 1.5*2^-8 rounds up
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0078125) and r[4] == -2.5
    assert np.all(mlp_ref.bf16_round(r) == r)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    D, H, rows = 12, 7, 9
    npar = H * D + H + H * H + H + 3 * H + 3
    p = rng.standard_normal(npar) * 0.4
    X = rng.standard_normal((rows, D))
    T = rng.standard_normal((rows, 3))
    loss, g = mlp_ref.loss_and_grad(p, X, T, D, H, bf16_weights=False)
    h = 1e-6
    for k in rng.choice(npar, 40, replace=False):
        e = np.zeros(npar)
        e[k] = h
        lp, _ = mlp_ref.loss_and_grad(p + e, X, T, D, H, bf16_weights=False)
        lm, _ = mlp_ref.loss_and_grad(p - e, X, T, D, H, bf16_weights=False)
        assert abs((lp - lm) / (2 * h) - g[k]) < 1e-7 * max(1.0, abs(g[k]))
    assert abs(loss - np.mean((mlp_ref.forward(p, X, D, H, False) - T) ** 2)) < 1e-14
