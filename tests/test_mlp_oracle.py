"""Pins the float64 MLP restatement (oracle/mlp.py): analytic gradients match
central finite differences, and the bf16 rounding helper rounds to nearest
even."""
import numpy as np

from oracle import mlp


def test_bf16_round_ties_and_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 3.140625, -2.5, 0.0], np.float32)
    r = mlp.bf16_round(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                  # tie -> even (mantissa ...0)
    assert r[2] == 1.0 + 2 ** -7        # 1.5 ulp rounds up
    assert r[3] == 3.140625 and r[4] == -2.5 and r[5] == 0.0


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    p = mlp.init_params(3)
    X = rng.choice([-1.0, 1.0], size=(16, 256))
    t = rng.normal(size=16)
    loss, g = mlp.loss_and_grads(p, X, t, bf16=False)
    eps = 1e-6
    idx = np.concatenate([rng.integers(0, mlp.N_PARAMS, 40), [mlp.N_PARAMS - 1]])
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        lp, _ = mlp.loss_and_grads(pp, X, t, bf16=False)
        lm, _ = mlp.loss_and_grads(pm, X, t, bf16=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_forward_is_linear_in_head():
    p = mlp.init_params(1)
    X = np.ones((4, 256))
    y, c = mlp.forward(p, X, bf16=False)
    P = mlp.unpack(p)
    assert np.allclose(y, c["A2"] @ P["w3"] + P["b3"][0])
