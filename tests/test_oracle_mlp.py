"""The oracle's learned-controller network (CPU).

Its hand-written backward pass is pinned against central finite
differences of its own loss (a dropped tanh derivative, a transposed weight
index or a missing 1/n fails it), and its bfloat16 forward against the
float forward within the bfloat16 rounding level.
"""
import numpy as np

import oracle_lib as O


def _setup(n=40, seed=0):
    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, 0.3, O.MLP_PARAMS)
    mean = rng.normal(0.0, 1.0, O.MLP_IN)
    std = rng.uniform(0.5, 2.0, O.MLP_IN)
    raw = mean + std * rng.normal(size=(n, O.MLP_IN))
    t = rng.normal(0.0, 0.5, n)
    return p, mean, std, raw, t


def test_gradient_matches_finite_differences():
    p, mean, std, raw, t = _setup()
    loss, g = O.mlp_loss_grad(p, mean, std, raw, t)
    rng = np.random.default_rng(1)
    idx = np.concatenate([[0, 639, 640, 703, 704, 4799, 4800, 4863, 4864, 4927, 4928],
                          rng.integers(0, O.MLP_PARAMS, 60)])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (O.mlp_loss_grad(pp, mean, std, raw, t)[0] -
              O.mlp_loss_grad(pm, mean, std, raw, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i])) + 1e-9, i
    y = O.mlp_forward(p, mean, std, raw)
    assert abs(loss - 0.5 * np.mean((y - t) ** 2)) < 1e-14


def test_forward_closed_form_single_path():
    """With only one hidden path switched on, the network is
    y = w3 tanh(w2 tanh(w1 x_norm + b1) + b2) + b3."""
    p = np.zeros(O.MLP_PARAMS)
    mean, std = np.zeros(O.MLP_IN), np.ones(O.MLP_IN)
    p[0 * 10 + 3] = 0.7          # W1[0][3]
    p[640 + 0] = -0.2            # b1[0]
    p[704 + 5 * 64 + 0] = 1.3    # W2[5][0]
    p[4800 + 5] = 0.1            # b2[5]
    p[4864 + 5] = -2.0           # w3[5]
    p[4928] = 0.25               # b3
    raw = np.zeros((1, 10))
    raw[0, 3] = 0.9
    want = -2.0 * np.tanh(1.3 * np.tanh(0.7 * 0.9 - 0.2) + 0.1) + 0.25
    assert abs(O.mlp_forward(p, mean, std, raw)[0] - want) < 1e-15


def test_bf16_forward_close_to_float():
    p, mean, std, raw, _ = _setup(200, seed=2)
    y = O.mlp_forward(p, mean, std, raw)
    yb = O.mlp_forward(p, mean, std, raw, bf16=True)
    assert np.max(np.abs(y - yb)) < 0.05 * max(1.0, np.max(np.abs(y)))
    assert np.any(y != yb)
