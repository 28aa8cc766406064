"""The CPU oracle of the surrogate MLP (oracle/mlp_oracle.py), pinned by
mathematics: central finite differences of the loss for the gradient, and
known answers for the bfloat16 rounding."""
import numpy as np
import pytest

from oracle import mlp_oracle as O


def test_parameter_count():
    assert O.P == 34566


def test_bf16_round_known_answers():
    v = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 3.0e-3, 0.0])
    r = O.bf16_round(v)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # halfway: ties to even
    assert r[2] == 1.0 + 2.0 ** -6          # halfway: ties to even (upward)
    assert r[3] == -2.5 and r[5] == 0.0
    assert abs(r[4] - 3.0e-3) <= 3.0e-3 * 2.0 ** -8


@pytest.mark.parametrize("n", [1, 7, 20])
def test_gradient_matches_finite_differences(n):
    rng = np.random.default_rng(n)
    p = O.init_params(3)
    x = rng.normal(size=(n, O.IN))
    t = rng.normal(size=(n, O.OUT))
    loss, g = O.loss_grad(p, x, t)
    idx = rng.choice(O.P, size=60, replace=False)
    # every block of parameters is probed
    blocks = np.cumsum([int(np.prod(s)) for _, s in O.SHAPES])
    idx = np.concatenate([idx, blocks - 1, blocks[:-1]])
    h = 1e-6
    for k in idx:
        e = np.zeros(O.P)
        e[k] = h
        lp, _ = O.loss_grad(p + e, x, t)
        lm, _ = O.loss_grad(p - e, x, t)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g).max()), (k, fd, g[k])


def test_forward_shapes_and_linear_output():
    p = O.init_params(0)
    w = O.unpack(p)
    x = np.zeros((3, O.IN))
    y, hs = O.forward(p, x)
    # zero inputs: h1 = tanh(b1), ..., y = W4 h3 + b4
    h1 = np.tanh(w["b1"])
    h2 = np.tanh(w["W2"] @ h1 + w["b2"])
    h3 = np.tanh(w["W3"] @ h2 + w["b3"])
    assert np.allclose(y, w["W4"] @ h3 + w["b4"])
    assert [h.shape for h in hs] == [(3, O.HID)] * 3
