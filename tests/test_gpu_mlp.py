"""The surrogate MLP's HIP kernels against the CPU oracle (oracle/mlp_oracle.py,
itself pinned by finite differences): forward outputs and loss gradients in
the FP32 mode (FP32 products) and the BF16 mode (BF16 operands in MFMA tiles,
FP32 accumulation; the oracle rounds the same operands), then training."""
import numpy as np
import pytest

from oracle import mlp_oracle as O

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, 5)).astype(np.float32), rng.normal(size=(n, 6)).astype(np.float32)


@pytest.mark.parametrize("n", [1, 16, 37, 200])
def test_forward_fp32_matches_oracle(lbm, n):
    x, _ = _data(n, n)
    with lbm.MLP(nets=2, mode="fp32", seed=5) as m:
        y = m.forward(x)
        for k in range(2):
            p = m.params(k).astype(np.float64)
            yo, _ = O.forward(p, x.astype(np.float64))
            assert np.abs(y[k] - yo).max() <= 2e-5 * max(1.0, np.abs(yo).max())


@pytest.mark.parametrize("n", [16, 37, 200])
def test_forward_bf16_matches_oracle(lbm, n):
    x, _ = _data(n, 100 + n)
    with lbm.MLP(nets=2, mode="bf16", seed=6) as m:
        y = m.forward(x)
        for k in range(2):
            p = m.params(k).astype(np.float64)
            yb, _ = O.forward(p, x.astype(np.float64), bf16=True)
            y64, _ = O.forward(p, x.astype(np.float64))
            scale = max(1.0, np.abs(yb).max())
            # same rounding of the operands: agreement to the FP32 accumulation and the
            # occasional operand that rounds to the neighbouring bfloat16
            assert np.abs(y[k] - yb).max() <= 5e-3 * scale
            # and BF16 operands stay close to the exact network
            assert np.abs(y[k] - y64).max() <= 5e-2 * scale


def test_bf16_copies_ready_right_after_create(lbm):
    """The largest handle, evaluated at once: the clears of the BF16 weight
    copies are ordered before the cast that fills them (both on the handle's
    stream), so every net sees its own weights."""
    x, _ = _data(40, 5)
    with lbm.MLP(nets=64, mode="bf16", seed=3) as m:
        y = m.forward(x)
        for k in (0, 31, 63):
            p = m.params(k).astype(np.float64)
            yb, _ = O.forward(p, x.astype(np.float64), bf16=True)
            assert np.abs(y[k] - yb).max() <= 5e-3 * max(1.0, np.abs(yb).max())


@pytest.mark.parametrize("n", [16, 45])
def test_gradient_fp32_matches_oracle(lbm, n):
    x, t = _data(n, 7 + n)
    with lbm.MLP(nets=3, mode="fp32", seed=8) as m:
        for k in (0, 2):
            loss, g = m.loss_grad(k, x, t)
            p = m.params(k).astype(np.float64)
            lo, go = O.loss_grad(p, x.astype(np.float64), t.astype(np.float64))
            assert loss == pytest.approx(lo, rel=2e-5)
            assert np.abs(g - go).max() <= 2e-5 * np.abs(go).max()


@pytest.mark.parametrize("n", [20, 64, 101])
def test_gradient_bf16_matches_oracle(lbm, n):
    """n below, at and past whole 32-sample blocks (padded samples must add
    nothing), for a net that is not the first of the handle."""
    x, t = _data(n, 11 + n)
    with lbm.MLP(nets=3, mode="bf16", seed=9) as m:
        for k in (0, 2):
            loss, g = m.loss_grad(k, x, t)
            p = m.params(k).astype(np.float64)
            lo, gb = O.loss_grad(p, x.astype(np.float64), t.astype(np.float64), bf16=True)
            assert loss == pytest.approx(lo, rel=1e-2)
            assert np.abs(g - gb).max() <= 2e-2 * np.abs(gb).max()
            # per-block agreement of the gradient direction
            blocks = np.cumsum([0] + [int(np.prod(s)) for _, s in O.SHAPES])
            for a, b in zip(blocks[:-1], blocks[1:]):
                u, v = g[a:b].astype(np.float64), gb[a:b]
                assert u @ v / (np.linalg.norm(u) * np.linalg.norm(v) + 1e-30) > 0.999


@pytest.mark.parametrize("mode", ["fp32", "bf16"])
def test_one_adam_step_of_every_net(lbm, mode):
    """One epoch of one full batch: every net (trained together in one launch)
    takes one Adam step, which moves each parameter by lr * g / (|g| + eps)
    (bias-corrected first step) -- lr * sign(g) wherever |g| >> eps -- with g
    the oracle's gradient of that net at its initial parameters."""
    x, t = _data(96, 21)
    lr = 1e-3
    with lbm.MLP(nets=4, mode=mode, seed=12) as m:
        p0 = [m.params(k).astype(np.float64) for k in range(4)]
        m.train(x, t, epochs=1, batch=96, lr=lr)
        for k in range(4):
            _, g = O.loss_grad(p0[k], x.astype(np.float64), t.astype(np.float64), bf16=(mode == "bf16"))
            step = m.params(k).astype(np.float64) - p0[k]
            # entries well clear of the BF16 gradient's own error (2e-2 of the largest)
            big = np.abs(g) > (1e-3 if mode == "fp32" else 5e-2) * np.abs(g).max()
            assert big.sum() > 200
            # float32 parameters of size ~0.1 resolve the step to ~1e-8
            assert np.abs(step[big] + lr * np.sign(g[big])).max() <= 2e-2 * lr
            assert np.abs(step).max() <= lr * (1 + 1e-5)


def test_set_params_round_trip(lbm):
    p = O.init_params(4).astype(np.float32)
    with lbm.MLP(nets=2, mode="fp32", seed=1) as m:
        m.set_params(1, p)
        assert np.array_equal(m.params(1), p)
        x, _ = _data(9, 3)
        y = m.forward(x)
        yo, _ = O.forward(p.astype(np.float64), x.astype(np.float64))
        assert np.abs(y[1] - yo).max() <= 2e-5 * max(1.0, np.abs(yo).max())


@pytest.mark.parametrize("mode", ["fp32", "bf16"])
def test_training_fits_a_smooth_map(lbm, mode):
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, size=(2048, 5)).astype(np.float32)
    t = np.stack([np.sin(2 * x[:, 0]) * x[:, 1], np.cos(x[:, 2] + x[:, 3]), x[:, 4] ** 2 - 0.3,
                  np.tanh(x[:, 0] - x[:, 2]), 0.5 * x[:, 1] * x[:, 3], np.sin(x[:, 4] + x[:, 0])], 1).astype(np.float32)
    with lbm.MLP(nets=4, mode=mode, seed=2) as m:
        l0 = np.array([m.loss_grad(k, x, t)[0] for k in range(4)])
        losses = m.train(x, t, epochs=60, batch=128, lr=2e-3)
        assert np.all(losses < 0.05 * l0), (l0, losses)
        assert np.all(losses < 2e-3), losses
        mean, spread = m.predict(x[:256])
        assert np.sqrt(np.mean((mean - t[:256]) ** 2)) < 0.05
        assert np.median(spread) < 0.1
        # the nets were initialised and shuffled differently
        assert not np.array_equal(m.params(0), m.params(1))
