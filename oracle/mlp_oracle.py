"""CPU oracle of the velocity-surrogate MLP (include/lbm_mlp.h) -- TEST
INFRASTRUCTURE ONLY: the tests compare the HIP kernels with it; no product
path imports it.

Float64 numpy restatement of the network 5 -> 128 -> 128 -> 128 -> 6 with
tanh hidden layers, the loss L = mean over samples and outputs of (y - t)^2
and its gradient by backpropagation, with the parameter layout of
lbm_mlp.h.  With bf16=True every operand of every matrix product (weights,
layer inputs, backpropagated errors) is first rounded to bfloat16 (round to
nearest even), as the BF16 mode feeds its MFMA tiles; products and sums stay
in float64.  There is no reference counterpart (the reference has no learned
part); the gradient is pinned by central finite differences
(tests/test_oracle_cpu.py).
"""
from __future__ import annotations

import numpy as np

IN, HID, OUT = 5, 128, 6
SHAPES = [("W1", (HID, IN)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (HID, HID)),
          ("b3", (HID,)), ("W4", (OUT, HID)), ("b4", (OUT,))]
P = sum(int(np.prod(s)) for _, s in SHAPES)
assert P == 34566


def unpack(p: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def pack(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1) for name, _ in SHAPES])


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _q(a, bf16):
    return bf16_round(a) if bf16 else np.asarray(a, dtype=np.float64)


def forward(p: np.ndarray, x: np.ndarray, bf16: bool = False):
    """Outputs (n, 6) and the hidden activations [h1, h2, h3]."""
    w = unpack(p)
    h, hs = np.asarray(x, dtype=np.float64), []
    for k in (1, 2, 3):
        h = np.tanh(_q(h, bf16) @ _q(w[f"W{k}"], bf16).T + w[f"b{k}"])
        hs.append(h)
    y = _q(h, bf16) @ _q(w["W4"], bf16).T + w["b4"]
    return y, hs


def loss_grad(p: np.ndarray, x: np.ndarray, t: np.ndarray, bf16: bool = False):
    """Loss and dL/dp (float64, layout of lbm_mlp.h)."""
    w = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    y, (h1, h2, h3) = forward(p, x, bf16)
    n = x.shape[0]
    d = y - t
    loss = float(np.mean(d * d))
    dy = 2.0 * d / (n * OUT)
    g = {}
    g["W4"] = _q(dy, bf16).T @ _q(h3, bf16)
    g["b4"] = dy.sum(0)
    dz = (_q(dy, bf16) @ _q(w["W4"], bf16)) * (1.0 - h3 * h3)
    for k, hp in ((3, h2), (2, h1)):
        g[f"W{k}"] = _q(dz, bf16).T @ _q(hp, bf16)
        g[f"b{k}"] = dz.sum(0)
        dz = (_q(dz, bf16) @ _q(w[f"W{k}"], bf16)) * (1.0 - hp * hp)
    g["W1"] = _q(dz, bf16).T @ _q(x, bf16)
    g["b1"] = dz.sum(0)
    return loss, pack(g)


def init_params(seed: int) -> np.ndarray:
    """Glorot-uniform weights, zero biases (an initial point for the tests)."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SHAPES:
        if name[0] == "W":
            a = np.sqrt(6.0 / (shape[0] + shape[1]))
            d[name] = rng.uniform(-a, a, size=shape)
        else:
            d[name] = rng.normal(0.0, 0.1, size=shape)
    return pack(d)
