"""CPU gradient oracle of the pair-ranking MLP (512 -> 256 -> 256 -> 1, ReLU).

Test infrastructure only: imported by tests/ as the checker of the HIP
forward, backward and Adam kernels.  No reference counterpart (the learned
ranking is an addition on top of the grading path).

Two modes share one set of formulas:
  quant=True   the device kernels' precision contract: BF16 inputs and weight
               copies, BF16 activations H1/H2 and back-propagated dZ2/dZ1,
               sums in float64 (the device sums in float32; tests allow for
               the difference);
  quant=False  the exact float64 network, whose gradient is pinned against
               torch autograd and central finite differences.
Flat parameter layout: W1[256][512] b1[256] W2[256][256] b2[256] w3[256] b3[1].
"""
from __future__ import annotations

import numpy as np

from .ranking import bf16_round, bf16_to_f32

IN, HID = 512, 256
SHAPES = (("W1", (HID, IN)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("w3", (HID,)), ("b3", (1,)))
NUM_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unflatten(flat: np.ndarray) -> dict:
    out, off = {}, 0
    for name, shape in SHAPES:
        size = int(np.prod(shape))
        out[name] = flat[off:off + size].reshape(shape)
        off += size
    return out


def flatten(parts: dict) -> np.ndarray:
    return np.concatenate([np.asarray(parts[name], dtype=np.float64).reshape(-1) for name, _ in SHAPES])


def _q(x: np.ndarray, quant: bool) -> np.ndarray:
    """Round to BF16 (returned as float64) in quant mode."""
    return bf16_to_f32(bf16_round(x.astype(np.float32))).astype(np.float64) if quant else x


def forward(flat: np.ndarray, X: np.ndarray, quant: bool = True):
    """X: float (rows, 512).  Returns (y, cache)."""
    p = unflatten(np.asarray(flat, dtype=np.float64))
    X = _q(np.asarray(X, dtype=np.float64), quant)
    W1 = _q(p["W1"], quant)
    W2 = _q(p["W2"], quant)
    H1 = _q(np.maximum(X @ W1.T + p["b1"], 0.0), quant)
    H2 = _q(np.maximum(H1 @ W2.T + p["b2"], 0.0), quant)
    y = H2 @ p["w3"] + p["b3"][0]
    return y, dict(X=X, W1=W1, W2=W2, H1=H1, H2=H2, p=p)


def loss_and_grad(flat: np.ndarray, X: np.ndarray, target: np.ndarray, quant: bool = True):
    """Mean squared error and its gradient in the flat layout."""
    y, c = forward(flat, X, quant)
    t = np.asarray(target, dtype=np.float64)
    rows = len(t)
    d = y - t
    loss = float(np.mean(d * d))
    dy = 2.0 * d / rows
    p = c["p"]
    dw3 = c["H2"].T @ dy
    db3 = np.array([dy.sum()])
    dZ2 = _q(dy[:, None] * p["w3"][None, :] * (c["H2"] > 0), quant)
    dW2 = dZ2.T @ c["H1"]
    db2 = dZ2.sum(axis=0)
    dZ1 = _q((dZ2 @ c["W2"]) * (c["H1"] > 0), quant)
    dW1 = dZ1.T @ c["X"]
    db1 = dZ1.sum(axis=0)
    return loss, flatten(dict(W1=dW1, b1=db1, W2=dW2, b2=db2, w3=dw3, b3=db3))


def adam_step(w, g, m, v, step: int, lr: float, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015) in float32, arranged as torch.optim.Adam:
    w -= (lr / bc1) m / (sqrt(v) / sqrt(bc2) + eps).  Returns (w, m, v)."""
    f = np.float32
    w, g, m, v = (np.asarray(a, dtype=f) for a in (w, g, m, v))
    m = f(beta1) * m + f(1 - beta1) * g
    v = f(beta2) * v + f(1 - beta2) * g * g
    bc1 = 1.0 - beta1 ** step
    bc2 = 1.0 - beta2 ** step
    w = w - f(lr / bc1) * m / (np.sqrt(v) * f(1.0 / np.sqrt(bc2)) + f(eps))
    return w.astype(f), m, v
