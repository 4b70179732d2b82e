"""CPU oracle of the amortization encoder (float64 numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
encoder (gaspacho2_amd/csrc/gp_encoder.hip).  The product never imports it.

The encoder has no counterpart in the reference (the reference starts X from
an external PCA file, src/ggplvm.cu:1088-1095), so this restatement is pinned
by mathematics rather than by reference vectors: with BF16 rounding switched
off its gradient is checked against central finite differences of its own
loss and against torch.autograd in float64 (tests/test_oracle_pins.py), and
Adam against the closed form of its first step (|update| = lr per entry).

With rnd=True it rounds exactly where the GPU path rounds to BF16 (round to
nearest even): the features, the weight copies, the hidden activations fed
to the next product and the backward signals fed to the products; it keeps
the GPU's FP32 values (pre-activations, bias gradients, loss) unrounded.  The
products themselves are float64 here and FP32-accumulated on the GPU.

Model:
  a0 = log1p(y / s); z1 = W1 a0 + b1; h1 = gelu(z1); z2 = W2 h1 + b2;
  h2 = gelu(z2); z3 = W3 h2 + b3 = [mu | logvar];
  loss = mean_i sum_q 1/2 [(x_iq - mu_iq)^2 exp(-s_iq) + s_iq],
  s = logvar clamped to [-10, 10] (zero gradient outside).
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

_SQRT1_2 = 0.70710678118654752
LV_CLAMP = 10.0      # log-variance clamp of the loss (zero gradient outside)
_INV_SQRT_2PI = 0.39894228040143268


def bf16(x) -> np.ndarray:
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z * _SQRT1_2))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z * _SQRT1_2)) + z * _INV_SQRT_2PI * np.exp(-0.5 * z * z)


def features(Y: np.ndarray, ls: np.ndarray, rnd: bool = True) -> np.ndarray:
    """log1p(y_ij / s_i) for a dense N x F count matrix."""
    a = np.log1p(Y / ls[:, None])
    return bf16(a) if rnd else a


def unpack(flat: np.ndarray, F: int, H1: int, H2: int, Q: int) -> dict:
    """Parameter array [W1 (H1 x Fp) | b1 | W2 (H2 x H1) | b2 | W3 (64 x H2) | b3 (64)]."""
    Fp, O = (F + 63) // 64 * 64, 64
    out, o = {}, 0
    flat = np.asarray(flat, dtype=np.float64)
    for name, shape in (("W1", (H1, Fp)), ("b1", (H1,)), ("W2", (H2, H1)), ("b2", (H2,)),
                        ("W3", (O, H2)), ("b3", (O,))):
        n = int(np.prod(shape))
        out[name] = flat[o:o + n].reshape(shape)
        o += n
    out["W1"] = out["W1"][:, :F]
    out["W3"] = out["W3"][:2 * Q]
    out["b3"] = out["b3"][:2 * Q]
    return out


def forward(p: dict, A0: np.ndarray, rnd: bool = True) -> dict:
    r = bf16 if rnd else (lambda v: v)
    z1 = A0 @ r(p["W1"]).T + p["b1"]
    h1 = r(gelu(z1))
    z2 = h1 @ r(p["W2"]).T + p["b2"]
    h2 = r(gelu(z2))
    z3 = h2 @ r(p["W3"]).T + p["b3"]
    return {"z1": z1, "h1": h1, "z2": z2, "h2": h2, "z3": z3}


def loss_grad(p: dict, A0: np.ndarray, X: np.ndarray, rnd: bool = True):
    """Loss (mean over the rows of A0) and gradients {W1, b1, W2, b2, W3, b3}."""
    r = bf16 if rnd else (lambda v: v)
    n, Q = X.shape
    f = forward(p, A0, rnd)
    mu, sr = f["z3"][:, :Q], f["z3"][:, Q:]
    s = np.clip(sr, -LV_CLAMP, LV_CLAMP)
    res = X - mu
    es = np.exp(-s)
    loss = 0.5 * np.sum(res * res * es + s) / n
    inside = (sr > -LV_CLAMP) & (sr < LV_CLAMP)
    dz3 = np.concatenate([-res * es, np.where(inside, 0.5 * (1.0 - res * res * es), 0.0)], axis=1) / n
    g = {}
    d3 = r(dz3)
    g["W3"] = d3.T @ f["h2"]
    g["b3"] = dz3.sum(0)
    dz2 = (d3 @ r(p["W3"])) * gelu_d(f["z2"])
    d2 = r(dz2)
    g["W2"] = d2.T @ f["h1"]
    g["b2"] = dz2.sum(0)
    dz1 = (d2 @ r(p["W2"])) * gelu_d(f["z1"])
    g["W1"] = r(dz1).T @ A0
    g["b1"] = dz1.sum(0)
    return loss, g


def adam(w, g, m, v, t: int, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
    """One Adam step (bias-corrected); returns new (w, m, v)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v
