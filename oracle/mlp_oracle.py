"""NumPy oracle of the candidate scorer (test infrastructure only: imported by tests/ and nothing
in the product path).

The scorer has no counterpart in the reference. It is the network described for the learned
proposal:
* 64 inputs, three hidden layers of 256 with GELU, and one sigmoid output; GELU(a) = a Phi(a)
  with Phi a polynomial fit of the normal CDF on |a| <= 3.5 (within 8.2e-4 of the erf form);
* BF16 weights and activations as matrix operands, with FP32 accumulation;
* trained with a weighted binary cross-entropy and Adam (Kingma & Ba 2015, Algorithm 1).

``bf16=True`` rounds every matrix operand to bfloat16 (round to nearest even) at the points
where the HIP kernels do, so the two can be compared closely. ``bf16=False`` gives the plain
float64 network that the central-difference gradient check pins.

Parameter vector layout (float32, shared with the kernels):
    Wt1 [256 x 64], b1 [256], Wt2 [256 x 256], b2 [256], Wt3 [256 x 256], b3 [256], w4 [256], b4 [1]
Wt is stored output-major (row n holds the weights into output n).
"""

from __future__ import annotations

import math

import numpy as np

IN, HID = 64, 256
SHAPES = [("Wt1", (HID, IN)), ("b1", (HID,)), ("Wt2", (HID, HID)), ("b2", (HID,)),
          ("Wt3", (HID, HID)), ("b3", (HID,)), ("w4", (HID,)), ("b4", (1,))]
OFFSETS = {}
_o = 0
for _name, _shape in SHAPES:
    OFFSETS[_name] = (_o, _shape)
    _o += int(np.prod(_shape))
NPARAM = _o
WEIGHTS_BF16 = ["Wt1", "Wt2", "Wt3"]  # the matrices the kernels read as bf16 operands


def unpack(theta):
    return {k: theta[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in OFFSETS.items()}


def pack(d):
    out = np.zeros(NPARAM, dtype=np.float64)
    for k, (o, s) in OFFSETS.items():
        out[o:o + int(np.prod(s))] = np.asarray(d[k], dtype=np.float64).ravel()
    return out


def init_params(seed=0):
    """He-uniform hidden weights, small output weights, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SHAPES:
        if name.startswith("Wt"):
            lim = math.sqrt(6.0 / shape[1])
            d[name] = rng.uniform(-lim, lim, shape)
        elif name == "w4":
            lim = math.sqrt(3.0 / shape[0])
            d[name] = rng.uniform(-lim, lim, shape)
        else:
            d[name] = np.zeros(shape)
    return pack(d).astype(np.float32)


def bf16(x):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    u = np.where(nan, 0x7FC0, u).astype(np.uint32) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


# GELU(a) = a / 2 + h(a), h even: h = u P(u) (u = a^2) on |a| <= 3.5 and |a| / 2 beyond, with P
# the float32 coefficients the kernels use (fitted with 3.5^2 P(3.5^2) = 3.5 / 2, so the pieces
# meet); the kernels' branch-free arrangement, evaluated in float64 here
GELU_CUT = 3.5
GELU_COEF = [float(np.float32(c)) for c in (0.39518553018569946, -0.061023153364658356,
                                             0.0072402348741889, -0.0005149735952727497,
                                             1.9077817341894843e-05, -2.691017755296343e-07)]


def gelu(a):
    a = np.asarray(a, dtype=np.float64)
    ac = np.clip(a, -GELU_CUT, GELU_CUT)
    u = ac * ac
    p = np.zeros_like(a)
    for c in reversed(GELU_COEF):
        p = p * u + c
    return 0.5 * (np.abs(a) - np.abs(ac)) + (0.5 * a + p * u)


def gelu_grad(a):
    a = np.asarray(a, dtype=np.float64)
    u = a * a
    q = np.zeros_like(a)
    for k in reversed(range(len(GELU_COEF))):
        q = q * u + (2 * k + 2) * GELU_COEF[k]
    body = 0.5 + a * q
    return np.where(a > GELU_CUT, 1.0, np.where(a < -GELU_CUT, 0.0, body))


def forward(theta, X, bf16_ops=True):
    """Returns a dict with pre-activations a1..a3, activations h0..h3 (the operands) and logits z."""
    P = unpack(np.asarray(theta, dtype=np.float64))
    r = bf16 if bf16_ops else (lambda v: np.asarray(v, dtype=np.float64))
    h0 = r(X)
    a1 = h0 @ r(P["Wt1"]).T + P["b1"]
    h1 = r(gelu(a1))
    a2 = h1 @ r(P["Wt2"]).T + P["b2"]
    h2 = r(gelu(a2))
    a3 = h2 @ r(P["Wt3"]).T + P["b3"]
    h3 = r(gelu(a3))
    z = h3 @ P["w4"] + P["b4"][0]
    return dict(h0=h0, a1=a1, h1=h1, a2=a2, h2=h2, a3=a3, h3=h3, z=z)


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def loss(theta, X, y, pos_weight=1.0, bf16_ops=True):
    """Weighted mean binary cross-entropy with logits: (1/B) sum w_b (softplus(z) - y z)."""
    z = forward(theta, X, bf16_ops)["z"]
    w = np.where(y > 0, pos_weight, 1.0)
    sp = np.maximum(z, 0) + np.log1p(np.exp(-np.abs(z)))
    return float(np.sum(w * (sp - y * z)) / len(z))


def backward(theta, X, y, pos_weight=1.0, bf16_ops=True):
    """Gradient of ``loss`` as a parameter vector (float64), with the kernels' rounding points:
    the back-propagated pre-activation gradients are bf16 operands of the next products."""
    P = unpack(np.asarray(theta, dtype=np.float64))
    r = bf16 if bf16_ops else (lambda v: np.asarray(v, dtype=np.float64))
    F = forward(theta, X, bf16_ops)
    B = len(y)
    w = np.where(y > 0, pos_weight, 1.0)
    dz = w * (sigmoid(F["z"]) - y) / B
    g = {}
    g["w4"] = F["h3"].T @ dz
    g["b4"] = np.array([dz.sum()])
    da3 = r(dz[:, None] * P["w4"][None, :] * gelu_grad(F["a3"]))
    g["Wt3"] = da3.T @ F["h2"]
    g["b3"] = da3.sum(0)
    da2 = r((da3 @ r(P["Wt3"])) * gelu_grad(F["a2"]))
    g["Wt2"] = da2.T @ F["h1"]
    g["b2"] = da2.sum(0)
    da1 = r((da2 @ r(P["Wt2"])) * gelu_grad(F["a1"]))
    g["Wt1"] = da1.T @ F["h0"]
    g["b1"] = da1.sum(0)
    return pack(g)


def adam(theta, grad, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1), step counted from 1. Returns new
    (theta, m, v) in float64."""
    m = b1 * m + (1.0 - b1) * grad
    v = b2 * v + (1.0 - b2) * grad * grad
    mh = m / (1.0 - b1 ** step)
    vh = v / (1.0 - b2 ** step)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v
