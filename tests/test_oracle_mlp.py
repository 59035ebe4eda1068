"""Pins of the scorer oracle (oracle/mlp_oracle.py) on the CPU.

* The analytic backward pass equals central finite differences of the loss (float64 network, no
  bf16 rounding), for every parameter tensor.
* Adam's first step moves every parameter by -lr * g / (|g| + eps) (m_hat = g, v_hat = g^2 after
  bias correction), and later steps follow Kingma & Ba's recurrences written out by hand.
* bf16 rounding agrees with torch's float32 -> bfloat16 conversion (round to nearest even),
  including ties, infinities and NaN.
* GELU is within 8.2e-4 of the erf form a (1 + erf(a / sqrt 2)) / 2 (math.erf) everywhere,
  GELU(0) = 0, GELU(x) - GELU(-x) = x, and its derivative matches a central difference away
  from the cut at |a| = 3.5.
"""

import math

import numpy as np
import pytest
import torch

from oracle import mlp_oracle as mo


def _small_problem(seed=0, B=7):
    rng = np.random.default_rng(seed)
    theta = mo.init_params(seed).astype(np.float64)
    theta += rng.normal(0, 0.05, theta.shape)  # non-zero biases
    X = rng.uniform(-1, 1, (B, mo.IN))
    y = (rng.uniform(size=B) < 0.4).astype(np.float64)
    return theta, X, y


def test_backward_matches_central_differences():
    theta, X, y = _small_problem()
    g = mo.backward(theta, X, y, pos_weight=2.5, bf16_ops=False)
    rng = np.random.default_rng(1)
    for name, (off, shape) in mo.OFFSETS.items():
        size = int(np.prod(shape))
        picks = rng.choice(size, min(size, 6), replace=False)
        for p in picks:
            j = off + p
            h = 1e-5
            tp, tm = theta.copy(), theta.copy()
            tp[j] += h
            tm[j] -= h
            fd = (mo.loss(tp, X, y, 2.5, False) - mo.loss(tm, X, y, 2.5, False)) / (2 * h)
            assert abs(fd - g[j]) <= 1e-7 + 1e-5 * abs(fd), (name, p, fd, g[j])


def test_planted_backward_error_is_caught():
    """A dropped GELU' factor in layer 2 changes the gradient far beyond the tolerance."""
    theta, X, y = _small_problem(3)
    P = mo.unpack(theta)
    F = mo.forward(theta, X, bf16_ops=False)
    dz = (mo.sigmoid(F["z"]) - y) / len(y)
    da3 = dz[:, None] * P["w4"][None, :] * mo.gelu_grad(F["a3"])
    wrong_da2 = da3 @ P["Wt3"]  # missing * gelu_grad(a2)
    wrong = wrong_da2.T @ F["h1"]
    right = mo.unpack(mo.backward(theta, X, y, 1.0, False))["Wt2"]
    assert np.max(np.abs(wrong - right)) > 1e-4 * np.max(np.abs(right))


def test_adam_closed_form():
    rng = np.random.default_rng(2)
    th = rng.normal(size=50)
    g = rng.normal(size=50)
    m = np.zeros(50)
    v = np.zeros(50)
    lr, eps = 1e-3, 1e-8
    th1, m1, v1 = mo.adam(th, g, m, v, 1, lr=lr, eps=eps)
    np.testing.assert_allclose(th1, th - lr * g / (np.abs(g) + eps), rtol=1e-12, atol=1e-15)
    g2 = rng.normal(size=50)
    th2, m2, v2 = mo.adam(th1, g2, m1, v1, 2, lr=lr, eps=eps)
    mm = 0.9 * (0.1 * g) + 0.1 * g2
    vv = 0.999 * (0.001 * g * g) + 0.001 * g2 * g2
    exp = th1 - lr * (mm / (1 - 0.81)) / (np.sqrt(vv / (1 - 0.999 ** 2)) + eps)
    np.testing.assert_allclose(th2, exp, rtol=1e-12, atol=1e-15)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.normal(size=10000).astype(np.float32),
                        np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -(1 + 2 ** -8), 0.0, -0.0,
                                  np.inf, -np.inf, 3.4e38, 1e-40], dtype=np.float32)])
    mine = mo.bf16(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(mine, ref)
    assert np.isnan(mo.bf16(np.array([np.nan], dtype=np.float32)))[0]
    assert mo.bf16(np.array([1 + 2 ** -8], dtype=np.float32))[0] == 1.0  # tie to even


def test_gelu_polynomial_form():
    x = np.linspace(-8, 8, 16001)
    erf_form = np.array([0.5 * v * (1 + math.erf(v / math.sqrt(2))) for v in x])
    assert np.max(np.abs(mo.gelu(x) - erf_form)) <= 8.2e-4
    assert mo.gelu(np.array([0.0]))[0] == 0.0
    np.testing.assert_allclose(mo.gelu(x) - mo.gelu(-x), x, atol=1e-13)
    # the fitted piece meets the tails: a above 3.5, 0 below -3.5 (to the float32 coefficients)
    far = np.abs(x) >= 3.5
    np.testing.assert_allclose(mo.gelu(x[far]), np.maximum(x[far], 0.0), atol=2e-6)
    xs = x[np.abs(np.abs(x) - 3.5) > 1e-3]
    h = 1e-6
    fd = (mo.gelu(xs + h) - mo.gelu(xs - h)) / (2 * h)
    np.testing.assert_allclose(mo.gelu_grad(xs), fd, atol=1e-7)


def test_parameter_layout_is_the_kernels():
    assert mo.NPARAM == 148481
    assert mo.OFFSETS["Wt2"][0] == 256 * 64 + 256
    assert mo.OFFSETS["b4"][0] == mo.NPARAM - 1


@pytest.mark.parametrize("seed", [0, 5])
def test_product_init_equals_oracle_init(seed):
    from paces_amd.core import scorer
    np.testing.assert_array_equal(scorer.init_params(seed), mo.init_params(seed))
