"""Scorer kernels (paces_mlp_*) against the NumPy oracle, through the C ABI.

The oracle rounds the matrix operands to bf16 where the kernels do and sums in float64. The
kernels sum in FP32, so an activation near a bf16 rounding boundary may round one ulp apart.
Tolerances:
* logits: |z_gpu - z_oracle| <= 1e-2 + 1e-2 |z|;
* gradients: per tensor, ||g_gpu - g_oracle|| <= 2e-2 ||g_oracle||;
* Adam: 1e-5 relative (FP32 against float64);
* BF16 copies: exact.
"""

import numpy as np
import pytest

from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    from paces_amd.config import devices
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not devices.configure_called:
        devices.configure()
    from paces_amd.core import scorer
    return dict(torch=torch, dev=devices.vector_dev, scorer=scorer)


def _bits(env, X):
    torch = env["torch"]
    return torch.from_numpy(np.ascontiguousarray(X, dtype=np.float32)).to(torch.bfloat16) \
        .view(torch.int16).to(env["dev"])


def _theta(seed):
    rng = np.random.default_rng(seed + 100)
    th = mo.init_params(seed).astype(np.float64)
    th += rng.normal(0, 0.02, th.shape)
    return th.astype(np.float32)


@pytest.mark.parametrize("tile", [64, 128, 256])
@pytest.mark.parametrize("n", [1, 64, 1000, 4099])
def test_infer_matches_oracle(env, n, tile):
    sc = env["scorer"]
    th = _theta(n)
    X = np.random.default_rng(n).uniform(-1, 1, (n, mo.IN)).astype(np.float32)
    s = sc.CandidateScorer(env["dev"], params=th)
    z = s.infer(_bits(env, X), want_logit=True, tile_rows=tile).cpu().numpy()
    p = s.infer(_bits(env, X), tile_rows=tile).cpu().numpy()
    zo = mo.forward(th.astype(np.float64), X)["z"]
    assert np.all(np.abs(z - zo) <= 1e-2 + 1e-2 * np.abs(zo)), np.max(np.abs(z - zo))
    np.testing.assert_allclose(p, 1 / (1 + np.exp(-z.astype(np.float64))), rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("B,gather", [(1000, False), (777, True), (8192, True)])
def test_gradient_matches_oracle(env, B, gather):
    torch, sc = env["torch"], env["scorer"]
    rng = np.random.default_rng(B)
    th = _theta(B)
    n = B * 2 if gather else B
    X = rng.uniform(-1, 1, (n, mo.IN)).astype(np.float32)
    y = (rng.uniform(size=n) < 0.3).astype(np.uint8)
    idx = rng.choice(n, B, replace=False).astype(np.int32) if gather else None
    s = sc.CandidateScorer(env["dev"], params=th)
    d_idx = torch.from_numpy(idx).to(env["dev"]) if gather else None
    g = s.grad(_bits(env, X), torch.from_numpy(y).to(env["dev"]), d_idx, pos_weight=2.0)
    g = g.cpu().numpy().astype(np.float64)
    Xs, ys = (X[idx], y[idx]) if gather else (X, y)
    go = mo.backward(th.astype(np.float64), Xs, ys.astype(np.float64), pos_weight=2.0)
    lo = mo.loss(th.astype(np.float64), Xs, ys.astype(np.float64), 2.0)
    assert abs(float(s.loss.item()) - lo) <= 1e-3 * max(1.0, abs(lo))
    for name, (off, shape) in mo.OFFSETS.items():
        k = int(np.prod(shape))
        a, b = g[off:off + k], go[off:off + k]
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        assert err <= 2e-2, (name, err)


def test_adam_step_matches_oracle(env):
    torch, sc = env["torch"], env["scorer"]
    rng = np.random.default_rng(9)
    th = _theta(9)
    s = sc.CandidateScorer(env["dev"], params=th)
    m = np.zeros(mo.NPARAM)
    v = np.zeros(mo.NPARAM)
    ref = th.astype(np.float64)
    for step in (1, 2, 3):
        g = rng.normal(0, 1e-2, mo.NPARAM).astype(np.float32)
        s.grads.copy_(torch.from_numpy(g))
        s.adam_step(lr=1e-3)
        ref, m, v = mo.adam(ref, g.astype(np.float64), m, v, step, lr=1e-3)
        got = s.params.cpu().numpy().astype(np.float64)
        np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-7)
    wb = s.wbf.cpu().view(torch.bfloat16).to(torch.float64).numpy()
    P = mo.unpack(s.params.cpu().numpy().astype(np.float64))
    exp = np.concatenate([mo.bf16(P[k]).ravel() for k in mo.WEIGHTS_BF16])
    np.testing.assert_array_equal(wb, exp)


def test_training_learns_a_separable_rule(env):
    """Labels from a fixed nonlinear rule of the inputs: 300 Adam steps on the device bring the
    loss well below its start and the accuracy above 0.9."""
    torch, sc = env["torch"], env["scorer"]
    rng = np.random.default_rng(11)
    n = 65536
    X = rng.uniform(-1, 1, (n, mo.IN)).astype(np.float32)
    y = ((X[:, 0] * X[:, 1] + 0.5 * X[:, 2] - 0.3 * X[:, 3] ** 2) > 0).astype(np.uint8)
    s = sc.CandidateScorer(env["dev"], seed=3)
    Xb = _bits(env, X)
    yd = torch.from_numpy(y).to(env["dev"])
    hist = s.train(Xb, yd, epochs=8, batch=2048, lr=2e-3, pos_weight=1.0)
    p = s.infer(Xb).cpu().numpy()
    acc = float(np.mean((p >= 0.5) == (y == 1)))
    print(f"[mlp training] loss by epoch {['%.3f' % h for h in hist]}, accuracy {acc:.3f}")
    assert hist[-1] < 0.5 * hist[0]
    assert acc > 0.9
