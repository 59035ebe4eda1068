"""Candidate scorer: the BF16-MFMA network of libpaces_hip.so (paces_mlp_*).

No counterpart in the reference. The network maps PACES_MLP_IN features of a candidate row to
the probability that the row survives the next top-maxstates selection. Training uses retained
and discarded labels logged from the plain method (``LabelLog``). Forward, backward, Adam and
inference all run as the library's HIP kernels; this module only holds the buffers and the loop.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy
import torch

from .. import _hip
from .._hip import check, ptr, stream

IN = 64
HIDDEN = 256
NPARAM = 148481
NBF16 = 147456

_SHAPES = [("Wt1", (HIDDEN, IN)), ("b1", (HIDDEN,)), ("Wt2", (HIDDEN, HIDDEN)), ("b2", (HIDDEN,)),
           ("Wt3", (HIDDEN, HIDDEN)), ("b3", (HIDDEN,)), ("w4", (HIDDEN,)), ("b4", (1,))]


def init_params(seed=0):
    """He-uniform hidden weights (limit sqrt(6 / fan_in)), output weights uniform with limit
    sqrt(3 / 256), zero biases; float32 vector in the library's parameter layout."""
    rng = numpy.random.default_rng(seed)
    parts = []
    for name, shape in _SHAPES:
        if name.startswith("Wt"):
            lim = math.sqrt(6.0 / shape[1])
            parts.append(rng.uniform(-lim, lim, shape).ravel())
        elif name == "w4":
            lim = math.sqrt(3.0 / shape[0])
            parts.append(rng.uniform(-lim, lim, shape).ravel())
        else:
            parts.append(numpy.zeros(int(numpy.prod(shape))))
    out = numpy.concatenate(parts).astype(numpy.float32)
    assert out.size == NPARAM
    return out


def to_bf16_bits(x):
    """float tensor -> bfloat16 bit patterns as int16 (round to nearest even), on its device."""
    return x.to(torch.bfloat16).view(torch.int16)


class CandidateScorer:
    """Parameters, Adam moments and BF16 weight copies of the scorer on one device."""

    def __init__(self, device, params=None, seed=0):
        self.device = device
        p = init_params(seed) if params is None else numpy.asarray(params, dtype=numpy.float32)
        if p.shape != (NPARAM,):
            raise ValueError(f"the scorer has {NPARAM} parameters")
        self.params = torch.from_numpy(p.copy()).to(device)
        self.wbf = torch.empty(NBF16, dtype=torch.int16, device=device)
        self.m = torch.zeros(NPARAM, dtype=torch.float32, device=device)
        self.v = torch.zeros(NPARAM, dtype=torch.float32, device=device)
        self.grads = torch.zeros(NPARAM, dtype=torch.float32, device=device)
        self.loss = torch.zeros(1, dtype=torch.float32, device=device)
        self.step = 0
        self.tile_rows = 0  # 0: the library default; the autotuner may set 64, 128 or 256
        self._ws = _hip.Workspace(device)
        self._lib = _hip.load()
        self.refresh_bf16()

    def refresh_bf16(self):
        check(self._lib.paces_mlp_weights_bf16(ptr(self.params), ptr(self.wbf), stream()),
              "mlp_weights_bf16")

    def state_dict(self):
        return {"params": self.params.cpu().numpy(), "m": self.m.cpu().numpy(),
                "v": self.v.cpu().numpy(), "step": numpy.array([self.step])}

    def load_state_dict(self, d):
        self.params.copy_(torch.from_numpy(numpy.asarray(d["params"], dtype=numpy.float32)))
        self.m.copy_(torch.from_numpy(numpy.asarray(d["m"], dtype=numpy.float32)))
        self.v.copy_(torch.from_numpy(numpy.asarray(d["v"], dtype=numpy.float32)))
        self.step = int(numpy.asarray(d["step"]).ravel()[0])
        self.refresh_bf16()

    def grad(self, X, labels, idx=None, pos_weight=1.0):
        """Gradient of the weighted BCE over rows X[idx] (all rows when idx is None) into
        self.grads; the loss lands in self.loss. X: (n, IN) int16 bf16 bits; labels: uint8."""
        B = int(idx.numel()) if idx is not None else int(X.shape[0])
        if X.dtype != torch.int16 or X.shape[1] != IN or labels.dtype != torch.uint8:
            raise ValueError("X must be (n, 64) bf16 bits in int16 and labels uint8")
        wsp, wsn = self._ws.get(self._lib.paces_mlp_train_workspace_bytes(B))
        check(self._lib.paces_mlp_grad(ptr(X), ptr(idx) if idx is not None else None,
                                       ptr(labels), B, float(pos_weight), ptr(self.params),
                                       ptr(self.wbf), ptr(self.grads), ptr(self.loss), wsp, wsn,
                                       stream()), "mlp_grad")
        return self.grads

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        check(self._lib.paces_mlp_adam(ptr(self.params), ptr(self.grads), ptr(self.m),
                                       ptr(self.v), self.step, float(lr), float(beta1),
                                       float(beta2), float(eps), ptr(self.wbf), stream()),
              "mlp_adam")

    def infer(self, X, want_logit=False, tile_rows=None):
        """Retain probabilities (or logits) of the rows of X. tile_rows (64, 128 or 256) overrides
        the workgroup tile; by default the autotuned choice, else the library default."""
        n = int(X.shape[0])
        out = torch.empty(max(n, 1), dtype=torch.float32, device=self.device)
        tile = self.tile_rows if tile_rows is None else int(tile_rows)
        check(self._lib.paces_mlp_infer_tiled(ptr(X) if n else None, n, ptr(self.params),
                                              ptr(self.wbf), None if want_logit else ptr(out),
                                              ptr(out) if want_logit else None, tile, stream()),
              "mlp_infer")
        return out[:n]

    def train(self, X, labels, epochs=1, batch=8192, lr=1e-3, pos_weight=1.0, seed=0,
              log_every=0):
        """Minibatch Adam over (X, labels) on the device; rows are drawn by a seeded permutation
        per epoch and gathered by the kernels. The default unweighted loss keeps the output a
        calibrated retain probability, so a pruning threshold reads as one. Returns the mean
        loss of each epoch."""
        n = int(X.shape[0])
        if n == 0:
            return []
        gen = torch.Generator(device="cpu")
        gen.manual_seed(seed)
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=gen).to(torch.int32).to(self.device)
            tot = torch.zeros(1, dtype=torch.float32, device=self.device)
            nb = 0
            for s in range(0, n, batch):
                idx = perm[s:s + batch]
                self.grad(X, labels, idx, pos_weight)
                self.adam_step(lr)
                tot += self.loss
                nb += 1
                if log_every and nb % log_every == 0:
                    print(f"      scorer epoch {ep} batch {nb}: loss {float(self.loss.item()):.4f}",
                          flush=True)
            hist.append(float(tot.item()) / nb)
        return hist
