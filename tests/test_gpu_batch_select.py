"""The batched per-trajectory top-k select against a NumPy stable sort, through the C ABI.

Each trajectory's segment keeps its k rows of largest |psi|^2; among equal keys the later rows
are kept (the reference sorts stably ascending and takes the end). The keys are drawn from a few
values, so most cuts fall inside a run of ties; segments include empty ones and k of 0, 1,
R_t - 1, R_t and R_t + 5. The row-emitting form (paces_batch_select_topk, keys formed from psi or
given) and the flag form (paces_batch_select_flags, with the expanded-basis bit) must both give
the reference's rows bit for bit.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _reference(w, seg, kpt):
    keep = np.zeros(w.shape[0], dtype=bool)
    for t in range(len(seg) - 1):
        a, b = seg[t], seg[t + 1]
        k = min(int(kpt[t]), b - a)
        if k <= 0:
            continue
        order = np.argsort(w[a:b], kind="stable")  # ascending, equal keys in row order
        keep[a + order[b - a - k:]] = True
    return keep


def _case(seed, sizes, kinds):
    rng = np.random.default_rng(seed)
    seg = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    R = int(seg[-1])
    levels = np.array([0.0, 1e-12, 3e-5, 0.01, 0.25, 0.5])
    amp = levels[rng.integers(0, len(levels), R)]
    # a few distinct |psi|^2 values, each on the real or the imaginary axis
    psi = np.zeros((R, 2))
    swap = rng.integers(0, 2, R).astype(bool)
    psi[~swap, 0] = np.sqrt(amp[~swap])
    psi[swap, 1] = -np.sqrt(amp[swap])
    kpt = []
    for t, n in enumerate(sizes):
        kind = kinds[t % len(kinds)]
        kpt.append({"zero": 0, "one": 1, "less": max(n - 1, 0), "all": n, "more": n + 5,
                    "half": n // 2, "third": n // 3}[kind])
    return psi, seg, np.array(kpt, dtype=np.int64)


@pytest.mark.parametrize("seed,sizes", [
    (1, [0, 1, 7, 300, 2048, 2049, 5000, 0, 12345, 3]),
    (2, [70000, 0, 0, 131072, 9, 40000]),
    (3, [1] * 40 + [100] * 20),
])
def test_batch_select_matches_stable_sort(seed, sizes):
    import torch
    from paces_amd import _hip
    from paces_amd.config import devices
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not devices.configure_called:
        devices.configure()
    lib = _hip.load()
    dev = devices.vector_dev
    kinds = ["half", "one", "less", "all", "more", "third", "zero"]
    psi, seg, kpt = _case(seed, sizes, kinds)
    R, ntraj = psi.shape[0], len(sizes)
    w = psi[:, 0] * psi[:, 0] + psi[:, 1] * psi[:, 1]
    exp = _reference(w, seg, kpt)
    rng = np.random.default_rng(seed + 100)
    W = 2
    rows = np.sort(rng.integers(0, 2 ** 62, size=(R, W), dtype=np.int64), axis=0)
    d_psi = torch.from_numpy(psi.copy()).to(dev)
    d_rows = torch.from_numpy(rows).to(dev)
    d_seg = torch.from_numpy(seg).to(dev)
    d_k = torch.from_numpy(kpt).to(dev)
    ws = _hip.Workspace(dev)
    wsp, wsn = ws.get(lib.paces_batch_select_workspace_bytes(R, ntraj))
    keys = torch.from_numpy(w.view(np.int64).copy()).to(dev)
    for given_keys in (False, True):
        out = torch.empty((max(R, 1), W), dtype=torch.int64, device=dev)
        cnt = torch.empty(1, dtype=torch.int32, device=dev)
        _hip.check(lib.paces_batch_select_topk_keys(
            _hip.ptr(d_psi), _hip.ptr(keys) if given_keys else None, _hip.ptr(d_rows), R, W,
            _hip.ptr(d_seg), ntraj, _hip.ptr(d_k), _hip.ptr(out), _hip.ptr(cnt), wsp, wsn,
            _hip.stream()))
        n = int(cnt.item())
        assert n == int(exp.sum())
        np.testing.assert_array_equal(out[:n].cpu().numpy(), rows[exp])
    in_be = torch.from_numpy((np.arange(R) % 3 == 0).astype(np.uint8)).to(dev)
    flags = torch.empty(max(R, 1), dtype=torch.uint8, device=dev)
    cnt = torch.empty(1, dtype=torch.int32, device=dev)
    _hip.check(lib.paces_batch_select_flags(
        _hip.ptr(d_psi), _hip.ptr(keys), R, _hip.ptr(d_seg), ntraj, _hip.ptr(d_k),
        _hip.ptr(in_be), _hip.ptr(flags), _hip.ptr(cnt), wsp, wsn, _hip.stream()))
    f = flags[:R].cpu().numpy()
    np.testing.assert_array_equal(f & 1, exp.astype(np.uint8))
    np.testing.assert_array_equal(f >> 7, (np.arange(R) % 3 == 0).astype(np.uint8))
    assert int(cnt.item()) == int(exp.sum())
