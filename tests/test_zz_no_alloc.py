"""paces_reassign_noalloc (the no-allocation form of paces_reassign, its tile bounds in a static
device array) against paces_reassign through the C ABI: bit-identical amplitudes on two calls in
a row, the second smaller (the array keeps the first call's contents), and sizes beyond the
array's capacity rejected before any launch."""

import ctypes as C

import numpy as np
import pytest

from oracle import paces_oracle as po

pytestmark = pytest.mark.gpu


def _rows(rng, n, W):
    """n sorted unique rows of W uint64 words (as int64 bit patterns)."""
    raw = rng.integers(0, 1 << 62, size=(int(n * 1.1), W), dtype=np.int64).view(np.uint64)
    raw[:, 0] >>= np.uint64(rng.integers(0, 40))  # share leading words so searches go deep
    rows = po.unique_rows(raw)[:n]
    return rows.view(np.int64)


def _call(lib, fn, old, psi_old, new, W, ws=None):
    import torch
    from paces_amd import _hip
    out = torch.full((new.shape[0],), complex("nan+nanj"), dtype=torch.complex128,
                     device=new.device)
    args = [_hip.ptr(old), old.shape[0], _hip.ptr(torch.view_as_real(psi_old)), _hip.ptr(new),
            new.shape[0], W, _hip.ptr(torch.view_as_real(out))]
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    if ws is None:
        rc = fn(*args, st)
    else:
        wsp, wsn = ws.get(lib.paces_reassign_workspace_bytes(new.shape[0]))
        rc = fn(*args, wsp, wsn, st)
    assert rc == 0, lib.paces_strerror(rc)
    torch.cuda.synchronize()
    return out


def test_reassign_noalloc_equals_reassign_twice():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from paces_amd import _hip
    lib = _hip.load()
    dev = torch.device("cuda", torch.cuda.current_device())
    ws = _hip.Workspace(dev)
    rng = np.random.default_rng(7)
    for (M, R, W) in ((200_000, 300_000, 2), (80_000, 50_000, 3)):
        print(f"[no-alloc] M={M} R={R} W={W}", flush=True)
        old_h = _rows(rng, M, W)
        # half of the new rows are old rows, the rest new
        pick = old_h[rng.choice(M, size=min(M, R // 2), replace=False)]
        fresh = _rows(rng, R - pick.shape[0], W)
        new_h = po.unique_rows(np.concatenate([pick, fresh]).view(np.uint64))
        new_h = np.ascontiguousarray(new_h.view(np.int64))
        old = torch.from_numpy(old_h).to(dev)
        new = torch.from_numpy(new_h).to(dev)
        psi = torch.from_numpy(rng.normal(size=M) + 1j * rng.normal(size=M)).to(dev)
        want = _call(lib, lib.paces_reassign, old, psi, new, W, ws)
        got = _call(lib, lib.paces_reassign_noalloc, old, psi, new, W)
        assert torch.equal(torch.view_as_real(got), torch.view_as_real(want))
        hits = int((want != 0).sum().item())
        assert hits >= pick.shape[0]  # the shared rows carried their amplitudes over
    too_many = ((4 << 20) // 16 - 1) * 256 + 1
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.paces_reassign_noalloc(_hip.ptr(old), old.shape[0],
                                    _hip.ptr(torch.view_as_real(psi)), _hip.ptr(new), too_many, W,
                                    _hip.ptr(torch.view_as_real(psi)), st)
    assert rc == -5 and b"scratch" in lib.paces_strerror(rc)
