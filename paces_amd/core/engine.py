"""Device primitives of the step, each a thin call into libpaces_hip.so.

Rows are (n, W) torch.int64 tensors holding uint64 bit patterns; complex vectors are
torch.complex128 tensors. Every function runs on the current device and stream.
"""

from __future__ import annotations

import ctypes as C
import math

import torch

from .. import _hip
from .._hip import check, ptr, stream

lib = _hip.load()


class CsrMatrix:
    """Canonical CSR of one Hamiltonian term (int32 indptr/indices, float64 data; float32 data
    for the batched mixed-precision engine, which reads nothing else)."""

    def __init__(self, indptr, indices, data, shape):
        self.indptr = indptr
        self.indices = indices
        self.data = data
        self.shape = shape

    @property
    def nnz(self):
        return int(self.indices.numel())

    def dot(self, x):
        """M.x for a complex128 vector (the cuSPARSE product the reference calls)."""
        zero = torch.zeros(self.shape[0], dtype=torch.float64, device=x.device)
        m = self if self.data.dtype == torch.float64 else \
            CsrMatrix(self.indptr, self.indices, self.data.double(), self.shape)
        return spmv(zero, [m], x)

    def to_scipy(self):
        import scipy.sparse
        return scipy.sparse.csr_matrix((self.data.cpu().numpy(), self.indices.cpu().numpy(),
                                        self.indptr.cpu().numpy()), shape=self.shape)


class CodedCsrMatrix:
    """A term's CSR with coded values (paces_csr_coded_t): int32 indptr/indices and, for a vib
    term, one code byte per entry (the pair's upper phonon count; None for a hopping term). The
    float64 values -- the trajectory's term value, times sqrt(code) for a vib term -- are formed
    on first use of ``data`` by ``value_fn(self)``."""

    def __init__(self, indptr, indices, code, shape, term, value_fn):
        self.indptr = indptr
        self.indices = indices
        self.code = code
        self.shape = shape
        self.term = term
        self._value_fn = value_fn
        self._data = None

    @property
    def nnz(self):
        return int(self.indices.numel())

    @property
    def data(self):
        if self._data is None:
            self._data = self._value_fn(self)
        return self._data

    def dot(self, x):
        return CsrMatrix(self.indptr, self.indices, self.data, self.shape).dot(x)

    def to_scipy(self):
        return CsrMatrix(self.indptr, self.indices, self.data, self.shape).to_scipy()


class SlotMatrix:
    """Every off-diagonal term of a batched Hamiltonian in one (nslots, R) array of uint32 slots
    (paces_slots_t, held as torch.int32, slot-major): slot layout.slot_off[t] + r of row i is row
    i's entry of term t whose column has rank r among the term's ops -- its column in the low
    col_bits bits and, for a vib term, the pair's upper phonon count above them -- or SLOT_EMPTY
    (-1 as int32). ``term(t, value_fn)`` views term t as a coded CSR."""

    def __init__(self, slots, layout, shape):
        self.slots = slots
        self.layout = layout
        self.shape = shape
        layout.slots = slots.data_ptr()

    def term(self, t, value_fn):
        return SlotTermMatrix(self, t, value_fn)


class SlotTermMatrix(CodedCsrMatrix):
    """Term t of a SlotMatrix with the CodedCsrMatrix interface. indptr, indices and code are
    formed from the slots on first use (a row's occupied slots of the term, in slot order, are
    its CSR row: columns ascending); the batched series reads the slots themselves."""

    def __init__(self, sm, t, value_fn):
        self.slotmat = sm
        self.shape = sm.shape
        self.term = t
        self._value_fn = value_fn
        self._data = None
        self._csr = None

    def _form(self):
        if self._csr is None:
            lay = self.slotmat.layout
            a, b = lay.slot_off[self.term], lay.slot_off[self.term + 1]
            sub = self.slotmat.slots[a:b].t()  # (R, the term's slots)
            occ = sub != -1
            counts = occ.sum(dim=1, dtype=torch.int32)
            indptr = torch.zeros(sub.shape[0] + 1, dtype=torch.int32, device=sub.device)
            torch.cumsum(counts, 0, out=indptr[1:])
            words = sub[occ].long() & 0xFFFFFFFF
            indices = (words & ((1 << lay.col_bits) - 1)).to(torch.int32)
            code = (words >> lay.col_bits).to(torch.uint8) if lay.vib[self.term] else None
            self._csr = (indptr, indices, code)
        return self._csr

    @property
    def indptr(self):
        return self._form()[0]

    @property
    def indices(self):
        return self._form()[1]

    @property
    def code(self):
        return self._form()[2]


def csr_coded(mats, sqrt_len, term_value_ptr, sqrt_table_ptr):
    """paces_csr_coded_t of a list of CodedCsrMatrix (term order)."""
    s = _hip.CsrCoded()
    s.nterms = len(mats)
    if len(mats) > _hip.MAX_TERMS:
        raise ValueError("too many Hamiltonian terms")
    s.sqrt_len = int(sqrt_len)
    for t, m in enumerate(mats):
        s.indptr[t] = m.indptr.data_ptr()
        s.indices[t] = m.indices.data_ptr() if m.nnz else m.indptr.data_ptr()
        s.code[t] = (m.code.data_ptr() if m.nnz else m.indptr.data_ptr()) \
            if m.code is not None else None
    s.term_value = term_value_ptr
    s.sqrt_table = sqrt_table_ptr
    return s


def csr_set(mats):
    s = _hip.CsrSet()
    s.nterms = len(mats)
    if len(mats) > _hip.MAX_TERMS:
        raise ValueError("too many Hamiltonian terms")
    for t, m in enumerate(mats):
        if m.data.dtype != torch.float64:
            raise ValueError("the FP64 engine needs float64 CSR values")
        s.indptr[t] = m.indptr.data_ptr()
        s.indices[t] = m.indices.data_ptr() if m.nnz else m.indptr.data_ptr()
        s.data[t] = m.data.data_ptr() if m.nnz else m.indptr.data_ptr()
    return s


def lower_bound(book, q, want_hit=True, count_misses=False, q_sorted=False, ws=None):
    """Lexicographic lower_bound of q's rows in the sorted book (int32 indices, uint8 hits).
    q_sorted=True promises ascending queries and uses the partitioned LDS-staged kernel."""
    M, Q = book.shape[0], q.shape[0]
    W = q.shape[1]
    if book.shape[1] != W:
        raise ValueError("Incompatible array shapes.")
    idx = torch.empty(Q, dtype=torch.int32, device=q.device)
    hit = torch.empty(Q, dtype=torch.uint8, device=q.device) if want_hit else None
    miss = torch.zeros(1, dtype=torch.int32, device=q.device) if count_misses else None
    wsp, wsn = None, 0
    if q_sorted:
        ws = ws if ws is not None else _hip.Workspace(q.device)
        wsp, wsn = ws.get(lib.paces_lower_bound_workspace_bytes(Q))
    check(lib.paces_lower_bound(ptr(book) if M else None, M, ptr(q) if Q else None, Q, W,
                                1 if q_sorted else 0, ptr(idx) if Q else None,
                                ptr(hit) if (hit is not None and Q) else None, ptr(miss), wsp, wsn,
                                stream()), "lower_bound")
    return idx, hit, miss


def merge_unique(runs, W, ws, want_from0=False, want_cand_pos=False, diag=None, run_codes=None):
    """Sorted union of sorted unique row runs; returns (rows, from_run0 flags or None) and, with
    want_cand_pos, the output index of every input row as a third element. With diag =
    (layout, holstein struct, torch.float64 or torch.float32), the Holstein diagonal of the rows
    comes from the same pass (paces_merge_unique_diag) as a last element. With run_codes (one
    byte per run), each row's code (the code of the run its kept copy came from,
    paces_merge_unique_coded) is the last element instead; diag and run_codes do not combine."""
    if diag is not None and run_codes is not None:
        raise ValueError("merge_unique takes diag or run_codes, not both")
    if len(runs) > _hip.MAX_RUNS:
        raise ValueError("too many runs")
    R = _hip.Runs()
    R.K = len(runs)
    total = 0
    dev = None
    for k, (t, off, n) in enumerate(runs):
        dev = t.device
        R.ptr[k] = t.data_ptr() + off * 8 * W if n else 0
        R.len[k] = n
        total += n
    out = torch.empty((max(total, 1), W), dtype=torch.int64, device=dev)
    cnt = torch.empty(1, dtype=torch.int32, device=dev)
    from0 = torch.empty(max(total, 1), dtype=torch.uint8, device=dev) if want_from0 else None
    cpos = torch.empty(max(total, 1), dtype=torch.int32, device=dev) if want_cand_pos else None
    wsp, wsn = ws.get(lib.paces_merge_unique_workspace_bytes(total, R.K, W))
    d = None
    if run_codes is not None:
        if len(run_codes) != len(runs):
            raise ValueError("one code per run")
        codes = (C.c_uint8 * len(run_codes))(*run_codes)
        d = torch.empty(max(total, 1), dtype=torch.uint8, device=dev)
        check(lib.paces_merge_unique_coded(C.byref(R), W, ptr(out), ptr(cnt), ptr(from0),
                                           ptr(cpos), codes, ptr(d), wsp, wsn, stream()),
              "merge_unique_coded")
    elif diag is None:
        check(lib.paces_merge_unique(C.byref(R), W, ptr(out), ptr(cnt), ptr(from0), ptr(cpos),
                                     wsp, wsn, stream()), "merge_unique")
    else:
        lay, h, ddt = diag
        d = torch.empty(max(total, 1), dtype=ddt, device=dev)
        f64 = ddt == torch.float64
        check(lib.paces_merge_unique_diag(C.byref(R), W, ptr(out), ptr(cnt), ptr(from0),
                                          ptr(cpos), C.byref(lay), C.byref(h),
                                          ptr(d) if f64 else None, None if f64 else ptr(d), wsp,
                                          wsn, stream()), "merge_unique_diag")
    n = int(cnt.item())
    res = [out[:n], (from0[:n] if want_from0 else None)]
    if want_cand_pos:
        res.append(cpos)
    if d is not None:
        res.append(d[:n])
    return tuple(res)


def sort_rows(rows, ws, unique=True, want_perm=False):
    """Lexicographically sorted rows (LSD radix sort on the device), duplicates dropped when
    unique; with want_perm also the source index of each output row. The digits at which all
    rows agree are found first (one small device-to-host copy) and skipped."""
    n, W = rows.shape
    out = torch.empty((max(n, 1), W), dtype=torch.int64, device=rows.device)
    perm = torch.empty(max(n, 1), dtype=torch.int32, device=rows.device) if want_perm else None
    cnt = torch.empty(1, dtype=torch.int32, device=rows.device)
    if n == 0:
        return (out[:0], perm[:0]) if want_perm else out[:0]
    mask = torch.empty(W, dtype=torch.int64, device=rows.device)
    check(lib.paces_rows_varying_bits(ptr(rows), n, W, ptr(mask), stream()), "rows_varying_bits")
    host = (C.c_uint64 * W)(*[int(x) & 0xFFFFFFFFFFFFFFFF for x in mask.cpu().tolist()])
    wsp, wsn = ws.get(lib.paces_sort_rows_workspace_bytes(n, W))
    check(lib.paces_sort_rows(ptr(rows), n, W, host, 1 if unique else 0, ptr(out),
                              ptr(perm) if want_perm else None, ptr(cnt), wsp, wsn, stream()),
          "sort_rows")
    m = int(cnt.item()) if unique else n
    return (out[:m], perm[:m]) if want_perm else out[:m]


def select_topk(psi, rows, k, ws, want_index=False):
    """Rows of the k largest |psi| (ties keep the highest indices), in index order; with
    want_index also their indices (int32)."""
    R, W = rows.shape
    k = int(min(k, R))
    out = torch.empty((k, W), dtype=torch.int64, device=rows.device)
    idx = torch.empty(max(k, 1), dtype=torch.int32, device=rows.device) if want_index else None
    if k > 0:
        wsp, wsn = ws.get(lib.paces_select_workspace_bytes(R))
        check(lib.paces_select_topk(ptr(torch.view_as_real(psi)), ptr(rows), R, W, k, ptr(out),
                                    ptr(idx) if want_index else None, wsp, wsn, stream()),
              "select_topk")
    if want_index:
        return out, idx[:k]
    return out


def reassign(old_rows, old_psi, new_rows, ws):
    R, W = new_rows.shape
    M = old_rows.shape[0]
    out = torch.empty(R, dtype=torch.complex128, device=new_rows.device)
    if R:
        wsp, wsn = ws.get(lib.paces_reassign_workspace_bytes(R))
        check(lib.paces_reassign(ptr(old_rows) if M else None, M,
                                 ptr(torch.view_as_real(old_psi)) if M else None, ptr(new_rows), R,
                                 W, ptr(torch.view_as_real(out)), wsp, wsn, stream()), "reassign")
    return out


def spmv(diag, mats, x, out=None):
    R = x.shape[0]
    y = torch.empty_like(x) if out is None else out
    s = csr_set(mats)
    if R:
        check(lib.paces_spmv(R, ptr(diag), C.byref(s), ptr(torch.view_as_real(x)),
                             ptr(torch.view_as_real(y)), stream()), "spmv")
    return y


class Reducer:
    def __init__(self, device):
        self.ws = _hip.Workspace(device)
        self.out = torch.empty(2, dtype=torch.float64, device=device)

    def norm2(self, x):
        """sum |x_i|^2 (device, deterministic order) -> float."""
        R = x.shape[0]
        wsp, wsn = self.ws.get(lib.paces_reduce_workspace_bytes(R))
        check(lib.paces_norm2(ptr(torch.view_as_real(x)) if R else None, R, ptr(self.out), wsp,
                              wsn, stream()), "norm2")
        return float(self.out[0].item())

    def norm(self, x):
        return math.sqrt(self.norm2(x))

    def norm2_masked(self, x, mask):
        """sum of |x_i|^2 over rows with mask[i] != 0 (uint8 mask) -> float."""
        R = x.shape[0]
        wsp, wsn = self.ws.get(lib.paces_reduce_workspace_bytes(R))
        check(lib.paces_norm2_masked(ptr(torch.view_as_real(x)) if R else None,
                                     ptr(mask) if R else None, R, ptr(self.out), wsp, wsn,
                                     stream()), "norm2_masked")
        return float(self.out[0].item())

    def norm2_masked_pair(self, x, mask_a, mask_b):
        """The sums of |x_i|^2 over mask_a's and over mask_b's rows, read in one transfer."""
        R = x.shape[0]
        wsp, wsn = self.ws.get(lib.paces_reduce_workspace_bytes(R))
        for k, mask in enumerate((mask_a, mask_b)):
            check(lib.paces_norm2_masked(ptr(torch.view_as_real(x)) if R else None,
                                         ptr(mask) if R else None, R, ptr(self.out[k:]), wsp,
                                         wsn, stream()), "norm2_masked")
        h = self.out.tolist()
        return h[0], h[1]

    def vdot(self, x, y):
        R = x.shape[0]
        wsp, wsn = self.ws.get(lib.paces_reduce_workspace_bytes(R))
        check(lib.paces_vdot(ptr(torch.view_as_real(x)) if R else None,
                             ptr(torch.view_as_real(y)) if R else None, R, ptr(self.out), wsp,
                             wsn, stream()), "vdot")
        h = self.out.tolist()
        return complex(h[0], h[1])


def axpby(a, x, b, y):
    """y <- a*x + b*y (complex host scalars)."""
    R = x.shape[0]
    a, b = complex(a), complex(b)
    if R:
        check(lib.paces_axpby(R, a.real, a.imag, ptr(torch.view_as_real(x)), b.real, b.imag,
                              ptr(torch.view_as_real(y)), stream()), "axpby")
    return y


def taylor_coeff(delta_t, j):
    """Imaginary magnitude c of -1j*dt/(j+1) computed as the reference does in Python."""
    return -((-1j * delta_t) / float(j + 1)).imag


TAYLOR_TOL = 1.1102230246251565e-16  # 2**-53 (paces/core/time_evolution.py:484)


def _tuned_shape(tuner, kind, R, nterms, run):
    """Launch shape for a Taylor term over R rows: the tuner's choice (timing run(shape) on
    scratch buffers the first time a size bucket is seen), or the default shape."""
    if tuner is None or R == 0:
        return None
    from . import autotune
    key = tuner.key(kind, R, nterms)
    shape = tuner.pick(key, autotune.SHAPES, run)
    return autotune.launch_cfg(shape)


def simple_taylor(diag, mats, v, delta_t, m_star=100, tol=TAYLOR_TOL, chunk=6, ws=None,
                  tuner=None, coeff=None):
    """res = sum_j (-i dt H)^j v / j! until (|t_j| + |t_{j+1}|)_inf / |res|_inf <= tol
    (paces/core/time_evolution.py:482-512). Returns (res, [converged, j_ind, c1+c2, ratio]).
    With a tuner, the term kernel's launch shape is the tuned one (identical results).
    coeff(j), when given, is term j's factor c in t_j = -i c H t_(j-1) (default: the reference's
    dt / (j + 1))."""
    if coeff is None:
        def coeff(j):
            return taylor_coeff(delta_t, j)
    R = v.shape[0]
    res = torch.empty_like(v)
    bufs = (torch.empty_like(v), torch.empty_like(v))
    state = torch.zeros(C.sizeof(_hip.TaylorState), dtype=torch.uint8, device=v.device)
    s = csr_set(mats)
    st = stream()
    ws = ws if ws is not None else _hip.Workspace(v.device)
    wsp, wsn = ws.get(lib.paces_taylor_workspace_bytes(R))

    def run(shape):
        # tol < 0 never stops, so every launch does the full term; scratch result and state
        scratch = torch.empty_like(v)
        sstate = torch.zeros_like(state)
        cfg = _hip.LaunchCfg(*shape)
        return lambda: check(lib.paces_taylor_term_tuned(
            R, ptr(diag), C.byref(s), ptr(torch.view_as_real(v)), ptr(torch.view_as_real(bufs[0])),
            ptr(torch.view_as_real(scratch)), 1e-3, 0, -1.0, ptr(sstate), wsp, wsn,
            C.byref(cfg), st), "taylor_term_tuned")

    cfg = _tuned_shape(tuner, "tay64", R, len(mats), run)
    check(lib.paces_taylor_begin(R, ptr(torch.view_as_real(v)), ptr(torch.view_as_real(res)),
                                 ptr(state), wsp, wsn, st), "taylor_begin")
    x_in = v
    j = 0
    host = None
    while j < m_star:
        for _ in range(chunk):
            if j >= m_star:
                break
            x_out = bufs[j % 2]
            check(lib.paces_taylor_term_tuned(R, ptr(diag), C.byref(s),
                                              ptr(torch.view_as_real(x_in)),
                                              ptr(torch.view_as_real(x_out)),
                                              ptr(torch.view_as_real(res)),
                                              coeff(j), j, tol, ptr(state), wsp,
                                              wsn, C.byref(cfg) if cfg else None, st),
                  "taylor_term")
            x_in = x_out
            j += 1
        host = _hip.TaylorState.from_buffer_copy(bytes(state.cpu().numpy()))
        if host.done:
            break
    return res, [bool(host.done), int(host.j_ind), float(host.c1_plus_c2), float(host.ratio)]


TAYLOR_TOL_MIXED = 5.960464477539063e-08  # 2**-24: FP32 terms carry no more


def to_f32(t):
    """FP32 copy of an FP64 tensor (round to nearest); an FP32 tensor is returned as is."""
    if t.dtype == torch.float32:
        return t
    out = torch.empty(t.shape, dtype=torch.float32, device=t.device)
    n = t.numel()
    if n:
        check(lib.paces_convert_f64_to_f32(n, ptr(t), ptr(out), stream()), "convert_f64_to_f32")
    return out


class MixedMats:
    """FP32 copies of the diagonal and of each term's CSR values for the mixed-precision series
    (values built in FP32 already are used as they are)."""

    def __init__(self, diag, mats):
        self.diag = to_f32(diag)
        self.data = [to_f32(m.data) for m in mats]
        s = _hip.CsrSet32()
        s.nterms = len(mats)
        for t, m in enumerate(mats):
            s.indptr[t] = m.indptr.data_ptr()
            s.indices[t] = m.indices.data_ptr() if m.nnz else m.indptr.data_ptr()
            s.data[t] = self.data[t].data_ptr() if m.nnz else m.indptr.data_ptr()
        self.cset = s


def simple_taylor_mixed(mm: MixedMats, v, delta_t, m_star=100, tol=TAYLOR_TOL_MIXED, chunk=4,
                        ws=None, tuner=None):
    """The Taylor series with complex64 terms (FP32 matrix and accumulation) summed into a
    complex128 result. Returns (res, [converged, j_ind, c1+c2, ratio])."""
    R = v.shape[0]
    res = torch.empty_like(v)
    bufs = (torch.empty(R, dtype=torch.complex64, device=v.device),
            torch.empty(R, dtype=torch.complex64, device=v.device),
            torch.empty(R, dtype=torch.complex64, device=v.device))
    state = torch.zeros(C.sizeof(_hip.TaylorState), dtype=torch.uint8, device=v.device)
    st = stream()
    ws = ws if ws is not None else _hip.Workspace(v.device)
    wsp, wsn = ws.get(lib.paces_taylor_workspace_bytes(R))
    check(lib.paces_taylor_begin_mixed(R, ptr(torch.view_as_real(v)), ptr(torch.view_as_real(res)),
                                       ptr(torch.view_as_real(bufs[2])), ptr(state), wsp, wsn, st),
          "taylor_begin_mixed")

    def run(shape):
        scratch = torch.empty_like(v)
        sstate = torch.zeros_like(state)
        cfg = _hip.LaunchCfg(*shape)
        return lambda: check(lib.paces_taylor_term_mixed_tuned(
            R, ptr(mm.diag), C.byref(mm.cset), ptr(torch.view_as_real(bufs[2])),
            ptr(torch.view_as_real(bufs[0])), ptr(torch.view_as_real(scratch)), 1e-3, 0, -1.0,
            ptr(sstate), wsp, wsn, C.byref(cfg), st), "taylor_term_mixed_tuned")

    cfg = _tuned_shape(tuner, "tay32", R, mm.cset.nterms, run)
    x_in = bufs[2]
    j = 0
    host = None
    while j < m_star:
        for _ in range(chunk):
            if j >= m_star:
                break
            x_out = bufs[j % 2]
            check(lib.paces_taylor_term_mixed_tuned(R, ptr(mm.diag), C.byref(mm.cset),
                                                    ptr(torch.view_as_real(x_in)),
                                                    ptr(torch.view_as_real(x_out)),
                                                    ptr(torch.view_as_real(res)),
                                                    taylor_coeff(delta_t, j), j, tol, ptr(state),
                                                    wsp, wsn, C.byref(cfg) if cfg else None, st),
                  "taylor_term_mixed")
            x_in = x_out
            j += 1
        host = _hip.TaylorState.from_buffer_copy(bytes(state.cpu().numpy()))
        if host.done:
            break
    # the sum is stored on odd terms: a series that stopped on an even term (its output in
    # bufs[0]) gets that term added here
    check(lib.paces_taylor_finish_mixed(R, ptr(torch.view_as_real(bufs[0])),
                                        ptr(torch.view_as_real(res)), ptr(state), st),
          "taylor_finish_mixed")
    return res, [bool(host.done), int(host.j_ind), float(host.c1_plus_c2), float(host.ratio)]
