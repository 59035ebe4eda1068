"""CPU-side checks of the C ABI: the library loads and exports every symbol the header declares."""

import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "paces_hip.h")
LIB = os.path.join(ROOT, "paces_amd", "libpaces_hip.so")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(paces_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    assert "paces_lower_bound" in names and "paces_taylor_term" in names
    assert len(names) >= 30


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (paces_[a-z0-9_]+)", out))
    missing = [n for n in _declared() if n not in exported]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_ctypes_binding_covers_header_and_host_only_calls():
    from paces_amd import _hip
    lib = _hip.load()
    assert sorted(_declared()) == _hip.declared_symbols()
    assert lib.paces_version() >= 10000
    assert lib.paces_strerror(-4) == b"workspace too small"
    import ctypes as C
    h = _hip.Holstein()
    h.nchain, h.lattice_l, h.geo_dim, h.periodic, h.nterms = 16, 16, 1, 0, 2
    h.terms[0] = _hip.Term(_hip.TERM_HOP, 1, -1.0)
    h.terms[1] = _hip.Term(_hip.TERM_VIB, 0, 1.0)
    assert lib.paces_holstein_num_runs(C.byref(h)) == 4
    h.periodic = 1
    assert lib.paces_holstein_num_runs(C.byref(h)) == 6
    h.geo_dim, h.lattice_l, h.nchain = 2, 4, 16
    assert lib.paces_holstein_num_runs(C.byref(h)) == 10
    assert lib.paces_scan_workspace_bytes(10 ** 7) > 0
    # argument validation runs on the host and returns a status, no device work
    assert lib.paces_lower_bound(None, -1, None, 0, 1, 1, None, None, None, None, 0, None) == -1
    assert lib.paces_merge_unique(None, 1, None, None, None, None, None, 0, None) == -1
    # the split pairs entry points need the prior part, a valid one, and one value form
    h.periodic, h.geo_dim, h.lattice_l, h.nchain = 0, 1, 16, 16
    offs = (C.c_int64 * 5)(0, 1, 2, 3, 4)
    ws = C.c_void_p(256)
    nnz = C.c_void_p(256)
    assert lib.paces_holstein_pairs_count_split(8, 4, offs, None, None, None, None, C.byref(h),
                                                nnz, ws, 1 << 20, None) == -1
    bad = _hip.PriorImages(-1, offs, None, None, None)
    assert lib.paces_holstein_pairs_count_split(8, 4, offs, None, None, None, C.byref(bad),
                                                C.byref(h), nnz, ws, 1 << 20, None) == -1
    unsorted = _hip.PriorImages(2, (C.c_int64 * 5)(0, 3, 2, 3, 4), None, None, None)
    assert lib.paces_holstein_pairs_count_split(8, 4, offs, None, None, None, C.byref(unsorted),
                                                C.byref(h), nnz, ws, 1 << 20, None) == -1
    arr = (C.c_void_p * _hip.MAX_TERMS)(256, 256)
    good = _hip.PriorImages(2, offs, 256, 256, 256)
    for vals in ((arr, arr, None), (None, None, None)):
        assert lib.paces_holstein_pairs_fill_split(8, 4, offs, None, None, None, None,
                                                   C.byref(good), None, None, C.byref(h), arr,
                                                   arr, vals[0], vals[1], vals[2], ws, 1 << 20,
                                                   None) == -1


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_struct_layouts_match_the_library():
    """ctypes mirrors of structs the library reads from device memory have the C sizes."""
    import ctypes as C
    from paces_amd import _hip
    lib = _hip.load()
    assert lib.paces_batch_series_desc_bytes() == C.sizeof(_hip.BatchSeries)
    assert C.sizeof(_hip.BatchTaylor) == 40  # 3 doubles, 2 int32, the last exact norm
    assert C.sizeof(_hip.TaylorState) == 64


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_slot_layout_host_checks():
    """paces_holstein_slot_layout: a batched open chain gets four slots per row (the two hops,
    then the two vib ops, each term's slots in column order), a code field wide enough for the
    phonon counts and a column field for the rows; periodic or unbatched models and row counts
    beyond the column field are refused on the host."""
    import ctypes as C
    from paces_amd import _hip
    lib = _hip.load()
    h = _hip.Holstein()
    h.nchain, h.lattice_l, h.geo_dim, h.periodic, h.nterms = 16, 16, 1, 0, 2
    h.terms[0] = _hip.Term(_hip.TERM_HOP, 1, -1.0)
    h.terms[1] = _hip.Term(_hip.TERM_VIB, 0, 1.0)
    h.sqrt_table_len = 8
    h.batched, h.ntraj = 1, 4
    s = _hip.Slots()
    assert lib.paces_holstein_slot_layout(C.byref(h), 1000, C.byref(s)) == 0
    assert (s.nterms, s.nslots, s.col_bits, s.sqrt_len, s.stride) == (2, 4, 29, 8, 1000)
    assert list(s.slot_off)[:3] == [0, 2, 4] and list(s.vib)[:2] == [0, 1]
    assert lib.paces_holstein_slot_layout(C.byref(h), (1 << 29) - 1, C.byref(s)) == 0
    assert lib.paces_holstein_slot_layout(C.byref(h), 1 << 29, C.byref(s)) == _hip.E_UNSUPPORTED
    h.sqrt_table_len = 16  # cap 15: four code bits
    assert lib.paces_holstein_slot_layout(C.byref(h), 10, C.byref(s)) == 0
    assert s.col_bits == 28
    h.periodic = 1
    assert lib.paces_holstein_slot_layout(C.byref(h), 10, C.byref(s)) == _hip.E_UNSUPPORTED
    h.periodic, h.batched = 0, 0
    assert lib.paces_holstein_slot_layout(C.byref(h), 10, C.byref(s)) == _hip.E_UNSUPPORTED
    # the pairs entry point checks its arguments before any launch
    h.batched = 1
    offs = (C.c_int64 * 5)(0, 1, 2, 3, 4)
    assert lib.paces_holstein_pairs_slots(8, 4, offs, None, None, None, None, None, C.byref(h),
                                          None, None) == -1
    assert lib.paces_holstein_pairs_slots(8, 4, offs, None, None, None, None, None, C.byref(h),
                                          C.c_void_p(4096 + 8), None) == -1  # 16-byte aligned
