"""ctypes boundary to libpaces_hip.so (declared in include/paces_hip.h).

Device buffers are torch tensors on the current ROCm device; their data pointers and the current
stream handle cross the boundary, nothing else. The library must be built (``python -m
paces_amd.build``); there is no fallback: a missing library raises at import of the engine.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpaces_hip.so")

MAX_RUNS = 32
MAX_TERMS = 4
MAX_ROW_WORDS = 16
MAX_SITES = 512
TERM_HOP = 0
TERM_VIB = 1

_lib = None


E_UNSUPPORTED = -2  # PACES_E_UNSUPPORTED


class PacesError(RuntimeError):
    pass


class Layout(C.Structure):
    _fields_ = [("W", C.c_int32), ("nfields", C.c_int32), ("d_off", C.c_void_p),
                ("d_width", C.c_void_p), ("d_maxdim", C.c_void_p), ("d_sum_mask", C.c_void_p),
                ("sum_bits", C.c_int32)]


def sum_masks(offsets, widths, W, first=1):
    """Per (word, in-field significance) masks over fields first.. (see paces_layout_t)."""
    kb = max([int(b) for b in widths[first:]] + [1])
    m = np.zeros((W, kb), dtype=np.uint64)
    for f in range(first, len(widths)):
        b = int(widths[f])
        for k in range(b):  # significance 2^k is bit (off + b - 1 - k) from the MSB
            pos = int(offsets[f]) + b - 1 - k
            m[pos // 64, k] |= np.uint64(1) << np.uint64(63 - pos % 64)
    return m, kb


class Term(C.Structure):
    _fields_ = [("kind", C.c_int32), ("order", C.c_int32), ("value", C.c_double)]


class Holstein(C.Structure):
    _fields_ = [("nchain", C.c_int32), ("lattice_l", C.c_int32), ("geo_dim", C.c_int32),
                ("periodic", C.c_int32), ("nterms", C.c_int32), ("terms", Term * MAX_TERMS),
                ("eps_sys", C.c_double), ("hbar_omega", C.c_double), ("delta_eps", C.c_double),
                ("d_sqrt_table", C.c_void_p), ("sqrt_table_len", C.c_int32),
                ("batched", C.c_int32), ("ntraj", C.c_int32), ("d_traj_L", C.c_void_p),
                ("d_traj_value", C.c_void_p)]


class Runs(C.Structure):
    _fields_ = [("ptr", C.c_void_p * MAX_RUNS), ("len", C.c_int64 * MAX_RUNS), ("K", C.c_int32)]


class CsrSet(C.Structure):
    _fields_ = [("nterms", C.c_int32), ("indptr", C.c_void_p * MAX_TERMS),
                ("indices", C.c_void_p * MAX_TERMS), ("data", C.c_void_p * MAX_TERMS)]


class CsrSet32(C.Structure):
    _fields_ = [("nterms", C.c_int32), ("indptr", C.c_void_p * MAX_TERMS),
                ("indices", C.c_void_p * MAX_TERMS), ("data", C.c_void_p * MAX_TERMS)]


CODE_TABLE = 64


class CsrCoded(C.Structure):
    _fields_ = [("nterms", C.c_int32), ("sqrt_len", C.c_int32),
                ("indptr", C.c_void_p * MAX_TERMS), ("indices", C.c_void_p * MAX_TERMS),
                ("code", C.c_void_p * MAX_TERMS), ("term_value", C.c_void_p),
                ("sqrt_table", C.c_void_p)]


SLOT_EMPTY = 0xFFFFFFFF
MAX_SLOTS = 8


class Slots(C.Structure):
    """paces_slots_t: every off-diagonal entry in one uint32 slot per (row, op)."""
    _fields_ = [("nterms", C.c_int32), ("nslots", C.c_int32), ("col_bits", C.c_int32),
                ("sqrt_len", C.c_int32), ("slot_off", C.c_int32 * (MAX_TERMS + 1)),
                ("vib", C.c_int32 * MAX_TERMS), ("stride", C.c_int32), ("slots", C.c_void_p),
                ("term_value", C.c_void_p), ("sqrt_table", C.c_void_p)]


class PriorImages(C.Structure):
    """paces_prior_images_t: the previous expansion's part of the image list."""
    _fields_ = [("nprev", C.c_int64), ("run_offsets", C.POINTER(C.c_int64)),
                ("cand_pos", C.c_void_p), ("img_src", C.c_void_p), ("img_nup", C.c_void_p)]


class BatchTaylor(C.Structure):
    _fields_ = [("c1", C.c_double), ("c1_plus_c2", C.c_double), ("ratio", C.c_double),
                ("done", C.c_int32), ("j_ind", C.c_int32), ("rn", C.c_double)]


class LaunchCfg(C.Structure):
    _fields_ = [("block", C.c_int32), ("rows_per_thread", C.c_int32)]


SERIES_MAX_TERMS = 64


class BatchSeries(C.Structure):
    _fields_ = [("R", C.c_int64), ("diag", C.c_void_p), ("nterms_h", C.c_int32),
                ("pad0", C.c_int32), ("indptr", C.c_void_p * MAX_TERMS),
                ("indices", C.c_void_p * MAX_TERMS), ("data", C.c_void_p * MAX_TERMS),
                ("v", C.c_void_p), ("res", C.c_void_p), ("x", C.c_void_p * 3),
                ("seg_off", C.c_void_p), ("ts", C.c_void_p), ("hdr", C.c_void_p),
                ("ws", C.c_void_p), ("tol", C.c_double), ("c", C.c_double * SERIES_MAX_TERMS),
                ("coded", C.c_int32), ("sqrt_len", C.c_int32), ("code", C.c_void_p * MAX_TERMS),
                ("term_value", C.c_void_p), ("sqrt_table", C.c_void_p),
                ("slots", C.c_void_p), ("nslots", C.c_int32), ("col_bits", C.c_int32),
                ("slot_off", C.c_int32 * (MAX_TERMS + 1)), ("slot_vib", C.c_int32 * MAX_TERMS),
                ("slot_stride", C.c_int32), ("xt", C.c_void_p * SERIES_MAX_TERMS),
                ("vmax", C.c_void_p)]


class TaylorState(C.Structure):
    _fields_ = [("max_term2", C.c_uint64), ("max_res2", C.c_uint64), ("c1", C.c_double),
                ("c1_plus_c2", C.c_double), ("ratio", C.c_double), ("done", C.c_int32),
                ("j_ind", C.c_int32), ("rn", C.c_double), ("pad", C.c_int32 * 2)]


_P, _I64, _I32, _D, _SZ, _F = (C.c_void_p, C.c_int64, C.c_int32, C.c_double, C.c_size_t,
                               C.c_float)

_SIGNATURES = {
    "paces_strerror": (C.c_char_p, [C.c_int]),
    "paces_version": (C.c_int, []),
    "paces_scan_workspace_bytes": (_SZ, [_I64]),
    "paces_exclusive_scan_u8": (C.c_int, [_P, _I64, _P, _P, _SZ, _P]),
    "paces_exclusive_scan_i32": (C.c_int, [_P, _I64, _P, _P, _SZ, _P]),
    "paces_pack_rows": (C.c_int, [_P, _I64, C.POINTER(Layout), _P, _P]),
    "paces_get_field": (C.c_int, [_P, _I64, C.POINTER(Layout), _I32, _P, _P]),
    "paces_add_field": (C.c_int, [_P, _I64, C.POINTER(Layout), _I32, _I32, _P]),
    "paces_lower_bound_workspace_bytes": (_SZ, [_I64]),
    "paces_lower_bound": (C.c_int, [_P, _I64, _P, _I64, _I32, _I32, _P, _P, _P, _P, _SZ, _P]),
    "paces_merge_unique_workspace_bytes": (_SZ, [_I64, _I32, _I32]),
    "paces_merge_unique": (C.c_int, [C.POINTER(Runs), _I32, _P, _P, _P, _P, _P, _SZ, _P]),
    "paces_merge_unique_coded": (C.c_int, [C.POINTER(Runs), _I32, _P, _P, _P, _P, _P, _P, _P,
                                           _SZ, _P]),
    "paces_holstein_inverse_skip_codes": (C.c_int, [C.POINTER(Holstein), _P, _I32]),
    "paces_merge_unique_diag": (C.c_int, [C.POINTER(Runs), _I32, _P, _P, _P, _P,
                                          C.POINTER(Layout), C.POINTER(Holstein), _P, _P, _P,
                                          _SZ, _P]),
    "paces_select_workspace_bytes": (_SZ, [_I64]),
    "paces_select_topk": (C.c_int, [_P, _P, _I64, _I32, _I64, _P, _P, _P, _SZ, _P]),
    "paces_holstein_num_runs": (C.c_int, [C.POINTER(Holstein)]),
    "paces_holstein_images_workspace_bytes": (_SZ, [_I64, C.POINTER(Holstein)]),
    "paces_holstein_images_count_skip": (C.c_int, [_P, _I64, _P, C.POINTER(Layout),
                                                   C.POINTER(Holstein), _P, _P, _SZ, _P]),
    "paces_holstein_images_emit_skip": (C.c_int, [_P, _I64, _P, C.POINTER(Layout),
                                                  C.POINTER(Holstein), _P, _P, _P, _P, _SZ, _P]),
    "paces_holstein_images_count": (C.c_int, [_P, _I64, C.POINTER(Layout), C.POINTER(Holstein),
                                              _P, _P, _SZ, _P]),
    "paces_holstein_images_emit": (C.c_int, [_P, _I64, C.POINTER(Layout), C.POINTER(Holstein),
                                             _P, _P, _P, _P, _SZ, _P]),
    "paces_holstein_pairs_workspace_bytes": (_SZ, [_I64, C.POINTER(Holstein)]),
    "paces_holstein_pairs_count": (C.c_int, [_I64, _I64, C.POINTER(C.c_int64), _P, _P, _P,
                                             C.POINTER(Holstein), _P, _P, _SZ, _P]),
    "paces_holstein_pairs_fill": (C.c_int, [_I64, _I64, C.POINTER(C.c_int64), _P, _P, _P, _P,
                                            _P, C.POINTER(Layout), C.POINTER(Holstein),
                                            C.POINTER(_P), C.POINTER(_P), C.POINTER(_P), _P, _SZ,
                                            _P]),
    "paces_holstein_pairs_fill32": (C.c_int, [_I64, _I64, C.POINTER(C.c_int64), _P, _P, _P, _P,
                                              _P, C.POINTER(Layout), C.POINTER(Holstein),
                                              C.POINTER(_P), C.POINTER(_P), C.POINTER(_P), _P,
                                              _SZ, _P]),
    "paces_holstein_pairs_fill_coded": (C.c_int, [_I64, _I64, C.POINTER(C.c_int64), _P, _P, _P,
                                                  _P, C.POINTER(Holstein), C.POINTER(_P),
                                                  C.POINTER(_P), C.POINTER(_P), _P, _SZ, _P]),
    "paces_holstein_pairs_count_split": (C.c_int, [_I64, _I64, C.POINTER(C.c_int64), _P, _P,
                                                   _P, C.POINTER(PriorImages),
                                                   C.POINTER(Holstein), _P, _P, _SZ, _P]),
    "paces_holstein_pairs_fill_split": (C.c_int, [_I64, _I64, C.POINTER(C.c_int64), _P, _P, _P,
                                                  _P, C.POINTER(PriorImages), _P,
                                                  C.POINTER(Layout), C.POINTER(Holstein),
                                                  C.POINTER(_P), C.POINTER(_P), C.POINTER(_P),
                                                  C.POINTER(_P), C.POINTER(_P), _P, _SZ, _P]),
    "paces_holstein_slot_layout": (C.c_int, [C.POINTER(Holstein), _I64, C.POINTER(Slots)]),
    "paces_holstein_pairs_slots": (C.c_int, [_I64, _I64, C.POINTER(C.c_int64), _P, _P, _P, _P,
                                             C.POINTER(PriorImages), C.POINTER(Holstein), _P,
                                             _P]),
    "paces_holstein_slots_place_explicit": (C.c_int, [_I64, C.POINTER(C.c_int64), _P, _P, _P,
                                                      _P, _P, C.POINTER(Holstein), _P, _P]),
    "paces_slots_mark": (C.c_int, [_P, _I64, _I32, _I32, _I64, _P, _I32, _I32, _I32, _P, _P,
                                   _P]),
    "paces_flag_indices_workspace_bytes": (_SZ, [_I64]),
    "paces_flag_indices": (C.c_int, [_P, _I64, _I32, _P, _P, _P, _SZ, _P]),
    "paces_gather_rows": (C.c_int, [_P, _I32, _P, _I64, _P, _P]),
    "paces_mark_hits": (C.c_int, [_P, _P, _I64, _P, _I32, _P, _P]),
    "paces_move_i32": (C.c_int, [_P, _P, _P, _I64, _P, _P]),
    "paces_fill_u8_at": (C.c_int, [_P, _P, _I64, _I32, _P]),
    "paces_warm_positions_workspace_bytes": (_SZ, [_I64]),
    "paces_warm_positions": (C.c_int, [_P, _I64, _P, _I64, _P, _P, _P, _P, _SZ, _P]),
    "paces_warm_emit": (C.c_int, [_P, _I32, _P, _P, _I32, _I64, _P, _I32, _P, _P, _I64, _I32,
                                  _I32, _P, _P, _P, _P, _P, _I64, _P, _P]),
    "paces_warm_emit_new": (C.c_int, [_P, _I32, _I64, _P, _P, _I32, _P, _I32, _P, _P, _P, _P,
                                      _P, _I64, _P, _P]),
    "paces_holstein_csr_workspace_bytes": (_SZ, [_I64, C.POINTER(Holstein)]),
    "paces_holstein_csr_count": (C.c_int, [_P, _I64, _P, C.POINTER(Layout), C.POINTER(Holstein),
                                           _P, _P, _SZ, _P]),
    "paces_holstein_csr_fill": (C.c_int, [_P, _I64, C.POINTER(Layout), C.POINTER(Holstein),
                                          C.POINTER(_P), C.POINTER(_P), C.POINTER(_P), _P, _SZ,
                                          _P]),
    "paces_holstein_diag": (C.c_int, [_P, _I64, C.POINTER(Layout), C.POINTER(Holstein), _P, _P]),
    "paces_holstein_diag32": (C.c_int, [_P, _I64, C.POINTER(Layout), C.POINTER(Holstein), _P,
                                        _P]),
    "paces_holstein_observables_workspace_bytes": (_SZ, [_I64, C.POINTER(Holstein)]),
    "paces_holstein_observables": (C.c_int, [_P, _P, _I64, C.POINTER(Layout),
                                             C.POINTER(Holstein), _P, _P, _SZ, _P]),
    "paces_reassign_workspace_bytes": (_SZ, [_I64]),
    "paces_reassign": (C.c_int, [_P, _I64, _P, _P, _I64, _I32, _P, _P, _SZ, _P]),
    "paces_reassign_noalloc": (C.c_int, [_P, _I64, _P, _P, _I64, _I32, _P, _P]),
    "paces_spmv": (C.c_int, [_I64, _P, C.POINTER(CsrSet), _P, _P, _P]),
    "paces_taylor_workspace_bytes": (_SZ, [_I64]),
    "paces_taylor_begin": (C.c_int, [_I64, _P, _P, _P, _P, _SZ, _P]),
    "paces_taylor_term": (C.c_int, [_I64, _P, C.POINTER(CsrSet), _P, _P, _P, _D, _I32, _D, _P,
                                    _P, _SZ, _P]),
    "paces_convert_f64_to_f32": (C.c_int, [_I64, _P, _P, _P]),
    "paces_taylor_begin_mixed": (C.c_int, [_I64, _P, _P, _P, _P, _P, _SZ, _P]),
    "paces_taylor_term_mixed": (C.c_int, [_I64, _P, C.POINTER(CsrSet32), _P, _P, _P, C.c_float,
                                          _I32, _D, _P, _P, _SZ, _P]),
    "paces_batch_segments": (C.c_int, [_P, _I64, _I32, _I32, _P, _P]),
    "paces_segment_vdot": (C.c_int, [_P, _P, _P, _I32, _P, _P]),
    "paces_pair_workspace_bytes": (_SZ, []),
    "paces_sum_f64": (C.c_int, [_P, _I64, _P, _P, _SZ, _P]),
    "paces_shift_f64": (C.c_int, [_P, _I64, _D, _P, _P]),
    "paces_csr_shifted_onenorm": (C.c_int, [_I64, _P, _D, C.POINTER(CsrSet), _P, _P, _SZ, _P]),
    "paces_pair_abs_sums": (C.c_int, [_P, _I64, _P, _P, _SZ, _P]),
    "paces_pair_dots": (C.c_int, [_P, _P, _I64, _P, _P, _SZ, _P]),
    "paces_pair_signs": (C.c_int, [_P, _I64, _P, _P]),
    "paces_pair_rowmax": (C.c_int, [_P, _I64, _P, _P]),
    "paces_argmax_after": (C.c_int, [_P, _I64, _D, _I64, _P, _P, _SZ, _P]),
    "paces_pair_random_column": (C.c_int, [_P, _I64, _I32, C.c_uint64, _D, _P]),
    "paces_pair_fill": (C.c_int, [_P, _I64, _D, _D, _I64, _I64, _P]),
    "paces_vacuum_sum": (C.c_int, [_P, _I32, _I32, _P, _P, _I32, _P, _P]),
    "paces_batch_select_workspace_bytes": (_SZ, [_I64, _I32]),
    "paces_batch_select_topk": (C.c_int, [_P, _P, _I64, _I32, _P, _I32, _P, _P, _P, _P, _SZ, _P]),
    "paces_holstein_observables_batched_workspace_bytes": (_SZ, [_I64, _I32, _I32]),
    "paces_holstein_observables_batched": (C.c_int, [_P, _P, _I64, _P, C.POINTER(Layout),
                                                     C.POINTER(Holstein), _P, _P, _SZ, _P]),
    "paces_batch_taylor_workspace_bytes": (_SZ, [_I64, _I32]),
    "paces_batch_taylor_begin": (C.c_int, [_I64, _P, _P, _P, _I32, _P, _I32, _P, _P, _SZ, _P]),
    "paces_batch_taylor_term": (C.c_int, [_I64, _P, C.POINTER(CsrSet), C.POINTER(CsrSet32), _P,
                                          _P, _P, _D, _I32, _D, _P, _I32, _P, _P, _P, _SZ, _P]),
    "paces_reduce_workspace_bytes": (_SZ, [_I64]),
    "paces_norm2": (C.c_int, [_P, _I64, _P, _P, _SZ, _P]),
    "paces_vdot": (C.c_int, [_P, _P, _I64, _P, _P, _SZ, _P]),
    "paces_axpby": (C.c_int, [_I64, _D, _D, _P, _D, _D, _P, _P]),
    "paces_taylor_term_tuned": (C.c_int, [_I64, _P, C.POINTER(CsrSet), _P, _P, _P, _D, _I32,
                                          _D, _P, _P, _SZ, C.POINTER(LaunchCfg), _P]),
    "paces_taylor_finish_mixed": (C.c_int, [_I64, _P, _P, _P, _P]),
    "paces_taylor_term_mixed_tuned": (C.c_int, [_I64, _P, C.POINTER(CsrSet32), _P, _P, _P, _F,
                                                _I32, _D, _P, _P, _SZ, C.POINTER(LaunchCfg),
                                                _P]),
    "paces_batch_taylor_begin_tuned": (C.c_int, [_I64, _P, _P, _P, _I32, _P, _I32, _P, _P, _SZ,
                                                 C.POINTER(LaunchCfg), _P]),
    "paces_batch_taylor_term_tuned": (C.c_int, [_I64, _P, C.POINTER(CsrSet),
                                                C.POINTER(CsrSet32), _P, _P, _P, _D, _I32, _D,
                                                _P, _I32, _P, _P, _P, _SZ, C.POINTER(LaunchCfg),
                                                _P]),
    "paces_batch_taylor_term_coded": (C.c_int, [_I64, _P, _I32, C.POINTER(CsrCoded), _P, _P,
                                                _P, _D, _I32, _D, _P, _I32, _P, _P, _P, _SZ,
                                                C.POINTER(LaunchCfg), _P]),
    "paces_batch_select_flags": (C.c_int, [_P, _P, _I64, _P, _I32, _P, _P, _P, _P, _P, _SZ,
                                           _P]),
    "paces_batch_select_topk_keys": (C.c_int, [_P, _P, _P, _I64, _I32, _P, _I32, _P, _P, _P, _P,
                                               _SZ, _P]),
    "paces_holstein_observables_batched_finish": (C.c_int, [_P, _P, _I64, _P, C.POINTER(Layout),
                                                            C.POINTER(Holstein), _P, _I32, _I32,
                                                            _I32, _P, _P, _P, _P, _P, _SZ, _P]),
    "paces_batch_taylor_term_slots": (C.c_int, [_I64, _P, _I32, C.POINTER(Slots), _P, _P,
                                                _P, _D, _I32, _D, _P, _I32, _P, _P, _P, _SZ,
                                                C.POINTER(LaunchCfg), _P]),
    "paces_batch_taylor_finish": (C.c_int, [_I64, _P, _I32, _I32, _P, _P, _I32, _P, _P, _SZ,
                                            C.POINTER(LaunchCfg), _P]),
    "paces_rows_varying_bits": (C.c_int, [_P, _I64, _I32, _P, _P]),
    "paces_sort_rows_workspace_bytes": (_SZ, [_I64, _I32]),
    "paces_sort_rows": (C.c_int, [_P, _I64, _I32, _P, _I32, _P, _P, _P, _P, _SZ, _P]),
    "paces_batch_series_desc_bytes": (_SZ, []),
    "paces_batch_series_launch": (C.c_int, [_P, _I64, _I32, _I32, _I32, _I32, _I32,
                                            C.POINTER(LaunchCfg), _P]),
    "paces_mlp_weights_bf16": (C.c_int, [_P, _P, _P]),
    "paces_mlp_train_workspace_bytes": (_SZ, [_I64]),
    "paces_mlp_grad": (C.c_int, [_P, _P, _P, _I64, _F, _P, _P, _P, _P, _P, _SZ, _P]),
    "paces_mlp_adam": (C.c_int, [_P, _P, _P, _P, _I32, _F, _F, _F, _F, _P, _P]),
    "paces_mlp_infer": (C.c_int, [_P, _I64, _P, _P, _P, _P, _P]),
    "paces_mlp_infer_tiled": (C.c_int, [_P, _I64, _P, _P, _P, _P, _I32, _P]),
    "paces_holstein_features_workspace_bytes": (_SZ, [_I64]),
    "paces_holstein_features": (C.c_int, [_P, _I64, _P, _P, _I64, _P, _P, _P, _P, _P, _P, _P,
                                          _SZ, _P]),
    "paces_holstein_features_added_workspace_bytes": (_SZ, [_I64]),
    "paces_holstein_features_added": (C.c_int, [_P, _I64, _P, _I64, _P, _P, _P, _P, _P, _P, _P,
                                                _P, _P, _SZ, _P]),
    "paces_prune_plan": (C.c_int, [_P, _I64, _F, _P, _I64, _I32, C.c_uint64, _P, _P]),
    "paces_prune_plan_rows": (C.c_int, [_P, _P, _I64, _P, _I64, _F, _I32, C.c_uint64, _P, _P]),
    "paces_scatter_u8": (C.c_int, [_P, _I64, _P, _P, _P]),
    "paces_norm2_masked": (C.c_int, [_P, _P, _I64, _P, _P, _SZ, _P]),
    "paces_mark_indices": (C.c_int, [_P, _I64, _P, _P, _P]),
    "paces_compact_workspace_bytes": (_SZ, [_I64]),
    "paces_compact_kept": (C.c_int, [_P, _I64, _P, _I32, _P, _P, _P, _I64, _P, _P, _P, _P, _P,
                                     _P, _P, _SZ, _P]),
}


def declared_symbols():
    return sorted(_SIGNATURES)


def load(path: str = LIB_PATH):
    """Load the library (once) and bind every declared entry point."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise PacesError(f"{path} is missing: build it with `python -m paces_amd.build`")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = load().paces_strerror(rc).decode()
        raise PacesError(f"{what}: {msg} (status {rc})")


def ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def stream():
    """The current HIP stream of the current device, as a void* for the C ABI (torch's raw
    getters: the same stream as torch.cuda.current_stream(), without building a Stream object,
    which costs about 10 us per call)."""
    return C.c_void_p(torch._C._cuda_getCurrentRawStream(torch._C._cuda_getDevice()))


class Workspace:
    """One growing device byte buffer reused by consecutive (stream-ordered) library calls."""

    def __init__(self, device):
        self.device = device
        self.buf = torch.empty(0, dtype=torch.uint8, device=device)

    def get(self, nbytes: int):
        nbytes = max(int(nbytes), 256)
        if self.buf.numel() < nbytes:
            self.buf = torch.empty(int(nbytes * 1.25) + 4096, dtype=torch.uint8,
                                   device=self.device)
        return C.c_void_p(self.buf.data_ptr()), C.c_size_t(self.buf.numel())


def rows_to_numpy(t):
    """Device int64 rows (bit patterns of uint64 words) -> host uint64 ndarray."""
    return t.detach().cpu().numpy().view(np.uint64)


def rows_from_numpy(a, device):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.uint64))
    return torch.from_numpy(a.view(np.int64).copy()).to(device)
