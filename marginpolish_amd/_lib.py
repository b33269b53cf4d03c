"""ctypes binding of libmp_pairhmm.so (the C ABI in include/mp_pairhmm.h).

The HIP library is the only implementation of the path: when it is missing
or no gfx950 device is present, every call raises; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmp_pairhmm.so")

MP_OK = 0
STATUS = {0: "MP_OK", 1: "MP_EINVAL", 2: "MP_EPARAMS", 3: "MP_EANCHORS", 4: "MP_EHIP", 5: "MP_ENOMEM",
          6: "MP_EUNSUPPORTED", 7: "MP_ENODEVICE"}

LIST_MATCH, LIST_GAPX, LIST_GAPY = 0, 1, 2
MODE_POSTERIORS, MODE_MATCH_ONLY, MODE_FORWARD_ONLY, MODE_EXPECTATIONS = 0, 1, 2, 3
BATCH_CROP_REFERENCE = 1
BATCH_FP64 = 2
BATCH_NO_GRAPH = 4
BATCH_EXACT_CHAIN = 8
BATCH_REFERENCE_ARITHMETIC = 16


class MPError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (STATUS.get(status, status), msg))
        self.status = status


class BandParams(C.Structure):
    _fields_ = [("threshold", C.c_double), ("minDiagsBetweenTraceBack", C.c_int64),
                ("traceBackDiagonals", C.c_int64), ("diagonalExpansion", C.c_int64),
                ("splitMatrixBiggerThanThis", C.c_int64), ("dynamicAnchorExpansion", C.c_int32)]


class Pair(C.Structure):
    _fields_ = [("sX", C.POINTER(C.c_uint16)), ("lX", C.c_int64), ("sY", C.POINTER(C.c_uint16)),
                ("lY", C.c_int64), ("anchors", C.POINTER(C.c_int64)), ("nAnchors", C.c_int64),
                ("strand", C.c_int32), ("raggedLeft", C.c_int32), ("raggedRight", C.c_int32)]


# mp_pair as a numpy record (pointers as uint64), for filling pair arrays column by column
PAIR_DTYPE = np.dtype([("sX", "<u8"), ("lX", "<i8"), ("sY", "<u8"), ("lY", "<i8"), ("anchors", "<u8"),
                       ("nAnchors", "<i8"), ("strand", "<i4"), ("raggedLeft", "<i4"), ("raggedRight", "<i4")],
                      align=True)
assert PAIR_DTYPE.itemsize == C.sizeof(Pair) and all(
    PAIR_DTYPE.fields[f][1] == getattr(Pair, f).offset for f, _ in Pair._fields_)


class AlignedPair(C.Structure):
    _fields_ = [("weight", C.c_int64), ("x", C.c_int32), ("y", C.c_int32)]


ALIGNED_PAIR_DTYPE = [("weight", "<i8"), ("x", "<i4"), ("y", "<i4")]


class RleString(C.Structure):
    _fields_ = [("bases", C.c_void_p), ("repeatCounts", C.c_void_p), ("length", C.c_int64)]


RLE_STRING_DTYPE = np.dtype([("bases", "<u8"), ("repeatCounts", "<u8"), ("length", "<i8")], align=True)
assert RLE_STRING_DTYPE.itemsize == C.sizeof(RleString)


class PoaSizes(C.Structure):
    _fields_ = [(f, C.c_int64) for f in ("nodes", "maxRepeatCount", "inserts", "insertSymbols", "deletes",
                                         "nodeObservations", "insertObservations", "deleteObservations")]


OBSERVATION_DTYPE = np.dtype([("readNo", "<i8"), ("offset", "<i8"), ("weight", "<f8")])

EXPORTS = [
    "mp_last_error", "mp_version", "mp_model_load_json", "mp_model_default_nucleotide", "mp_model_band_params",
    "mp_model_tables", "mp_model_repeat_log10", "mp_model_max_repeat_count", "mp_model_free", "mp_engine_create",
    "mp_engine_free", "mp_engine_cu_count", "mp_pairhmm_posteriors", "mp_batch_create", "mp_batch_run",
    "mp_batch_run_plain", "mp_batch_kernel_ms", "mp_batch_cells", "mp_batch_npairs", "mp_batch_fetch_counts",
    "mp_batch_fetch_list", "mp_batch_free", "mp_band_construct", "mp_split_points", "mp_traceback_windows",
    "mp_batch_fetch_expectations", "mp_pairhmm_forward_log_prob", "mp_pairhmm_expectations",
    "mp_model_set_transitions", "mp_batch_set_model", "mp_model_set_transitions5", "mp_model_transitions5",
    "mp_model_num_states", "mp_model_set_emissions", "mp_model_set_repeat_log10", "mp_batch_fetch_emission_counts",
    "mp_batch_autotune", "mp_poa_reference_graph", "mp_poa_free", "mp_poa_augment", "mp_batch_poa_augment",
    "mp_poa_get_sizes", "mp_poa_get_nodes", "mp_poa_get_inserts", "mp_poa_get_deletes", "mp_poa_get_observations",
    "mp_bubble_allele_supports", "mp_remove_overlaps", "mp_merge_chunks", "mp_poa_estimate_repeat_counts",
    "mp_poa_anchor_alignments", "mp_poa_consensus", "mp_poa_scores", "mp_poas_estimate_repeat_counts",
    "mp_batch_list_ms", "mp_batch_fetch_total_log_prob", "mp_batch_stats", "mp_batch_rerun_stats",
    "mp_batch_prep_ms", "mp_pairhmm_forward_log_prob_noalloc",
]

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MPError(7, "libmp_pairhmm.so is not built (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    i64, i32, dbl = C.c_int64, C.c_int32, C.c_double
    P = C.POINTER
    sig = {
        "mp_last_error": (C.c_char_p, []),
        "mp_version": (C.c_char_p, []),
        "mp_model_load_json": (C.c_int, [C.c_char_p, P(vp)]),
        "mp_model_default_nucleotide": (C.c_int, [P(vp)]),
        "mp_model_band_params": (C.c_int, [vp, P(BandParams)]),
        "mp_model_tables": (C.c_int, [vp, i32, P(dbl), P(dbl), P(dbl), P(dbl), P(i32)]),
        "mp_model_repeat_log10": (dbl, [vp, i32, i32, i32, i32]),
        "mp_model_max_repeat_count": (i32, [vp]),
        "mp_model_free": (None, [vp]),
        "mp_engine_create": (C.c_int, [i32, P(vp)]),
        "mp_engine_free": (None, [vp]),
        "mp_engine_cu_count": (i32, [vp]),
        "mp_pairhmm_posteriors": (C.c_int, [vp, vp, P(Pair), P(BandParams), i32, P(AlignedPair), i64, P(i64),
                                            P(AlignedPair), i64, P(i64), P(AlignedPair), i64, P(i64), P(dbl)]),
        "mp_batch_create": (C.c_int, [vp, vp, P(Pair), i64, P(BandParams), i32, i32, P(vp)]),
        "mp_batch_run": (C.c_int, [vp]),
        "mp_batch_run_plain": (C.c_int, [vp]),
        "mp_batch_kernel_ms": (dbl, [vp]),
        "mp_batch_list_ms": (dbl, [vp]),
        "mp_batch_fetch_total_log_prob": (C.c_int, [vp, P(dbl)]),
        "mp_batch_stats": (C.c_int, [vp, P(i64), P(i64), P(i64)]),
        "mp_batch_rerun_stats": (C.c_int, [vp, P(i32), P(i64), P(i64)]),
        "mp_batch_prep_ms": (C.c_int, [vp, P(dbl), i32]),
        "mp_batch_cells": (i64, [vp]),
        "mp_batch_npairs": (i64, [vp]),
        "mp_batch_fetch_counts": (C.c_int, [vp, P(i64), P(dbl)]),
        "mp_batch_fetch_list": (C.c_int, [vp, i32, P(AlignedPair)]),
        "mp_batch_free": (None, [vp]),
        "mp_band_construct": (C.c_int, [P(i64), i64, i64, i64, i64, i32, P(i32), P(i32)]),
        "mp_split_points": (i64, [P(i64), i64, i64, i64, i64, i32, i32, P(i64), i64]),
        "mp_traceback_windows": (i64, [P(i32), P(i32), i64, P(BandParams), P(i32), i64]),
        "mp_batch_fetch_expectations": (C.c_int, [vp, P(dbl), P(dbl)]),
        "mp_pairhmm_forward_log_prob": (C.c_int, [vp, vp, P(Pair), P(BandParams), P(dbl)]),
        "mp_pairhmm_forward_log_prob_noalloc": (C.c_int, [vp, vp, P(Pair), P(BandParams), P(dbl)]),
        "mp_pairhmm_expectations": (C.c_int, [vp, vp, P(Pair), P(BandParams), P(dbl), P(dbl)]),
        "mp_model_set_transitions": (C.c_int, [vp, P(dbl), i32]),
        "mp_model_set_transitions5": (C.c_int, [vp, P(dbl)]),
        "mp_model_transitions5": (C.c_int, [vp, P(dbl)]),
        "mp_model_num_states": (i32, [vp]),
        "mp_model_set_emissions": (C.c_int, [vp, P(dbl), P(dbl), P(dbl)]),
        "mp_model_set_repeat_log10": (C.c_int, [vp, P(dbl), i32]),
        "mp_batch_fetch_emission_counts": (C.c_int, [vp, P(dbl), P(dbl), P(dbl), P(dbl)]),
        "mp_batch_autotune": (C.c_int, [vp, i32, P(i32), P(i32), P(C.c_float)]),
        "mp_batch_set_model": (C.c_int, [vp, vp]),
        "mp_poa_reference_graph": (C.c_int, [P(RleString), i64, P(vp)]),
        "mp_poa_free": (None, [vp]),
        "mp_poa_augment": (C.c_int, [vp, P(RleString), i32, i64, vp, i64, vp, i64, vp, i64, i32]),
        "mp_batch_poa_augment": (C.c_int, [vp, P(vp), i64, vp, vp, vp, i32, i32]),
        "mp_poa_get_sizes": (C.c_int, [vp, P(PoaSizes)]),
        "mp_poa_get_nodes": (C.c_int, [vp, vp, vp, vp, vp]),
        "mp_poa_get_inserts": (C.c_int, [vp, vp, vp, vp, vp, vp, vp, vp]),
        "mp_poa_get_deletes": (C.c_int, [vp, vp, vp, vp, vp]),
        "mp_poa_get_observations": (C.c_int, [vp, i32, vp, vp]),
        "mp_bubble_allele_supports": (C.c_int, [vp, vp, P(BandParams), i64, vp, vp, vp, vp, vp, i32, i64, vp,
                                                P(i64)]),
        "mp_remove_overlaps": (C.c_int, [vp, P(BandParams), i64, vp, vp, vp, vp, vp, vp, vp, vp]),
        "mp_poa_estimate_repeat_counts": (C.c_int, [vp, vp, vp, vp, i64, vp, i32]),
        "mp_poa_anchor_alignments": (C.c_int, [vp, vp, i64, vp, i64, vp, vp, i64]),
        "mp_poa_consensus": (C.c_int, [vp, i32, dbl, vp, i64, P(i64), vp]),
        "mp_poa_scores": (C.c_int, [vp, vp]),
        "mp_poas_estimate_repeat_counts": (C.c_int, [P(vp), i64, vp, vp, vp, vp, P(vp), i32]),
        "mp_merge_chunks": (C.c_int, [vp, P(BandParams), i64, vp, vp, i64, C.c_char_p, vp, i64, P(i64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(status):
    if status != MP_OK:
        raise MPError(status, lib().mp_last_error().decode(errors="replace"))
