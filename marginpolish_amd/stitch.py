"""Chunk stitching over the library (include/mp_stitch.h): removeOverlap
(polisher.c:1937-2006) for many boundaries in one batch and mergeContigChunks
(chunker.c:55-118)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import pairhmm
from ._lib import check, lib


def _strings(strs):
    bufs = [s.encode() if isinstance(s, str) else bytes(s) for s in strs]
    ptrs = (C.c_char_p * max(len(bufs), 1))(*bufs)
    lens = np.array([len(b) for b in bufs] or [0], dtype=np.int64)
    return bufs, ptrs, lens


def remove_overlaps(p, prefixes, suffixes, approx_overlaps, engine=None):
    """removeOverlap for every (prefix, suffix, approxOverlap): returns arrays
    (prefixCropEnd, suffixCropStart, overlapWeight)."""
    engine = engine or pairhmm.default_engine()
    n = len(prefixes)
    kp, pp, pl = _strings(prefixes)
    ks, ps, sl = _strings(suffixes)
    ao = np.ascontiguousarray(np.asarray(approx_overlaps, dtype=np.int64).reshape(-1))
    if ao.size != n or len(suffixes) != n:
        raise ValueError("one suffix and one overlap per prefix")
    ce, cs, w = (np.zeros(max(n, 1), np.int64) for _ in range(3))
    pc = p.c()
    check(lib().mp_remove_overlaps(engine.handle, C.byref(pc), n, C.cast(pp, C.c_void_p), pl.ctypes.data,
                                   C.cast(ps, C.c_void_p), sl.ctypes.data, ao.ctypes.data, ce.ctypes.data,
                                   cs.ctypes.data, w.ctypes.data))
    del kp, ks
    return ce[:n], cs[:n], w[:n]


def removeOverlap(prefixString, suffixString, approxOverlap, polishParams, engine=None):
    """polisher.c:1937-2006: (overlapWeight, prefixStringCropEnd, suffixStringCropStart)."""
    ce, cs, w = remove_overlaps(polishParams.p, [prefixString], [suffixString], [approxOverlap], engine)
    return int(w[0]), int(ce[0]), int(cs[0])


def mergeContigChunks(chunks, overlap, polishParams, missingChunkSpacer="", engine=None):
    """chunker.c:55-118: the polished chunk strings of one contig, merged."""
    engine = engine or pairhmm.default_engine()
    keep, ptrs, lens = _strings(chunks)
    total = C.c_int64()
    pc = polishParams.p.c()
    cap = int(lens.sum()) + 10 * len(chunks) + len(missingChunkSpacer) * len(chunks) + 1
    out = C.create_string_buffer(cap)
    check(lib().mp_merge_chunks(engine.handle, C.byref(pc), len(chunks), C.cast(ptrs, C.c_void_p),
                                lens.ctypes.data, int(overlap), missingChunkSpacer.encode(), out, cap,
                                C.byref(total)))
    del keep
    return out.raw[:total.value].decode()
