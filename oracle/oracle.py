"""ctypes wrapper of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
import this module: as the checker of the HIP path, and as the timed CPU
baseline.  The shipped library never imports it.

Model construction restates the parameter path of the reference:
  parser.c:154-408          polish params keys, RLE emission swap (391-401)
  stateMachine.c:206-266    hmm JSON (type, emissionsType, transitions, emissions)
  stateMachine.c:632-689    symmetric / asymmetric 3-state load
  stateMachine.c:454-470    reverse-complement of the emission tables
  stateMachine.c:467-489    default nucleotide machine constants
  parser.c:105-152          repeat-count table (only *_F keys)
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

LOGADD_POLY = 0
LOGADD_EXACT = 1
MAX_REPEAT = 51  # MAXIMUM_REPEAT_LENGTH, margin.h:123


class OrModel(C.Structure):
    _fields_ = [
        ("matchContinue", C.c_double), ("matchFromGapX", C.c_double), ("matchFromGapY", C.c_double),
        ("gapOpenX", C.c_double), ("gapOpenY", C.c_double), ("gapExtendX", C.c_double),
        ("gapExtendY", C.c_double), ("gapSwitchToX", C.c_double), ("gapSwitchToY", C.c_double),
        ("match", C.c_double * 16), ("gapX", C.c_double * 4), ("gapY", C.c_double * 4),
        ("rle", C.c_int), ("strand", C.c_int), ("maxRepeat", C.c_int),
        ("repeatLog10", C.POINTER(C.c_double)),
        ("nStates", C.c_int), ("trans5", C.c_double * 25),
    ]


class OrParams(C.Structure):
    _fields_ = [
        ("threshold", C.c_double), ("minDiagsBetweenTraceBack", C.c_int64),
        ("traceBackDiagonals", C.c_int64), ("diagonalExpansion", C.c_int64),
        ("splitMatrixBiggerThanThis", C.c_int64), ("dynamicAnchorExpansion", C.c_int),
    ]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def use_native():
    """Build the oracle for the host it runs on (-march=native, into
    _build_native/) and use that build from here on: the CPU baseline is timed
    with it.  Returns its path."""
    global _lib, LIB_PATH
    subprocess.check_call(["make", "-s", "-C", HERE, "native"])
    LIB_PATH = os.path.join(HERE, "_build_native", "liboracle.so")
    _lib = None
    lib()
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        i64p = C.POINTER(C.c_int64)
        u16p = C.POINTER(C.c_uint16)
        dp = C.POINTER(C.c_double)
        L.or_logadd.restype = C.c_double
        L.or_logadd.argtypes = [C.c_double, C.c_double, C.c_int]
        L.or_band.restype = C.c_int
        L.or_band.argtypes = [i64p, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int, i64p, i64p]
        L.or_split_points.restype = C.c_int64
        L.or_split_points.argtypes = [i64p, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int, C.c_int,
                                      i64p, C.c_int64]
        L.or_posteriors.restype = C.c_void_p
        L.or_posteriors.argtypes = [C.POINTER(OrModel), u16p, C.c_int64, u16p, C.c_int64, i64p, C.c_int64,
                                    C.POINTER(OrParams), C.c_int, C.c_int, C.c_int, C.c_int]
        L.or_expectations.restype = C.c_void_p
        L.or_expectations.argtypes = [C.POINTER(OrModel), u16p, C.c_int64, u16p, C.c_int64, i64p, C.c_int64,
                                      C.POINTER(OrParams), C.c_int, C.c_int, C.c_int]
        L.or_forward_prob.restype = C.c_double
        L.or_forward_prob.argtypes = [C.POINTER(OrModel), u16p, C.c_int64, u16p, C.c_int64, i64p, C.c_int64,
                                      C.POINTER(OrParams), C.c_int, C.c_int, C.c_int]
        L.or_full_dp.restype = C.c_int
        L.or_full_dp.argtypes = [C.POINTER(OrModel), u16p, C.c_int64, u16p, C.c_int64, C.c_int64, C.c_int,
                                 dp, dp, dp, dp]
        L.or_result_status.restype = C.c_int
        L.or_result_status.argtypes = [C.c_void_p]
        L.or_result_count.restype = C.c_int64
        L.or_result_count.argtypes = [C.c_void_p, C.c_int]
        L.or_result_copy.restype = None
        L.or_result_copy.argtypes = [C.c_void_p, C.c_int, i64p]
        L.or_result_total.restype = C.c_double
        L.or_result_total.argtypes = [C.c_void_p]
        L.or_result_likelihood.restype = C.c_double
        L.or_result_likelihood.argtypes = [C.c_void_p]
        L.or_result_cells.restype = C.c_int64
        L.or_result_cells.argtypes = [C.c_void_p]
        L.or_result_expectations.restype = None
        L.or_result_expectations.argtypes = [C.c_void_p, dp]
        L.or_result_emissions.restype = None
        L.or_result_emissions.argtypes = [C.c_void_p, dp]
        L.or_result_free.restype = None
        L.or_result_free.argtypes = [C.c_void_p]
        L.or_batch_cells.restype = C.c_int64
        L.or_batch_cells.argtypes = [C.POINTER(OrModel), C.c_int64, u16p, i64p, i64p, u16p, i64p, i64p,
                                     i64p, i64p, i64p, C.POINTER(C.c_int), C.POINTER(OrParams), C.c_int,
                                     C.c_int, i64p]
        _lib = L
    return _lib


# --------------------------------------------------------------------------
# Models
# --------------------------------------------------------------------------

def _log(v):
    return math.log(v) if v > 0 else -math.inf


class Model:
    """A resolved 3-state machine (log space) plus the RLE repeat table.
    With trans5 (a 5x5 natural-log matrix [from, to] over M, short X, short
    Y, long X, long Y) it is the five-state machine instead, and `trans` is
    not used by the recursions."""

    def __init__(self, trans, match, gapX, gapY, rle=False, strand=1, repeat=None, trans5=None):
        self.trans5 = None if trans5 is None else np.asarray(trans5, dtype=np.float64).reshape(5, 5).copy()
        self.trans = dict(trans)
        self.match = list(match)
        self.gapX = list(gapX)
        self.gapY = list(gapY)
        self.rle = bool(rle)
        self.strand = int(strand)
        self.repeat = (np.ascontiguousarray(repeat, dtype=np.float64) if repeat is not None
                       else np.zeros(5 * MAX_REPEAT * MAX_REPEAT))
        self._c = None

    def c(self):
        if self._c is None:
            m = OrModel()
            for k, v in self.trans.items():
                setattr(m, k, v)
            m.match = (C.c_double * 16)(*self.match)
            m.gapX = (C.c_double * 4)(*self.gapX)
            m.gapY = (C.c_double * 4)(*self.gapY)
            m.rle = int(self.rle)
            m.strand = self.strand
            m.maxRepeat = MAX_REPEAT
            m.repeatLog10 = self.repeat.ctypes.data_as(C.POINTER(C.c_double))
            if self.trans5 is not None:
                m.nStates = 5
                m.trans5 = (C.c_double * 25)(*[float(v) for v in self.trans5.reshape(25)])
            else:
                m.nStates = 3
            self._c = m
        return self._c

    def reverse_complement(self):
        """nucleotideEmissions_reverseComplement (stateMachine.c:454-470):
        M'[i][j] = M[3-i][3-j], G'[i] = G[3-i]."""
        m = [self.match[(3 - i) * 4 + (3 - j)] for i in range(4) for j in range(4)]
        return Model(self.trans, m, self.gapX[::-1], self.gapY[::-1], self.rle, self.strand, self.repeat,
                     self.trans5)


def default_nucleotide_model():
    """stateMachine3_construct + nucleotideEmissions_construct defaults
    (stateMachine.c:413-440, 611-620)."""
    tr = dict(matchContinue=-0.030064059121770816, matchFromGapX=-1.272871422049609,
              matchFromGapY=-1.272871422049609, gapOpenX=-4.21256642, gapOpenY=-4.21256642,
              gapExtendX=-0.3388262689231553, gapExtendY=-0.3388262689231553,
              gapSwitchToX=-4.910694825551255, gapSwitchToY=-4.910694825551255)
    M, TV, TS = -1.8917761142, -4.3459578861, -3.760242452
    match = [M, TV, TS, TV, TV, M, TV, TS, TS, TV, M, TV, TV, TS, TV, M]
    gap = [-1.3862943611] * 4
    return Model(tr, match, gap, gap)


def transitions_from_hmm(hmm):
    """stateMachine3_loadSymmetric / loadAsymmetric (stateMachine.c:632-668)."""
    t = hmm["transitions"]
    T = lambda a, b: t[a * 3 + b]
    M, X, Y = 0, 1, 2
    if hmm["type"] == 2:
        mfg = _log((T(X, M) + T(Y, M)) / 2.0)
        go = _log((T(M, X) + T(M, Y)) / 2.0)
        ge = _log((T(X, X) + T(Y, Y)) / 2.0)
        gs = _log((T(Y, X) + T(X, Y)) / 2.0)
        return dict(matchContinue=_log(T(M, M)), matchFromGapX=mfg, matchFromGapY=mfg, gapOpenX=go,
                    gapOpenY=go, gapExtendX=ge, gapExtendY=ge, gapSwitchToX=gs, gapSwitchToY=gs)
    if hmm["type"] == 3:
        return dict(matchContinue=_log(T(M, M)), matchFromGapX=_log(T(X, M)), matchFromGapY=_log(T(Y, M)),
                    gapOpenX=_log(T(M, X)), gapOpenY=_log(T(M, Y)), gapExtendX=_log(T(X, X)),
                    gapExtendY=_log(T(Y, Y)), gapSwitchToX=_log(T(Y, X)), gapSwitchToY=_log(T(X, Y)))
    raise ValueError("unrecognised state machine type %r" % hmm["type"])


def repeat_table(rcm):
    """repeatSubMatrix_jsonParse (parser.c:105-152): only *_F keys, row=base."""
    tab = np.zeros((5, MAX_REPEAT, MAX_REPEAT))
    for key, val in rcm.items():
        if key in ("baseLogRepeatCounts_AT", "baseLogRepeatCounts_GC"):
            continue
        if len(key) != 31 or key[28] not in "ACGT" or key[30] != "F":
            raise ValueError("unrecognised repeat matrix key %s" % key)
        tab["ACGT".index(key[28])] = np.asarray(val, dtype=np.float64).reshape(MAX_REPEAT, MAX_REPEAT)
    return tab.reshape(-1)


class PolishParams:
    """The slice of PolishParams the realignment path reads (margin.h:534-580)."""

    def __init__(self, path):
        with open(path) as fh:
            js = json.load(fh)
        pol = js["polish"] if "polish" in js else js
        hmm = pol["hmmForwardStrandReadGivenReference"]
        tr = transitions_from_hmm(hmm)
        em = [_log(v) for v in hmm["emissions"]]
        self.useRepeatCountsInAlignment = bool(pol.get("useRepeatCountsInAlignment", False))
        rep = repeat_table(pol["repeatCountSubstitutionMatrix"]) if "repeatCountSubstitutionMatrix" in pol else None
        fwd = Model(tr, em[:16], em[16:20], em[20:24])
        rev = fwd.reverse_complement()
        if self.useRepeatCountsInAlignment:
            fwd = Model(fwd.trans, fwd.match, fwd.gapX, fwd.gapY, True, 1, rep)
            rev = Model(rev.trans, rev.match, rev.gapX, rev.gapY, True, 0, rep)
        self.forward = fwd
        self.reverse = rev
        pa = pol["pairwiseAlignmentParameters"]
        self.p = Params(threshold=pa.get("threshold", 0.01),
                        minDiagsBetweenTraceBack=pa.get("minDiagsBetweenTraceBack", 1000),
                        traceBackDiagonals=pa.get("traceBackDiagonals", 40),
                        diagonalExpansion=pa.get("diagonalExpansion", 20),
                        splitMatrixBiggerThanThis=pa.get("splitMatrixBiggerThanThis", 3000 * 3000),
                        dynamicAnchorExpansion=pa.get("dynamicAnchorExpansion", False))
        self.repeat = rep


class Params:
    """pairwiseAlignmentBandingParameters_construct defaults (pairwiseAligner.c:999-1011)."""

    def __init__(self, threshold=0.01, minDiagsBetweenTraceBack=1000, traceBackDiagonals=40,
                 diagonalExpansion=20, splitMatrixBiggerThanThis=3000 * 3000, dynamicAnchorExpansion=False):
        self.threshold = threshold
        self.minDiagsBetweenTraceBack = int(minDiagsBetweenTraceBack)
        self.traceBackDiagonals = int(traceBackDiagonals)
        self.diagonalExpansion = int(diagonalExpansion)
        self.splitMatrixBiggerThanThis = int(splitMatrixBiggerThanThis)
        self.dynamicAnchorExpansion = bool(dynamicAnchorExpansion)

    def c(self):
        return OrParams(self.threshold, self.minDiagsBetweenTraceBack, self.traceBackDiagonals,
                        self.diagonalExpansion, self.splitMatrixBiggerThanThis, int(self.dynamicAnchorExpansion))


# --------------------------------------------------------------------------
# Symbols (stateMachine.c:25-146)
# --------------------------------------------------------------------------

_CODE = np.full(256, 4, dtype=np.uint16)
for _i, _ch in enumerate("ACGT"):
    _CODE[ord(_ch)] = _i
    _CODE[ord(_ch.lower())] = _i


def symbols(seq):
    return _CODE[np.frombuffer(seq.encode(), dtype=np.uint8)].astype(np.uint16)


def rle_encode(seq):
    """RleString: collapsed bases and run lengths."""
    if not seq:
        return "", np.zeros(0, dtype=np.int64)
    b = np.frombuffer(seq.encode(), dtype=np.uint8)
    starts = np.flatnonzero(np.r_[True, b[1:] != b[:-1]])
    runs = np.diff(np.r_[starts, len(b)])
    return bytes(b[starts]).decode(), runs.astype(np.int64)


def rle_symbols(rle_bases, runs, max_exclusive=MAX_REPEAT):
    """rleString_constructSymbolString with includeRepeatCounts."""
    base = symbols(rle_bases).astype(np.uint32)
    rl = np.minimum(np.asarray(runs, dtype=np.int64), max_exclusive - 1).astype(np.uint32)
    return ((rl << 8) | base).astype(np.uint16)


# --------------------------------------------------------------------------
# Calls
# --------------------------------------------------------------------------

def _u16(a):
    a = np.ascontiguousarray(a, dtype=np.uint16)
    if a.size == 0:
        a = np.zeros(1, dtype=np.uint16)[:0]
    return a


def _anchors(anchors):
    if anchors is None or len(anchors) == 0:
        return np.zeros((1, 3), dtype=np.int64), 0
    a = np.ascontiguousarray(np.asarray(anchors, dtype=np.int64).reshape(-1, 3))
    return a, a.shape[0]


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def logadd(x, y, mode=LOGADD_POLY):
    return lib().or_logadd(x, y, mode)


def band(anchors, lX, lY, expansion=0, dynamic=False):
    L = lib()
    a, n = _anchors(anchors)
    xl = np.zeros(lX + lY + 1, dtype=np.int64)
    xr = np.zeros(lX + lY + 1, dtype=np.int64)
    st = L.or_band(_ptr(a, C.c_int64), n, lX, lY, expansion, int(dynamic), _ptr(xl, C.c_int64), _ptr(xr, C.c_int64))
    if st != 0:
        raise ValueError("invalid band")
    return xl, xr


def split_points(anchors, lX, lY, max_area, ragged_left=False, ragged_right=False):
    L = lib()
    a, n = _anchors(anchors)
    cap = 64
    while True:
        out = np.zeros((cap, 4), dtype=np.int64)
        k = L.or_split_points(_ptr(a, C.c_int64), n, lX, lY, max_area, int(ragged_left), int(ragged_right),
                              _ptr(out, C.c_int64), cap)
        if k <= cap:
            return [tuple(int(v) for v in r) for r in out[:k]]
        cap = k


def _take(res):
    L = lib()
    out = []
    for li in range(3):
        n = L.or_result_count(res, li)
        arr = np.zeros((max(n, 1), 3), dtype=np.int64)
        if n:
            L.or_result_copy(res, li, _ptr(arr, C.c_int64))
        out.append(arr[:n])
    return out


def posteriors(model, sX, sY, anchors, params, ragged_left=False, ragged_right=False, with_indels=True,
               mode=LOGADD_EXACT):
    """getAlignedPairsWithIndelsUsingAnchors / getAlignedPairsUsingAnchors.

    Returns (matches, gapX, gapY, total_log_prob, cells), each list an
    int64 array of rows (weight, x, y) in the reference's output order."""
    L = lib()
    sx, sy = _u16(sX), _u16(sY)
    a, n = _anchors(anchors)
    pc = params.c()
    r = L.or_posteriors(C.byref(model.c()), _ptr(sx, C.c_uint16), len(sx), _ptr(sy, C.c_uint16), len(sy),
                        _ptr(a, C.c_int64), n, C.byref(pc), int(ragged_left), int(ragged_right),
                        int(with_indels), mode)
    try:
        if L.or_result_status(r) != 0:
            raise ValueError("oracle status %d" % L.or_result_status(r))
        m, gx, gy = _take(r)
        return m, gx, gy, L.or_result_total(r), L.or_result_cells(r)
    finally:
        L.or_result_free(r)


def expectations(model, sX, sY, anchors, params, ragged_left=False, ragged_right=False, mode=LOGADD_EXACT):
    """getExpectationsUsingAnchors: 9 transition expectations + likelihood."""
    L = lib()
    sx, sy = _u16(sX), _u16(sY)
    a, n = _anchors(anchors)
    pc = params.c()
    r = L.or_expectations(C.byref(model.c()), _ptr(sx, C.c_uint16), len(sx), _ptr(sy, C.c_uint16), len(sy),
                          _ptr(a, C.c_int64), n, C.byref(pc), int(ragged_left), int(ragged_right), mode)
    try:
        if L.or_result_status(r) != 0:
            raise ValueError("oracle status %d" % L.or_result_status(r))
        e = np.zeros(9)
        L.or_result_expectations(r, _ptr(e, C.c_double))
        return e, L.or_result_likelihood(r)
    finally:
        L.or_result_free(r)


def emission_counts(model, sX, sY, anchors, params, ragged_left=False, ragged_right=False, mode=LOGADD_EXACT):
    """Expected emission counts of getExpectationsUsingAnchors' diagonals (the
    reference disables them, pairwiseAligner.c:359-362): (match 4x4 [x, y],
    gapX 4, gapY 4, repeat 4x51x51 [base, underlying, observed]) on the
    forward-strand parameters."""
    L = lib()
    sx, sy = _u16(sX), _u16(sY)
    a, n = _anchors(anchors)
    pc = params.c()
    r = L.or_expectations(C.byref(model.c()), _ptr(sx, C.c_uint16), len(sx), _ptr(sy, C.c_uint16), len(sy),
                          _ptr(a, C.c_int64), n, C.byref(pc), int(ragged_left), int(ragged_right), mode)
    try:
        if L.or_result_status(r) != 0:
            raise ValueError("oracle status %d" % L.or_result_status(r))
        e = np.zeros(16 + 4 + 4 + 4 * 51 * 51)
        L.or_result_emissions(r, _ptr(e, C.c_double))
        return e[:16].reshape(4, 4), e[16:20], e[20:24], e[24:].reshape(4, 51, 51)
    finally:
        L.or_result_free(r)


def forward_prob(model, sX, sY, anchors, params, ragged_left=False, ragged_right=False, mode=LOGADD_EXACT):
    """computeForwardProbability."""
    L = lib()
    sx, sy = _u16(sX), _u16(sY)
    a, n = _anchors(anchors)
    pc = params.c()
    return L.or_forward_prob(C.byref(model.c()), _ptr(sx, C.c_uint16), len(sx), _ptr(sy, C.c_uint16), len(sy),
                             _ptr(a, C.c_int64), n, C.byref(pc), int(ragged_left), int(ragged_right), mode)


def full_dp(model, sX, sY, expansion=2, mode=LOGADD_POLY, want_post=False):
    """Full-matrix forward/backward as test_diagonalDPCalculations drives it."""
    L = lib()
    sx, sy = _u16(sX), _u16(sY)
    f = C.c_double()
    b = C.c_double()
    cuts = np.zeros(len(sx) + len(sy) + 1)
    post = np.zeros((len(sx) + 1) * (len(sy) + 1)) if want_post else None
    st = L.or_full_dp(C.byref(model.c()), _ptr(sx, C.c_uint16), len(sx), _ptr(sy, C.c_uint16), len(sy),
                      expansion, mode, C.byref(f), C.byref(b), _ptr(cuts, C.c_double),
                      _ptr(post, C.c_double) if want_post else None)
    if st:
        raise ValueError("oracle status %d" % st)
    return f.value, b.value, cuts, (post.reshape(len(sx) + 1, len(sy) + 1) if want_post else None)


def batch_cells(models, pairs, params, mode=LOGADD_POLY, nthreads=1):
    """Timed CPU baseline: posteriors for every pair (outputs counted only).

    pairs: list of (sX, sY, anchors, strand).  Returns (cells, counts[n,3])."""
    L = lib()
    n = len(pairs)
    offX = np.zeros(n, np.int64); lenX = np.zeros(n, np.int64)
    offY = np.zeros(n, np.int64); lenY = np.zeros(n, np.int64)
    aoff = np.zeros(n, np.int64); an = np.zeros(n, np.int64)
    strand = np.zeros(n, np.int32)
    xs, ys, anc = [], [], []
    ox = oy = oa = 0
    for i, (sx, sy, a, s) in enumerate(pairs):
        offX[i], lenX[i], offY[i], lenY[i] = ox, len(sx), oy, len(sy)
        ox += len(sx); oy += len(sy)
        a = np.asarray(a, dtype=np.int64).reshape(-1, 3)
        aoff[i], an[i] = oa, a.shape[0]
        oa += a.shape[0]
        xs.append(np.asarray(sx, np.uint16)); ys.append(np.asarray(sy, np.uint16)); anc.append(a)
        strand[i] = 1 if s else 0
    X = np.ascontiguousarray(np.concatenate(xs + [np.zeros(1, np.uint16)]))
    Y = np.ascontiguousarray(np.concatenate(ys + [np.zeros(1, np.uint16)]))
    A = np.ascontiguousarray(np.concatenate(anc + [np.zeros((1, 3), np.int64)]))
    ms = (OrModel * 2)(models[0].c(), models[1].c())
    counts = np.zeros((n, 3), np.int64)
    pc = params.c()
    cells = L.or_batch_cells(ms, n, _ptr(X, C.c_uint16), _ptr(offX, C.c_int64), _ptr(lenX, C.c_int64),
                             _ptr(Y, C.c_uint16), _ptr(offY, C.c_int64), _ptr(lenY, C.c_int64),
                             _ptr(A, C.c_int64), _ptr(aoff, C.c_int64), _ptr(an, C.c_int64),
                             strand.ctypes.data_as(C.POINTER(C.c_int)), C.byref(pc), mode, nthreads,
                             _ptr(counts, C.c_int64))
    return cells, counts
