"""Reference-shaped Python interface of the MI355X pair-HMM engine.

Mirrors the entry points of impl/pairwiseAligner.c and the realign caller of
impl/polisher.c, with the same argument meaning:

  params_readParams(path)                        parser.c:503-528
  getAlignedPairsWithIndelsUsingAnchors(...)     pairwiseAligner.c:1105-1130
  getAlignedPairsUsingAnchors(...)               pairwiseAligner.c:1088-1103
  getAlignedPairs / getAlignedPairsWithIndels    pairwiseAligner.c:1132-1149
  computeForwardProbability(...)                 pairwiseAligner.c:813-865
  getExpectationsUsingAnchors / getExpectations  pairwiseAligner.c:1151-1164
  getAlignedPairsWithIndelsCroppingReference     polisher.c:605-654
  realign_all(...)                               the per-read loop of poa_realign
                                                 (polisher.c:669-700), batched

Lists are numpy structured arrays with fields (weight, x, y), in the
reference's output order.  Every call runs on the GPU through
libmp_pairhmm.so; errors raise MPError (the reference aborts instead).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import ALIGNED_PAIR_DTYPE, MPError, check, lib  # noqa: F401

MAX_REPEAT_EXCLUSIVE = 51  # MAXIMUM_REPEAT_LENGTH (margin.h:123)


@dataclass
class PairwiseAlignmentParameters:
    """pairwiseAligner.h:28-49 (fields on this path), defaults of
    pairwiseAlignmentBandingParameters_construct (pairwiseAligner.c:999-1011)."""
    threshold: float = 0.01
    minDiagsBetweenTraceBack: int = 1000
    traceBackDiagonals: int = 40
    diagonalExpansion: int = 20
    splitMatrixBiggerThanThis: int = 3000 * 3000
    dynamicAnchorExpansion: bool = False

    def c(self):
        return _lib.BandParams(self.threshold, self.minDiagsBetweenTraceBack, self.traceBackDiagonals,
                               self.diagonalExpansion, self.splitMatrixBiggerThanThis,
                               int(bool(self.dynamicAnchorExpansion)))


# Long-gap defaults of the five-state machine: open 0.0018 from M to each
# side, extension 0.9965 (mean long-gap length 1 / 0.0035 ~ 286 positions).
LONG_GAP_OPEN = 0.0018
LONG_GAP_EXTEND = 0.9965


def with_long_gaps(trans9, long_open=LONG_GAP_OPEN, long_extend=LONG_GAP_EXTEND):
    """5x5 natural-log transitions [from, to] over (M, short X, short Y,
    long X, long Y) from 3-state natural-log transitions (order of
    StateMachines.tables()['trans']): the short-gap rows and columns are the
    3-state ones, M opens each long gap with probability long_open (taken from
    M -> M), and a long gap extends with long_extend or returns to M."""
    t = [float(v) for v in trans9]
    mm = math.exp(t[0]) - 2.0 * long_open
    if not (mm > 0.0 and 0.0 <= long_open and 0.0 < long_extend < 1.0):
        raise ValueError("long-gap parameters leave no match-continue probability or are out of range")
    lg = lambda v: math.log(v) if v > 0.0 else -math.inf
    T = np.full((5, 5), -math.inf)
    T[0, 0], T[0, 1], T[0, 2], T[0, 3], T[0, 4] = math.log(mm), t[3], t[4], lg(long_open), lg(long_open)
    T[1, 0], T[1, 1], T[1, 2] = t[1], t[5], t[8]
    T[2, 0], T[2, 1], T[2, 2] = t[2], t[7], t[6]
    T[3, 0], T[3, 3] = math.log(1.0 - long_extend), math.log(long_extend)
    T[4, 0], T[4, 4] = math.log(1.0 - long_extend), math.log(long_extend)
    return T


class StateMachines:
    """Both strands' state machines and tables (an mp_model): the reference's
    3-state machine, or the five-state extension (M, short X, short Y,
    long X, long Y) made by set_transitions5 or a "type": 5 hmm."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def from_json(cls, path):
        h = C.c_void_p()
        check(lib().mp_model_load_json(path.encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def default_nucleotide(cls):
        """stateMachine3_constructNucleotide(threeState)."""
        h = C.c_void_p()
        check(lib().mp_model_default_nucleotide(C.byref(h)))
        return cls(h)

    @property
    def handle(self):
        return self._h

    def band_params(self):
        bp = _lib.BandParams()
        check(lib().mp_model_band_params(self._h, C.byref(bp)))
        return PairwiseAlignmentParameters(bp.threshold, bp.minDiagsBetweenTraceBack, bp.traceBackDiagonals,
                                           bp.diagonalExpansion, bp.splitMatrixBiggerThanThis,
                                           bool(bp.dynamicAnchorExpansion))

    def tables(self, strand):
        t = (C.c_double * 9)()
        m = (C.c_double * 16)()
        gx = (C.c_double * 4)()
        gy = (C.c_double * 4)()
        rle = C.c_int32()
        check(lib().mp_model_tables(self._h, int(strand), t, m, gx, gy, C.byref(rle)))
        return dict(trans=list(t), match=list(m), gapX=list(gx), gapY=list(gy), rle=bool(rle.value))

    def set_transitions(self, log_trans9, pin_end=False):
        """Replace both strands' natural-log transitions (order of
        tables()['trans']); pin_end keeps the end-state vector fixed."""
        t = (C.c_double * 9)(*[float(v) for v in log_trans9])
        check(lib().mp_model_set_transitions(self._h, t, int(bool(pin_end))))

    def set_emissions(self, log_match16, log_gap_x4, log_gap_y4):
        """Replace the forward strand's natural-log emissions (joint match
        [x * 4 + y], gap X, gap Y); the reverse strand takes their reverse
        complement."""
        f = lambda v, n: (C.c_double * n)(*[float(x) for x in np.asarray(v, dtype=np.float64).reshape(n)])
        check(lib().mp_model_set_emissions(self._h, f(log_match16, 16), f(log_gap_x4, 4), f(log_gap_y4, 4)))

    def set_repeat_log10(self, log10_repeat, enable=True):
        """Rows 0-3 of the repeat table (log10, [base][underlying][observed],
        4 x 51 x 51) and whether the match emission uses it."""
        if log10_repeat is None:
            check(lib().mp_model_set_repeat_log10(self._h, None, int(bool(enable))))
            return
        a = np.ascontiguousarray(np.asarray(log10_repeat, dtype=np.float64).reshape(4 * 51 * 51))
        check(lib().mp_model_set_repeat_log10(self._h, a.ctypes.data_as(C.POINTER(C.c_double)), int(bool(enable))))

    @property
    def num_states(self):
        return int(lib().mp_model_num_states(self._h))

    def set_transitions5(self, log_trans25):
        """Make both strands five-state with these natural-log transitions,
        a 5x5 matrix (or 25 values) [from, to] over (M, Xs, Ys, Xl, Yl)."""
        v = [float(x) for x in np.asarray(log_trans25, dtype=np.float64).reshape(25)]
        check(lib().mp_model_set_transitions5(self._h, (C.c_double * 25)(*v)))

    def transitions5(self):
        t = (C.c_double * 25)()
        check(lib().mp_model_transitions5(self._h, t))
        return np.array(list(t)).reshape(5, 5)

    def add_long_gaps(self, long_open=LONG_GAP_OPEN, long_extend=LONG_GAP_EXTEND):
        """Make the machine five-state: its 3-state transitions become the
        short-gap part and long gaps are added (with_long_gaps)."""
        self.set_transitions5(with_long_gaps(self.tables(1)["trans"], long_open, long_extend))

    def repeat_log10(self, base, strand, observed, underlying):
        return lib().mp_model_repeat_log10(self._h, base, strand, observed, underlying)

    def __del__(self):
        if getattr(self, "_h", None) and _lib._lib is not None:
            _lib._lib.mp_model_free(self._h)
            self._h = None


class PolishParams:
    """params_readParams for the realignment path: the read/reference state
    machines (forward + reverse strand) and the pairwise alignment params."""

    def __init__(self, path):
        self.machines = StateMachines.from_json(path)
        self.p = self.machines.band_params()
        self.useRepeatCountsInAlignment = self.machines.tables(1)["rle"]
        self.maximumRepeatLength = lib().mp_model_max_repeat_count(self.machines.handle)


def params_readParams(path):
    return PolishParams(path)


class Engine:
    """One gfx950 device and its stream."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().mp_engine_create(int(device), C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    @property
    def cu_count(self):
        return lib().mp_engine_cu_count(self._h)

    def close(self):
        if getattr(self, "_h", None):
            lib().mp_engine_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_engine = None


def default_engine():
    global _default_engine
    if _default_engine is None:
        _default_engine = Engine(0)
    return _default_engine


def _as_u16(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint16))


def _as_anchors(a):
    if a is None:
        return np.zeros((0, 3), dtype=np.int64)
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.int64).reshape(-1, 3))
    return arr


class PairArray:
    """Pairs laid out once as the library's mp_pair array (what a C caller
    hands to mp_batch_create): build it once and create batches from it
    without re-marshalling the Python tuples."""

    def __init__(self, pairs):
        n = len(pairs)
        arr = np.zeros(max(n, 1), dtype=_lib.PAIR_DTYPE)
        sxs = [_as_u16(pr[0]) for pr in pairs]
        sys_ = [_as_u16(pr[1]) for pr in pairs]
        ans = [_as_anchors(pr[2]) for pr in pairs]
        self._keep = (sxs, sys_, ans)
        if n:
            arr["sX"][:n] = [a.ctypes.data for a in sxs]
            arr["lX"][:n] = [a.size for a in sxs]
            arr["sY"][:n] = [a.ctypes.data for a in sys_]
            arr["lY"][:n] = [a.size for a in sys_]
            arr["anchors"][:n] = [a.ctypes.data for a in ans]
            arr["nAnchors"][:n] = [a.shape[0] for a in ans]
            arr["strand"][:n] = [int(pr[3]) for pr in pairs]
            arr["raggedLeft"][:n] = [int(bool(pr[4])) if len(pr) > 4 else 0 for pr in pairs]
            arr["raggedRight"][:n] = [int(bool(pr[5])) if len(pr) > 5 else 0 for pr in pairs]
        self.array = arr
        self.n = n

    def __len__(self):
        return self.n


class Batch:
    """A device-resident batch of read/reference pairs (mp_batch)."""

    def __init__(self, machines, pairs, p, mode=_lib.MODE_POSTERIORS, crop_reference=False, engine=None,
                 fp64=False, graph=True, exact_chain=False, reference_arithmetic=False):
        """pairs: iterable of (sX, sY, anchors, strand[, raggedLeft, raggedRight]),
        or a PairArray built from one.
        fp64: FP64 arithmetic on the per-cell-exponent kernel (the precision
        comparison path; 3-state machines run with the long gaps unreachable).
        graph: replay the batch's captured hipGraphs on every run (False:
        enqueue the launches directly).
        exact_chain: run a sub-problem's traceback windows on one forward
        chain (the reference's order) instead of as window jobs.
        reference_arithmetic: the reference's own FP64 log-space arithmetic
        (polynomial logAdd, its operation order and total cadence), so the
        weights are the reference's."""
        self.engine = engine or default_engine()
        # the mp_pair array (a numpy record array of the same layout)
        pa = pairs if isinstance(pairs, PairArray) else PairArray(pairs)
        n = len(pa)
        self._pairs = pa
        arr = pa.array.ctypes.data_as(C.POINTER(_lib.Pair))
        h = C.c_void_p()
        pc = p.c()
        check(lib().mp_batch_create(self.engine.handle, machines.handle, arr, n, C.byref(pc), int(mode),
                                    (_lib.BATCH_CROP_REFERENCE if crop_reference else 0) |
                                    (_lib.BATCH_FP64 if fp64 else 0) | (0 if graph else _lib.BATCH_NO_GRAPH) |
                                    (_lib.BATCH_EXACT_CHAIN if exact_chain else 0) |
                                    (_lib.BATCH_REFERENCE_ARITHMETIC if reference_arithmetic else 0),
                                    C.byref(h)))
        self._h = h
        self.npairs = n
        self.mode = mode

    @property
    def handle(self):
        return self._h

    def run(self):
        check(lib().mp_batch_run(self._h))
        return self

    def autotune(self, repeats=3):
        """Pick the fastest launch geometry for this batch (cached in the
        engine for batches of the same shape); returns (narrow workgroup
        threads, five-state waves per workgroup or 0, best kernel ms)."""
        t, w, ms = C.c_int32(), C.c_int32(), C.c_float()
        check(lib().mp_batch_autotune(self._h, int(repeats), C.byref(t), C.byref(w), C.byref(ms)))
        return t.value, w.value, ms.value

    def run_plain(self):
        check(lib().mp_batch_run_plain(self._h))
        return self

    @property
    def kernel_ms(self):
        return lib().mp_batch_kernel_ms(self._h)

    @property
    def list_ms(self):
        """Device time of the last run's list assembly (ms)."""
        return lib().mp_batch_list_ms(self._h)

    @property
    def cells(self):
        return lib().mp_batch_cells(self._h)

    def stats(self):
        """(device jobs, posterior-record capacity, records of the last run)."""
        j, c, r = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().mp_batch_stats(self._h, C.byref(j), C.byref(c), C.byref(r)))
        return j.value, c.value, r.value

    def prep_ms(self):
        """Host time (ms) of the batch's creation phases: checks and split
        points, symbols and jobs, bands and windows, window jobs, device
        order and record regions, slots, upload and device buffers, list
        buffers."""
        v = (C.c_double * 8)()
        check(lib().mp_batch_prep_ms(self._h, v, 8))
        return list(v)

    def rerun_stats(self):
        """The last run's record-region re-runs: (re-runs, jobs over their
        capacity, records beyond their regions)."""
        n, j, m = C.c_int32(), C.c_int64(), C.c_int64()
        check(lib().mp_batch_rerun_stats(self._h, C.byref(n), C.byref(j), C.byref(m)))
        return n.value, j.value, m.value

    def total_log_prob(self):
        """Sum of the pairs' log-likelihoods, accumulated in FP64 on the device."""
        v = C.c_double()
        check(lib().mp_batch_fetch_total_log_prob(self._h, C.byref(v)))
        return v.value

    def counts(self):
        c = np.zeros((max(self.npairs, 1), 3), dtype=np.int64)
        lp = np.zeros(max(self.npairs, 1), dtype=np.float64)
        check(lib().mp_batch_fetch_counts(self._h, c.ctypes.data_as(C.POINTER(C.c_int64)),
                                          lp.ctypes.data_as(C.POINTER(C.c_double))))
        return c[: self.npairs], lp[: self.npairs]

    def fetch_arrays(self):
        """The three lists of every pair at once, as the C ABI hands them over:
        (lists, starts, logProb), lists[l] the structured array of list l
        (matches, gapX, gapY) in pair order, pair i's entries
        lists[l][starts[l][i]:starts[l][i + 1]]."""
        counts, lp = self.counts()
        lists, starts = [], []
        for li in range(3):
            tot = int(counts[:, li].sum())
            buf = np.empty(max(tot, 1), dtype=ALIGNED_PAIR_DTYPE)
            if tot:
                check(lib().mp_batch_fetch_list(self._h, li, buf.ctypes.data_as(C.POINTER(_lib.AlignedPair))))
            st = np.zeros(self.npairs + 1, dtype=np.int64)
            np.cumsum(counts[:, li], out=st[1:])
            lists.append(buf[:tot])
            starts.append(st)
        return lists, starts, lp

    def fetch(self):
        """Per pair: (matches, gapX, gapY, logProb)."""
        lists, starts, lp = self.fetch_arrays()
        per = []
        for li in range(3):
            st = starts[li].tolist()
            per.append([lists[li][a:e] for a, e in zip(st[:-1], st[1:])])
        return [(per[0][i], per[1][i], per[2][i], float(lp[i])) for i in range(self.npairs)]

    def set_model(self, machines):
        """Re-read a model's transitions and emission tables (same pairs)."""
        check(lib().mp_batch_set_model(self._h, machines.handle))

    def expectations(self):
        """Expectation mode: (npairs, 9) expected transition counts T[from*3+to]
        and the per-pair likelihood increment."""
        t = np.zeros((max(self.npairs, 1), 9), dtype=np.float64)
        lk = np.zeros(max(self.npairs, 1), dtype=np.float64)
        check(lib().mp_batch_fetch_expectations(self._h, t.ctypes.data_as(C.POINTER(C.c_double)),
                                                lk.ctypes.data_as(C.POINTER(C.c_double))))
        return t[: self.npairs], lk[: self.npairs]

    def emission_counts(self):
        """Expectation mode: expected emission counts summed over the pairs,
        on the forward-strand parameters: (match 4x4 [x, y], gap X 4, gap Y 4,
        repeat table 4x51x51 [base, underlying, observed])."""
        m, gx, gy = np.zeros(16), np.zeros(4), np.zeros(4)
        rl = np.zeros(4 * 51 * 51)
        P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        check(lib().mp_batch_fetch_emission_counts(self._h, P(m), P(gx), P(gy), P(rl)))
        return m.reshape(4, 4), gx, gy, rl.reshape(4, 51, 51)

    def close(self):
        if getattr(self, "_h", None):
            lib().mp_batch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _single(machines, strand, sX, sY, anchors, p, ragged_left, ragged_right, mode, engine=None):
    b = Batch(machines, [(sX, sY, anchors, strand, ragged_left, ragged_right)], p, mode=mode, engine=engine)
    b.run_plain()
    m, gx, gy, lp = b.fetch()[0]
    b.close()
    return m, gx, gy, lp


def getAlignedPairsWithIndelsUsingAnchors(machines, strand, sX, sY, anchorPairs, p, alignmentHasRaggedLeftEnd=False,
                                          alignmentHasRaggedRightEnd=False, engine=None):
    """Returns (alignedPairs, gapXPairs, gapYPairs)."""
    m, gx, gy, _ = _single(machines, strand, sX, sY, anchorPairs, p, alignmentHasRaggedLeftEnd,
                           alignmentHasRaggedRightEnd, _lib.MODE_POSTERIORS, engine)
    return m, gx, gy


def getAlignedPairsUsingAnchors(machines, strand, sX, sY, anchorPairs, p, alignmentHasRaggedLeftEnd=False,
                                alignmentHasRaggedRightEnd=False, engine=None):
    m, _, _, _ = _single(machines, strand, sX, sY, anchorPairs, p, alignmentHasRaggedLeftEnd,
                         alignmentHasRaggedRightEnd, _lib.MODE_MATCH_ONLY, engine)
    return m


def getAlignedPairs(machines, strand, sX, sY, p, alignmentHasRaggedLeftEnd=False, alignmentHasRaggedRightEnd=False,
                    engine=None):
    return getAlignedPairsUsingAnchors(machines, strand, sX, sY, None, p, alignmentHasRaggedLeftEnd,
                                       alignmentHasRaggedRightEnd, engine)


def getAlignedPairsWithIndels(machines, strand, sX, sY, p, alignmentHasRaggedLeftEnd=False,
                              alignmentHasRaggedRightEnd=False, engine=None):
    return getAlignedPairsWithIndelsUsingAnchors(machines, strand, sX, sY, None, p, alignmentHasRaggedLeftEnd,
                                                 alignmentHasRaggedRightEnd, engine)


def computeForwardProbability(machines, strand, sX, sY, anchorPairs, p, alignmentHasRaggedLeftEnd=False,
                              alignmentHasRaggedRightEnd=False, engine=None):
    """Forward log-likelihood with the static band of p.diagonalExpansion."""
    b = Batch(machines, [(sX, sY, anchorPairs, strand, alignmentHasRaggedLeftEnd, alignmentHasRaggedRightEnd)], p,
              mode=_lib.MODE_FORWARD_ONLY, engine=engine)
    b.run_plain()
    _, lp = b.counts()
    b.close()
    return float(lp[0])


def getExpectationsUsingAnchors(machines, strand, sX, sY, anchorPairs, p, alignmentHasRaggedLeftEnd=False,
                                alignmentHasRaggedRightEnd=False, engine=None):
    """Expected transition counts (9, [from*3+to]) and the likelihood increment."""
    b = Batch(machines, [(sX, sY, anchorPairs, strand, alignmentHasRaggedLeftEnd, alignmentHasRaggedRightEnd)], p,
              mode=_lib.MODE_EXPECTATIONS, engine=engine)
    b.run_plain()
    t, lk = b.expectations()
    b.close()
    return t[0], float(lk[0])


def getExpectations(machines, strand, sX, sY, p, alignmentHasRaggedLeftEnd=False, alignmentHasRaggedRightEnd=False,
                    engine=None):
    return getExpectationsUsingAnchors(machines, strand, sX, sY, None, p, alignmentHasRaggedLeftEnd,
                                       alignmentHasRaggedRightEnd, engine)


def getAlignedPairsWithIndelsCroppingReference(params, ref_sym, read_sym, read_strand, anchors, engine=None):
    """polisher.c:605-654 for one read: returns (matches, inserts, deletes) in
    uncropped reference coordinates (gap-X -> deletes, gap-Y -> inserts)."""
    b = Batch(params.machines, [(ref_sym, read_sym, anchors, read_strand)], params.p, crop_reference=True,
              engine=engine)
    b.run_plain()
    m, gx, gy, _ = b.fetch()[0]
    b.close()
    return m, gy, gx


def realign_all(params, pairs, engine=None):
    """The per-read loop of poa_realign over every read of every chunk, as one
    batched launch.  pairs: (ref_sym, read_sym, anchors, strand) per read.
    Returns per read (matches, inserts, deletes)."""
    b = Batch(params.machines, pairs, params.p, crop_reference=True, engine=engine)
    b.run()
    out = [(m, gy, gx) for (m, gx, gy, _) in b.fetch()]
    b.close()
    return out


_SYMBOL_OF_CHAR = np.full(256, 4, dtype=np.uint16)
for _c, _v in zip(b"ACGTacgt", (0, 1, 2, 3, 0, 1, 2, 3)):
    _SYMBOL_OF_CHAR[_c] = _v


def rleString_constructSymbolString(rle, includeRepeatCounts=True, maxRepeatCountExclusive=MAX_REPEAT_EXCLUSIVE):
    """stateMachine.c:124-146: uint16 symbols (run length << 8 | base) of a
    run-length string (bases, repeatCounts), run lengths capped at
    maxRepeatCountExclusive - 1."""
    from .poa import _as_rle
    b, c = _as_rle(rle)
    sym = _SYMBOL_OF_CHAR[b]
    if includeRepeatCounts:
        sym = sym | (np.minimum(c, maxRepeatCountExclusive - 1).astype(np.uint16) << 8)
    return sym.astype(np.uint16)


def pack_bubbles(bubbles):
    """Flatten [(alleles, [(read, strand)])] (run-length strings as (bases,
    repeatCounts)) into the packed arrays bubble_allele_supports_packed takes."""
    al = [a for alleles, _ in bubbles for a in alleles]
    rd = [r for _, reads in bubbles for r, _ in reads]

    def cat(strings):
        bl = [np.frombuffer(b, np.uint8) if isinstance(b, (bytes, bytearray)) else np.asarray(b, np.uint8)
              for b, _ in strings]
        lens = np.array([x.size for x in bl], dtype=np.int64)
        off = np.zeros(lens.size + 1, dtype=np.int64)
        np.cumsum(lens, out=off[1:])
        b = np.concatenate(bl) if bl else np.zeros(0, np.uint8)
        c = np.concatenate([np.asarray(c, np.uint64) for _, c in strings]) if strings else np.zeros(0, np.uint64)
        return b, c, off

    ab, ac, aoff = cat(al)
    rb, rc, roff = cat(rd)
    a_start = np.zeros(len(bubbles) + 1, dtype=np.int64)
    r_start = np.zeros(len(bubbles) + 1, dtype=np.int64)
    np.cumsum([len(a) for a, _ in bubbles], out=a_start[1:])
    np.cumsum([len(r) for _, r in bubbles], out=r_start[1:])
    strands = np.array([int(bool(st)) for _, reads in bubbles for _, st in reads] or [0], dtype=np.int32)
    return {"allele_bases": ab, "allele_counts": ac, "allele_off": aoff, "allele_start": a_start,
            "read_bases": rb, "read_counts": rc, "read_off": roff, "read_start": r_start, "read_strand": strands}


def _rle_records(bases, counts, off):
    n = off.size - 1
    rec = np.zeros(max(n, 1), dtype=_lib.RLE_STRING_DTYPE)
    if n:
        rec["bases"][:n] = bases.ctypes.data + off[:-1]
        rec["repeatCounts"][:n] = counts.ctypes.data + 8 * off[:-1]
        rec["length"][:n] = np.diff(off)
    return rec


def bubble_allele_supports_packed(machines, p, pk, include_repeat_counts=True,
                                  max_repeat_count_exclusive=MAX_REPEAT_EXCLUSIVE, engine=None):
    """mp_bubble_allele_supports on packed bubbles (pack_bubbles): returns the
    flat float32 supports (bubble-major, allele-major within a bubble) and the
    number of forward computations run."""
    engine = engine or default_engine()
    for k in ("allele_bases", "read_bases"):
        if pk[k].size == 0:
            pk[k] = np.zeros(1, np.uint8)
    for k in ("allele_counts", "read_counts"):
        if pk[k].size == 0:
            pk[k] = np.zeros(1, np.uint64)
    A = _rle_records(pk["allele_bases"], pk["allele_counts"], pk["allele_off"])
    R = _rle_records(pk["read_bases"], pk["read_counts"], pk["read_off"])
    a_st, r_st = pk["allele_start"], pk["read_start"]
    n = a_st.size - 1
    out = np.zeros(max(int((np.diff(a_st) * np.diff(r_st)).sum()), 1), dtype=np.float32)
    npairs = C.c_int64()
    pc = p.c()
    check(lib().mp_bubble_allele_supports(engine.handle, machines.handle, C.byref(pc), n, a_st.ctypes.data,
                                          A.ctypes.data, r_st.ctypes.data, R.ctypes.data,
                                          pk["read_strand"].ctypes.data, int(bool(include_repeat_counts)),
                                          int(max_repeat_count_exclusive), out.ctypes.data, C.byref(npairs)))
    return out, npairs.value


def bubble_allele_supports(machines, p, bubbles, include_repeat_counts=True,
                           max_repeat_count_exclusive=MAX_REPEAT_EXCLUSIVE, engine=None):
    """The per-read x per-allele forward log-likelihoods of bubbleGraph.c:895-928
    for many bubbles in one forward-only launch (mp_bubble_allele_supports).

    bubbles: list of (alleles, reads): alleles a list of run-length strings
    (bases, repeatCounts), reads a list of ((bases, repeatCounts), strand).
    Returns one float32 array (alleleNo, readNo) per bubble, and the number of
    forward computations run (repeated read substrings of a bubble reuse the
    first one's values, as the reference does)."""
    pk = pack_bubbles(bubbles)
    flat, npairs = bubble_allele_supports_packed(machines, p, pk, include_repeat_counts,
                                                 max_repeat_count_exclusive, engine)
    a_st, r_st = pk["allele_start"], pk["read_start"]
    res, at = [], 0
    for b in range(len(bubbles)):
        na, nr = int(a_st[b + 1] - a_st[b]), int(r_st[b + 1] - r_st[b])
        res.append(flat[at:at + na * nr].reshape(na, nr).copy())
        at += na * nr
    return res, npairs
