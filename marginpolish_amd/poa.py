"""POA augmentation from the pair-HMM's posterior lists: the host mirror of
the reference's poa_getReferenceGraph / poa_augment / the augment loop of
poa_realign (impl/polisher.c:118-131, 316-534, 656-703), over the native
implementation in libmp_pairhmm.so (include/mp_poa.h).

Run-length strings are (bases, repeatCounts): bases a str / bytes / uint8
array of characters, repeatCounts integers >= 1 (RleString, inc/margin.h).
Aligned-pair lists are numpy arrays of ALIGNED_PAIR_DTYPE or sequences of
(weight, x, y) tuples.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib

MAXIMUM_REPEAT_LENGTH = 51  # margin.h:123 (exclusive)
_BASES = np.frombuffer(b"ACGTN", dtype=np.uint8)


def rle_from_symbols(sym, runs=None):
    """An RLE string from uint16 pair-HMM symbols (run length << 8 | base):
    bases as characters; repeat counts from `runs` (uncapped) when given,
    else the symbols' own (capped) run lengths."""
    sym = np.asarray(sym, dtype=np.uint16)
    bases = _BASES[np.minimum(sym & 0xFF, 4)]
    counts = np.asarray(runs if runs is not None else (sym >> 8), dtype=np.uint64)
    return bases, counts


def _as_rle(r):
    bases, counts = r
    if isinstance(bases, str):
        bases = bases.encode()
    b = np.ascontiguousarray(np.frombuffer(bases, dtype=np.uint8) if isinstance(bases, (bytes, bytearray))
                             else np.asarray(bases, dtype=np.uint8))
    c = np.ascontiguousarray(np.asarray(counts, dtype=np.uint64))
    if b.shape != c.shape or b.ndim != 1:
        raise ValueError("bases and repeatCounts must be 1-d of equal length")
    return b, c


class PackedReads:
    """Run-length reads laid out once as the library's mp_rle_string array
    (the reads do not change between realignment iterations)."""

    def __init__(self, reads):
        self._keep = [_as_rle(r) for r in reads]
        self.records = np.empty(max(len(self._keep), 1), dtype=_lib.RLE_STRING_DTYPE)
        if self._keep:
            self.records["bases"][: len(self._keep)] = [b.ctypes.data for b, _ in self._keep]
            self.records["repeatCounts"][: len(self._keep)] = [c.ctypes.data for _, c in self._keep]
            self.records["length"][: len(self._keep)] = [b.size for b, _ in self._keep]

    def __len__(self):
        return len(self._keep)


def _rle_struct(b, c):
    return _lib.RleString(b.ctypes.data, c.ctypes.data, b.size)


def _as_pairs(t):
    if isinstance(t, np.ndarray) and t.dtype == np.dtype(_lib.ALIGNED_PAIR_DTYPE):
        return np.ascontiguousarray(t)
    a = np.zeros(len(t), dtype=_lib.ALIGNED_PAIR_DTYPE)
    if len(t):
        arr = np.asarray(t, dtype=np.int64).reshape(-1, 3)
        a["weight"], a["x"], a["y"] = arr[:, 0], arr[:, 1], arr[:, 2]
    return a


class Poa:
    """A POA graph (struct _Poa, inc/margin.h:590-597) held by the library."""

    def __init__(self, reference, max_repeat_count=MAXIMUM_REPEAT_LENGTH):
        b, c = _as_rle(reference)
        h = C.c_void_p()
        check(lib().mp_poa_reference_graph(C.byref(_rle_struct(b, c)), int(max_repeat_count), C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def augment(self, read, read_strand, read_no, matches, inserts, deletes, compare_repeat_counts=True):
        """poa_augment (polisher.c:316-534)."""
        b, c = _as_rle(read)
        m, i, d = _as_pairs(matches), _as_pairs(inserts), _as_pairs(deletes)
        check(lib().mp_poa_augment(self._h, C.byref(_rle_struct(b, c)), int(bool(read_strand)), int(read_no),
                                   m.ctypes.data, m.size, i.ctypes.data, i.size, d.ctypes.data, d.size,
                                   int(bool(compare_repeat_counts))))

    def sizes(self):
        s = _lib.PoaSizes()
        check(lib().mp_poa_get_sizes(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in _lib.PoaSizes._fields_}

    def nodes(self):
        """Per node: bases (str), repeatCounts, baseWeights (n, 5), repeatCountWeights (n, maxRepeatCount)."""
        s = self.sizes()
        n, mrc = s["nodes"], s["maxRepeatCount"]
        bases = np.empty(n, dtype=np.uint8)
        rc = np.empty(n, dtype=np.uint64)
        bw = np.empty((n, 5), dtype=np.float64)
        rw = np.empty((n, mrc), dtype=np.float64)
        check(lib().mp_poa_get_nodes(self._h, bases.ctypes.data, rc.ctypes.data, bw.ctypes.data, rw.ctypes.data))
        return {"bases": bases.tobytes().decode(), "repeatCounts": rc, "baseWeights": bw, "repeatCountWeights": rw}

    def inserts(self):
        """Structure-of-arrays over every PoaInsert (nodes ascending, list order)."""
        s = self.sizes()
        k, ns = s["inserts"], s["insertSymbols"]
        node = np.empty(k, np.int64)
        wf, wr = np.empty(k), np.empty(k)
        st, ln = np.empty(k, np.int64), np.empty(k, np.int64)
        bases = np.empty(max(ns, 1), np.uint8)
        counts = np.empty(max(ns, 1), np.uint64)
        check(lib().mp_poa_get_inserts(self._h, node.ctypes.data, wf.ctypes.data, wr.ctypes.data, st.ctypes.data,
                                       ln.ctypes.data, bases.ctypes.data, counts.ctypes.data))
        return {"node": node, "weightForwardStrand": wf, "weightReverseStrand": wr, "symbolStart": st,
                "symbolLength": ln, "bases": bases[:ns], "repeatCounts": counts[:ns]}

    def deletes(self):
        s = self.sizes()
        k = s["deletes"]
        node, ln = np.empty(k, np.int64), np.empty(k, np.int64)
        wf, wr = np.empty(k), np.empty(k)
        check(lib().mp_poa_get_deletes(self._h, node.ctypes.data, ln.ctypes.data, wf.ctypes.data, wr.ctypes.data))
        return {"node": node, "length": ln, "weightForwardStrand": wf, "weightReverseStrand": wr}

    def observations(self, kind):
        """kind 0 = nodes, 1 = inserts, 2 = deletes: (start, records)."""
        s = self.sizes()
        ent = {0: s["nodes"], 1: s["inserts"], 2: s["deletes"]}[kind]
        tot = {0: s["nodeObservations"], 1: s["insertObservations"], 2: s["deleteObservations"]}[kind]
        start = np.empty(ent + 1, np.int64)
        obs = np.empty(max(tot, 1), dtype=_lib.OBSERVATION_DTYPE)
        check(lib().mp_poa_get_observations(self._h, int(kind), start.ctypes.data, obs.ctypes.data))
        return start, obs[:tot]

    def estimate_repeat_counts(self, machines, reads, read_strands, threads=0):
        """poa_estimateRepeatCountsUsingBayesianModel (polisher.c:1557-1566): the
        maximum-likelihood run length of every reference node (nodes 1..n) from
        its base observations; reads[readNo] = (bases, repeatCounts) with
        uncapped counts, read_strands[readNo] = 1 for forward reads."""
        keep = [_as_rle(r) for r in reads]
        rs = np.empty(max(len(keep), 1), dtype=_lib.RLE_STRING_DTYPE)
        for i, (b, c) in enumerate(keep):
            rs[i] = (b.ctypes.data, c.ctypes.data, b.size)
        st = np.ascontiguousarray(np.asarray(list(read_strands) or [0], dtype=np.int32))
        out = np.zeros(max(self.sizes()["nodes"] - 1, 1), dtype=np.uint64)
        check(lib().mp_poa_estimate_repeat_counts(self._h, machines.handle, rs.ctypes.data, st.ctypes.data,
                                                  len(keep), out.ctypes.data, int(threads)))
        del keep
        return out[: self.sizes()["nodes"] - 1]

    def anchor_alignments(self, n_reads, anchor_params, poa_to_consensus=None):
        """poa_getAnchorAlignments (polisher.c:536-592): per read an (n, 3)
        int64 array of (consensus position, read offset, expansion) anchors.
        anchor_params = minPosteriorProbForAlignmentAnchors (threshold,
        expansion pairs); poa_to_consensus maps node i to consensus position
        poa_to_consensus[i - 1] (-1 = unaligned), None = identity."""
        ap = np.ascontiguousarray(np.asarray(anchor_params, dtype=np.float64))
        cm = None if poa_to_consensus is None else np.ascontiguousarray(np.asarray(poa_to_consensus, np.int64))
        counts = np.zeros(max(n_reads, 1), dtype=np.int64)
        cmp_ = None if cm is None else cm.ctypes.data
        cap = max(self.sizes()["nodeObservations"], 1)  # every anchor comes from one base observation
        buf = np.empty((cap, 3), dtype=np.int64)
        check(lib().mp_poa_anchor_alignments(self._h, cmp_, int(n_reads), ap.ctypes.data, ap.size,
                                             counts.ctypes.data, buf.ctypes.data, cap))
        ends = np.cumsum(counts[:n_reads])
        return [buf[e - c:e] for e, c in zip(ends, counts[:n_reads])]

    def consensus(self, use_run_length_encoding=True, reference_base_penalty=0.5):
        """poa_getConsensus (polisher.c:1040-1275): the expanded consensus string
        and poaToConsensusMap (node i -> position in the run-length consensus,
        -1 where none)."""
        n = self.sizes()["nodes"]
        total = C.c_int64()
        cmap = np.zeros(max(n - 1, 1), dtype=np.int64)
        lib().mp_poa_consensus(self._h, int(bool(use_run_length_encoding)), float(reference_base_penalty), None, 0,
                               C.byref(total), cmap.ctypes.data)
        buf = C.create_string_buffer(max(total.value, 1))
        check(lib().mp_poa_consensus(self._h, int(bool(use_run_length_encoding)), float(reference_base_penalty), buf,
                                     total.value, C.byref(total), cmap.ctypes.data))
        return buf.raw[:total.value].decode(), cmap[: n - 1]

    def scores(self):
        """(reference match weight, disagreement weight, insert weight, delete
        weight) as poa_realignIterative scores the graph (polisher.c:706-758)."""
        out = np.zeros(4)
        check(lib().mp_poa_scores(self._h, out.ctypes.data))
        return tuple(float(v) for v in out)

    def score(self):
        m, d, i, x = self.scores()
        return m - (x + i + d)

    def close(self):
        if getattr(self, "_h", None):
            lib().mp_poa_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poa_getReferenceGraph(reference, maxRepeatCount=MAXIMUM_REPEAT_LENGTH):
    return Poa(reference, maxRepeatCount)


def poa_augment(poa, read, readStrand, readNo, matches, inserts, deletes, poaConstructCompareRepeatCounts=True):
    poa.augment(read, readStrand, readNo, matches, inserts, deletes, poaConstructCompareRepeatCounts)


def poa_estimateRepeatCountsUsingBayesianModel(poa, machines, reads, read_strands):
    return poa.estimate_repeat_counts(machines, reads, read_strands)


def augment_batch(batch, poas, pair_poa, reads, pair_read_no=None, compare_repeat_counts=True, threads=0):
    """The augment loop of poa_realign (polisher.c:668-698) for every pair of
    a posterior batch that has run, straight from the batch's lists: pair i
    is read pair_read_no[i] (default: its index) of poas[pair_poa[i]], with
    run-length string reads[i] = (bases, repeatCounts)."""
    n = batch.npairs
    if len(reads) != n or len(pair_poa) != n:
        raise ValueError("one graph index and one read string per pair")
    packed = reads if isinstance(reads, PackedReads) else PackedReads(reads)
    rs = packed.records
    pp = np.ascontiguousarray(np.asarray(pair_poa, dtype=np.int64))
    rn = np.ascontiguousarray(np.asarray(pair_read_no if pair_read_no is not None else np.arange(n), dtype=np.int64))
    handles = (C.c_void_p * max(len(poas), 1))(*[p.handle for p in poas])
    check(lib().mp_batch_poa_augment(batch.handle, handles, len(poas), pp.ctypes.data, rn.ctypes.data,
                                     rs.ctypes.data, int(bool(compare_repeat_counts)), int(threads)))
    del packed


def estimate_repeat_counts_many(graphs, machines, reads_per_graph, strands_per_graph, threads=0):
    """poa_estimateRepeatCountsUsingBayesianModel for many graphs in one
    call (graphs in parallel): reads_per_graph[g] = graph g's reads by readNo
    (a list of (bases, counts) or a PackedReads), strands_per_graph[g] their
    strands.  Returns one uint64 array of run lengths per graph."""
    recs, strands, start = [], [], [0]
    keep = []
    for rs, st in zip(reads_per_graph, strands_per_graph):
        pk = rs if isinstance(rs, PackedReads) else PackedReads(rs)
        keep.append(pk)
        recs.append(pk.records[: len(pk)])
        strands.append(np.asarray(list(st), dtype=np.int32))
        start.append(start[-1] + len(pk))
    allrec = np.concatenate(recs) if recs else np.zeros(1, _lib.RLE_STRING_DTYPE)
    if allrec.size == 0:
        allrec = np.zeros(1, _lib.RLE_STRING_DTYPE)
    allst = np.concatenate(strands).astype(np.int32) if strands else np.zeros(1, np.int32)
    if allst.size == 0:
        allst = np.zeros(1, np.int32)
    st_arr = np.asarray(start, dtype=np.int64)
    outs = [np.zeros(max(g.sizes()["nodes"] - 1, 1), dtype=np.uint64) for g in graphs]
    hs = (C.c_void_p * max(len(graphs), 1))(*[g.handle for g in graphs])
    ptrs = (C.c_void_p * max(len(graphs), 1))(*[o.ctypes.data for o in outs])
    check(lib().mp_poas_estimate_repeat_counts(hs, len(graphs), machines.handle, allrec.ctypes.data, allst.ctypes.data,
                                               st_arr.ctypes.data, ptrs, int(threads)))
    del keep
    return [o[: g.sizes()["nodes"] - 1] for o, g in zip(outs, graphs)]

