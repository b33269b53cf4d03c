"""The consensus pipeline of poa_realignAll / poa_realignIterative
(polisher.c:2034-2129) with the run lengths of marginPolish.c:554-560, in the
reference's arithmetic on the CPU (test infrastructure).

Posteriors come from the oracle in the reference's FP64 arithmetic (the
polynomial logAdd, O_faithful), one pair at a time as poa_realign calls the
aligner after getAlignedPairsWithIndelsCroppingReference (polisher.c:605-703).
The POA steps (augmentation, consensus, anchors, run lengths, score) come
either from the oracle's restatement (oracle/poa.py: `poa="oracle"`, fully
independent of the library) or, for workloads too large for pure Python, from
the library's host POA code (`poa="library"`), which equals the restatement
bit for bit (tests/test_poa_host.py).  The rounds mirror
marginpolish_amd.polish.realign_all: the initial graph, consensus rounds with
greedy convergence and score stopping per chunk, then the run lengths.

refs[c] = (bases, counts), reads[c] = [((bases, counts), strand)],
anchors[c][r] = RLE anchors of read r (x in the chunk's reference)."""
from concurrent.futures import ThreadPoolExecutor
import os

import numpy as np

from marginpolish_amd import crop
from oracle import oracle as O
from oracle import poa as OP

ACGTN = np.frombuffer(b"ACGTN", np.uint8)


def _bytes(b):
    if isinstance(b, (bytes, bytearray)):
        return bytes(b)
    if isinstance(b, str):
        return b.encode()
    return np.asarray(b, np.uint8).tobytes()


def _symbols(rle, cap, with_counts):
    """rleString_constructSymbolString (stateMachine.c:124-146), run lengths capped at cap - 1."""
    b, c = rle
    if not with_counts:
        return O.symbols(_bytes(b).decode())
    return O.rle_symbols(_bytes(b).decode(), np.asarray(c, np.int64), cap)


def _rle_of(expanded):
    b = np.frombuffer(expanded.encode(), np.uint8)
    if b.size == 0:
        return b.copy(), np.zeros(0, np.int64)
    new = np.r_[True, b[1:] != b[:-1]]
    starts = np.flatnonzero(new)
    return b[starts].copy(), np.diff(np.r_[starts, b.size]).astype(np.int64)


def _same(a, b):
    return _bytes(a[0]) == _bytes(b[0]) and np.array_equal(np.asarray(a[1], np.int64), np.asarray(b[1], np.int64))


class _OracleGraph:
    def __init__(self, ref, cap, op):
        self.op = op
        self.g = OP.Poa((_bytes(ref[0]).decode(), [int(v) for v in ref[1]]), cap)

    def augment(self, read, strand, no, m, ins, dels, cmp_counts):
        r = (_bytes(read[0]).decode(), [int(v) for v in read[1]])
        OP.augment(self.g, r, strand, no, [tuple(map(int, t)) for t in m], [tuple(map(int, t)) for t in ins],
                   [tuple(map(int, t)) for t in dels], cmp_counts)

    def consensus(self, rle, penalty):
        return OP.consensus(self.g, rle, penalty)

    def anchors(self, n, params, cmap):
        return [np.array(a, np.int64).reshape(-1, 3) for a in OP.anchor_alignments(self.g, n, params, cmap)]

    def repeat_counts(self, reads):
        rs = [(_bytes(r[0]).decode(), [int(v) for v in r[1]]) for r, _ in reads]
        return np.array(OP.estimate_repeat_counts(self.g, rs, [bool(s) for _, s in reads], self.op.forward.repeat),
                        np.int64)

    def score(self):
        """poa_getReferenceNodeTotalMatchWeight - poa_getTotalErrorWeight (polisher.c:706-758)."""
        match = dis = ins = dele = 0.0
        for node in self.g.nodes:
            rs = OP.char_to_symbol(node.base)
            for j, w in enumerate(node.base_weights):
                if j == rs:
                    match += w
                else:
                    dis += w
            for e in node.inserts:
                ins += (e[1] + e[2]) * len(e[0][0])
            for e in node.deletes:
                dele += (e[1] + e[2]) * e[0]
        return match - (dele + ins + dis)


class _LibraryGraph:
    def __init__(self, ref, cap, op, machines):
        from marginpolish_amd import poa as mpoa
        self.machines = machines
        self.g = mpoa.Poa((_bytes(ref[0]), np.asarray(ref[1], np.uint64)), cap)

    def augment(self, read, strand, no, m, ins, dels, cmp_counts):
        self.g.augment((_bytes(read[0]), np.asarray(read[1], np.uint64)), strand, no, m, ins, dels, cmp_counts)

    def consensus(self, rle, penalty):
        return self.g.consensus(rle, penalty)

    def anchors(self, n, params, cmap):
        return self.g.anchor_alignments(n, params, cmap)

    def repeat_counts(self, reads):
        return np.asarray(self.g.estimate_repeat_counts(self.machines, [(_bytes(r[0]), np.asarray(r[1], np.uint64))
                                                                        for r, _ in reads],
                                                        [int(s) for _, s in reads], threads=1), np.int64)

    def score(self):
        return self.g.score()


def _posteriors(op, jobs, threads, mode):
    """The oracle's lists of every (ref symbols, read symbols, anchors,
    strand): (matches, inserts = gap-Y, deletes = gap-X), x in the uncropped
    reference; mode O.LOGADD_POLY is the reference's arithmetic."""
    def one(job):
        rs, qs, an, strand = job
        sx, sy, a2, off = crop.crop_pair(rs, qs, an)
        m, gx, gy, _, _ = O.posteriors(op.forward if strand else op.reverse, sx, sy, a2, op.p, mode=mode)
        out = []
        for lst in (m, gy, gx):
            lst = np.asarray(lst, np.int64).reshape(-1, 3).copy()
            lst[:, 1] += off
            out.append(lst)
        return out
    with ThreadPoolExecutor(max_workers=threads) as pool:
        return list(pool.map(one, jobs))


def realign(op, cfg, refs, reads, anchors, poa, machines=None, threads=None, mode=O.LOGADD_POLY):
    """poa_realign (polisher.c:656-703) for every chunk."""
    threads = threads or max(1, min(16, os.cpu_count() or 1))
    cap = O.MAX_REPEAT if cfg.useRunLengthEncoding else 2
    wc = op.useRepeatCountsInAlignment
    jobs, where = [], []
    for c, (ref, rs, an) in enumerate(zip(refs, reads, anchors)):
        ref_sym = _symbols(ref, cap, wc)
        for r, ((read, strand), a) in enumerate(zip(rs, an)):
            jobs.append((ref_sym, _symbols(read, cap, wc), a, strand))
            where.append((c, r))
    lists = _posteriors(op, jobs, threads, mode)
    graphs = [(_OracleGraph(ref, cap, op) if poa == "oracle" else _LibraryGraph(ref, cap, op, machines))
              for ref in refs]
    for (c, r), (m, ins, dels) in zip(where, lists):
        read, strand = reads[c][r]
        graphs[c].augment(read, strand, r, m, ins, dels, cfg.poaConstructCompareRepeatCounts)
    return graphs


def realign_all(op, cfg, refs, reads, anchors, poa="oracle", machines=None, threads=None, stats=None,
                mode=O.LOGADD_POLY):
    """poa_realignAll's consensus path for every chunk; returns the polished
    expanded string of every chunk.  stats (a dict) receives the number of
    chunk realignments after the first ("realigned"); mode = the posteriors'
    arithmetic (O.LOGADD_POLY: the reference's; O.LOGADD_EXACT: exact)."""
    n = len(refs)
    graphs = realign(op, cfg, refs, reads, anchors, poa, machines, threads, mode)
    ref_strings = [(_bytes(r[0]), np.asarray(r[1], np.int64)) for r in refs]
    scores = [g.score() for g in graphs]
    if stats is not None:
        stats["realigned"] = 0
    active = [True] * n
    it = 0
    while it < cfg.maxPoaConsensusIterations and any(active):
        it += 1
        todo, new_refs, new_anchors = [], [], []
        for c in range(n):
            if not active[c]:
                continue
            cons, cmap = graphs[c].consensus(cfg.useRunLengthEncoding, cfg.referenceBasePenalty)
            cons_rle = _rle_of(cons) if cfg.useRunLengthEncoding else (
                np.frombuffer(cons.encode(), np.uint8).copy(), np.ones(len(cons), np.int64))
            if _same(cons_rle, ref_strings[c]):  # greedy convergence
                active[c] = False
                continue
            todo.append(c)
            new_refs.append(cons_rle)
            new_anchors.append(graphs[c].anchors(len(reads[c]), cfg.minPosteriorProbForAlignmentAnchors, cmap))
        if not todo:
            break
        new_graphs = realign(op, cfg, new_refs, [reads[c] for c in todo], new_anchors, poa, machines, threads, mode)
        if stats is not None:
            stats["realigned"] = stats.get("realigned", 0) + len(todo)
        for k, c in enumerate(todo):
            g2, ref2 = new_graphs[k], new_refs[k]
            if cfg.useRunLengthEncoding:
                ref2 = (_bytes(ref2[0]), graphs_counts(g2, reads[c]))
            s2 = g2.score()
            if s2 <= scores[c] and it > cfg.minPoaConsensusIterations:
                active[c] = False  # the score did not improve: keep the previous graph
                continue
            graphs[c], ref_strings[c], scores[c] = g2, ref2, s2
    out = []
    for c in range(n):
        b = np.frombuffer(ref_strings[c][0], np.uint8)
        cnt = graphs_counts(graphs[c], reads[c]) if cfg.useRunLengthEncoding else ref_strings[c][1]
        out.append(np.repeat(b, np.asarray(cnt, np.int64)).tobytes().decode())
    return out


def graphs_counts(graph, reads):
    return graph.repeat_counts(reads)


def remove_overlap(prefix, suffix, approx, op):
    """removeOverlap (polisher.c:1937-2006) on the oracle in the reference's
    arithmetic: the prefix's last `approx` characters against the suffix's
    first `approx`, the default nucleotide machine, no anchors, both ends
    ragged; the first match tuple of maximal weight is the cut.  Returns
    (weight, prefix crop end, suffix crop start)."""
    i = max(len(prefix) - approx, 0)
    j = min(approx, len(suffix))
    if len(prefix) - i == 0 or j == 0:
        return 0, len(prefix), 0
    mm, _, _, _, _ = O.posteriors(O.default_nucleotide_model(), O.symbols(prefix[i:]), O.symbols(suffix[:j]), [],
                                  op.p, True, True, with_indels=False, mode=O.LOGADD_POLY)
    best = None
    for w, x, y in mm:
        if best is None or w > best[0]:
            best = (int(w), int(x), int(y))
    if best is None:
        return 0, len(prefix), 0
    return best[0], best[1] + i, best[2]


def merge_contig_chunks(chunks, overlap, op, missing="", threads=None):
    """mergeContigChunks (chunker.c:50-118) with remove_overlap.  The cut of
    every boundary against the uncropped predecessor is computed up front
    (in parallel); replaying the sequential rules, a boundary whose
    predecessor was cropped into its own overlap window is aligned again
    with the cropped string."""
    threads = threads or max(1, min(16, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=threads) as pool:
        pre = list(pool.map(lambda k: remove_overlap(chunks[k - 1], chunks[k], overlap, op), range(1, len(chunks))))
    out = []
    start = 0  # crop start of the last string in `out` within its chunk
    for k, cur in enumerate(chunks):
        if not out:
            out.append(cur)
            start = 0
            continue
        prev = out[-1]
        if start <= len(chunks[k - 1]) - overlap:
            w, ce, cs = pre[k - 1]
            ce -= start
        else:
            w, ce, cs = remove_overlap(prev, cur, overlap, op)
        if w > 0:
            out[-1] = prev[:ce]
            out.append(cur[cs:])
            start = cs
        elif len(cur) == 0:
            out.append(missing)
            start = len(chunks[k]) + 1  # never equal to a chunk's window: always aligned again
        elif overlap == 0:
            out.append(cur)
            start = 0
        else:
            out.append("NNNNNNNNNN")
            out.append(cur)
            start = 0
    return "".join(out)
