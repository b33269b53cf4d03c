"""One polishing run over the library: poa_realign for every read of every
chunk as one GPU batch, the consensus iterations of poa_realignIterative
(polisher.c:2034-2106) batched across the chunks still iterating,
maximum-likelihood run lengths, and the contig merge (marginPolish.c:554-560,
chunker.c:55-118).

Scope: the consensus path of poa_realignIterative (hmmMNotRealign); the
bubble-graph polish path (poa_polish, maxRealignmentPolishIterations) is not
built here and is refused rather than skipped.
"""
from __future__ import annotations

import json
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import pairhmm
from . import poa as mpoa
from . import stitch


class PolishConfig:
    """The polish-section fields this pipeline reads (params JSON, margin.h:534-580)."""

    def __init__(self, path):
        pol = json.load(open(path))["polish"]
        self.useRunLengthEncoding = bool(pol.get("useRunLengthEncoding", True))
        self.referenceBasePenalty = float(pol.get("referenceBasePenalty", 0.5))
        self.poaConstructCompareRepeatCounts = bool(pol.get("poaConstructCompareRepeatCounts", True))
        self.minPosteriorProbForAlignmentAnchors = [float(v) for v in
                                                    pol.get("minPosteriorProbForAlignmentAnchors", [0.9, 10])]
        self.minPoaConsensusIterations = int(pol.get("minPoaConsensusIterations", 0))
        self.maxPoaConsensusIterations = int(pol.get("maxPoaConsensusIterations", 0))
        self.maxRealignmentPolishIterations = int(pol.get("maxRealignmentPolishIterations", 0))
        self.chunkBoundary = int(pol.get("chunkBoundary", 50))


def _rle_of(expanded):
    b = np.frombuffer(expanded.encode(), np.uint8)
    if b.size == 0:
        return b.copy(), np.zeros(0, np.uint64)
    new = np.r_[True, b[1:] != b[:-1]]
    starts = np.flatnonzero(new)
    return b[starts].copy(), np.diff(np.r_[starts, b.size]).astype(np.uint64)


def _same_rle(a, b):
    return (np.array_equal(np.asarray(a[0], np.uint8), np.asarray(b[0], np.uint8)) and
            np.array_equal(np.asarray(a[1], np.uint64), np.asarray(b[1], np.uint64)))


def realign(params, cfg, refs, reads, anchors, engine=None, reference_arithmetic=False):
    """poa_realign (polisher.c:656-703) for several chunks at once: refs[c] =
    (bases, counts), reads[c] = [((bases, counts), strand)], anchors[c][r] =
    RLE anchors of read r.  One batch; every chunk's graph built from the
    batch's lists (mp_batch_poa_augment).  reference_arithmetic: the pair-HMM
    in the reference's own FP64 arithmetic (its weights exactly)."""
    cap = params.maximumRepeatLength if cfg.useRunLengthEncoding else 2
    pairs, pair_poa, pair_no, rle_reads = [], [], [], []
    for c, (ref, rs, an) in enumerate(zip(refs, reads, anchors)):
        ref_sym = pairhmm.rleString_constructSymbolString(ref, params.useRepeatCountsInAlignment, cap)
        for r, ((read, strand), a) in enumerate(zip(rs, an)):
            pairs.append((ref_sym, pairhmm.rleString_constructSymbolString(read, params.useRepeatCountsInAlignment,
                                                                            cap), a, strand))
            pair_poa.append(c)
            pair_no.append(r)
            rle_reads.append(read)
    graphs = [mpoa.Poa(ref, cap) for ref in refs]
    if pairs:
        b = pairhmm.Batch(params.machines, pairs, params.p, crop_reference=True, engine=engine,
                          reference_arithmetic=reference_arithmetic)
        b.run()
        mpoa.augment_batch(b, graphs, pair_poa, rle_reads, pair_no, cfg.poaConstructCompareRepeatCounts)
        b.close()
    return graphs


def _estimate(params, cfg, graph, ref, reads):
    if not cfg.useRunLengthEncoding:
        return ref
    counts = graph.estimate_repeat_counts(params.machines, [r for r, _ in reads], [s for _, s in reads], threads=1)
    return np.asarray(ref[0], np.uint8), counts


def realign_all(params, cfg, refs, reads, anchors, engine=None, reference_arithmetic=False):
    """poa_realignAll (polisher.c:2109-2129) for every chunk: the initial
    graph, then up to maxPoaConsensusIterations consensus rounds (greedy
    stopping per chunk), then the run lengths of marginPolish.c:557-560.
    Returns per chunk (graph, polished expanded string)."""
    if cfg.maxRealignmentPolishIterations > 0:
        raise NotImplementedError("the bubble-graph polish iterations (poa_polish) are not built; set "
                                  "maxRealignmentPolishIterations to 0")
    n = len(refs)
    graphs = realign(params, cfg, refs, reads, anchors, engine, reference_arithmetic)
    ref_strings = [(np.frombuffer(bytes(r[0]), np.uint8) if isinstance(r[0], (bytes, bytearray))
                    else np.asarray(r[0], np.uint8), np.asarray(r[1], np.uint64)) for r in refs]
    # per-chunk host steps run on a thread pool (the library calls release the interpreter)
    pool = ThreadPoolExecutor(max_workers=max(1, min(16, os.cpu_count() or 1)))
    try:
        scores = list(pool.map(lambda g: g.score(), graphs))
        active = [True] * n
        it = 0
        while it < cfg.maxPoaConsensusIterations and any(active):
            it += 1

            def next_round(c):
                cons, cmap = graphs[c].consensus(cfg.useRunLengthEncoding, cfg.referenceBasePenalty)
                cons_rle = _rle_of(cons) if cfg.useRunLengthEncoding else (
                    np.frombuffer(cons.encode(), np.uint8).copy(), np.ones(len(cons), np.uint64))
                if _same_rle(cons_rle, ref_strings[c]):  # greedy convergence
                    return c, None, None
                return c, cons_rle, graphs[c].anchor_alignments(len(reads[c]),
                                                                cfg.minPosteriorProbForAlignmentAnchors, cmap)

            todo, new_refs, new_anchors = [], [], []
            for c, cons_rle, an in pool.map(next_round, [c for c in range(n) if active[c]]):
                if cons_rle is None:
                    active[c] = False
                else:
                    todo.append(c)
                    new_refs.append(cons_rle)
                    new_anchors.append(an)
            if not todo:
                break
            new_graphs = realign(params, cfg, new_refs, [reads[c] for c in todo], new_anchors, engine,
                                 reference_arithmetic)

            def rescore(k):
                c, g2, ref2 = todo[k], new_graphs[k], new_refs[k]
                return _estimate(params, cfg, g2, ref2, reads[c]), g2.score()

            for k, (ref2, s2) in enumerate(pool.map(rescore, range(len(todo)))):
                c = todo[k]
                if s2 <= scores[c] and it > cfg.minPoaConsensusIterations:
                    active[c] = False  # the score did not improve: keep the previous graph
                    continue
                graphs[c], ref_strings[c], scores[c] = new_graphs[k], ref2, s2
        final = list(pool.map(lambda c: _estimate(params, cfg, graphs[c], ref_strings[c], reads[c]), range(n)))
    finally:
        pool.shutdown()
    out = []
    for c in range(n):
        b, cnt = final[c]
        out.append((graphs[c], np.repeat(b, cnt.astype(np.int64)).tobytes().decode()))
    return out


def polish_contig(params, cfg, refs, reads, anchors, engine=None, missing_spacer="", reference_arithmetic=False):
    """Polish every chunk of one contig and merge them (chunker.c:55-118 with
    overlap 2 x chunkBoundary)."""
    res = realign_all(params, cfg, refs, reads, anchors, engine, reference_arithmetic)
    return stitch.mergeContigChunks([s for _, s in res], 2 * cfg.chunkBoundary, params, missing_spacer, engine)
