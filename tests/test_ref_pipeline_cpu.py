"""The CPU pipeline in the reference's arithmetic (tests/ref_pipeline.py)
gives the same polished strings with the oracle's POA restatement and with
the library's host POA code, over several consensus rounds."""
import numpy as np

from marginpolish_amd import sim
from oracle import oracle as O
from tests.conftest import PARAMS
from tests import ref_pipeline as RP


class Cfg:
    useRunLengthEncoding = True
    referenceBasePenalty = 0.5
    poaConstructCompareRepeatCounts = True
    minPosteriorProbForAlignmentAnchors = [0.9, 10, 0.95, 4, 0.99, 0]
    minPoaConsensusIterations = 0
    maxPoaConsensusIterations = 3
    maxRealignmentPolishIterations = 0


def workload_of(w):
    """One chunk: the simulated reference and its reads as run-length strings."""
    acgt = np.frombuffer(b"ACGT", np.uint8)
    p0 = w.pairs[0]
    ref = (acgt[p0.ref_sym & 255].tobytes(), np.asarray(p0.ref_runs, np.int64))
    reads = [((acgt[p.read_sym & 255].tobytes(), np.asarray(p.read_runs, np.int64)), int(p.strand)) for p in w.pairs]
    return [ref], [reads], [[p.anchors for p in w.pairs]]


def test_oracle_and_library_poa_agree():
    """A 2 kb draft at 3% error, 12x reads at 8%: three consensus rounds (the
    first realigns against a consensus that differs from the draft)."""
    from marginpolish_amd import pairhmm
    op = O.PolishParams(PARAMS)
    machines = pairhmm.PolishParams(PARAMS).machines
    truth, draft, chunks = sim.polish_workload(seed=3, genome_len=2000, depth=12, mean=1200, sd=400,
                                               chunk_size=4000, boundary=0, rle=True)
    refs, reads, anchors = [c.ref for c in chunks], [c.reads for c in chunks], [c.anchors for c in chunks]
    stats_a, stats_b = {}, {}
    a = RP.realign_all(op, Cfg, refs, reads, anchors, poa="oracle", stats=stats_a)
    b = RP.realign_all(op, Cfg, refs, reads, anchors, poa="library", machines=machines, stats=stats_b)
    assert a == b and stats_a == stats_b
    assert stats_a["realigned"] >= 1, stats_a


def _merge_sequential(chunks, overlap, op, missing):
    """chunker.c:50-118 one boundary at a time on the (cropped) predecessor."""
    pieces = []
    for cur in chunks:
        if not pieces:
            pieces.append(cur)
            continue
        prev = pieces[-1]
        w, ce, cs = RP.remove_overlap(prev, cur, overlap, op)
        if w > 0:
            pieces[-1] = prev[:ce]
            cur = cur[cs:]
        elif len(cur) == 0:
            cur = missing
        elif overlap != 0:
            pieces.append("NNNNNNNNNN")
        pieces.append(cur)
    return "".join(pieces)


def test_merge_replay_equals_sequential_merge():
    """merge_contig_chunks (cuts against uncropped predecessors, re-aligned
    where a crop reaches into the window) equals the one-boundary-at-a-time
    rules: short chunks whose crops reach into the next window, a missing
    chunk, an empty first chunk, overlap 0."""
    import random
    rng = random.Random(4)
    op = O.PolishParams(PARAMS)
    genome = "".join(rng.choice("ACGT") for _ in range(4000))
    size, bnd = 400, 40
    chunks = [genome[max(0, s - bnd):s + size + bnd] for s in range(0, len(genome), size)]
    chunks = ["".join(c if rng.random() > 0.01 else rng.choice("ACGT") for c in ch) for ch in chunks]
    chunks[3] = ""
    chunks[6] = chunks[6][:50]
    chunks[7] = chunks[7][:60]
    for ch, ov, miss in ((chunks, 2 * bnd, "NNNNN"), (chunks, 2 * bnd, ""), (["", chunks[1], chunks[2]], 2 * bnd, ""),
                         (chunks[:4], 0, "")):
        assert RP.merge_contig_chunks(ch, ov, op, miss, threads=3) == _merge_sequential(ch, ov, op, miss)
