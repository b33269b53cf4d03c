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
