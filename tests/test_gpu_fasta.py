"""Polished consensus against the reference's pipeline in the reference's
arithmetic (tests/ref_pipeline.py): the initial poa_realign, three consensus
rounds of poa_realignIterative (polisher.c:2034-2106, consensus path, the
bubble-graph step off on both sides: maxRealignmentPolishIterations = 0) and
the run lengths of marginPolish.c:554-560, per chunk.

- Config 0 (one 5 kb chunk of a draft at 3% error, 20 full-length reads at
  10%) and config 1 (a 400 kb draft in 40 chunks of 10 kb with 500 bp
  boundaries, 60x reads of mean 8 kb at 9%): the GPU pipeline with
  reference-arithmetic batches (MP_BATCH_REFERENCE_ARITHMETIC) gives FASTA
  identical to the reference pipeline's (the oracle's posteriors in the
  reference's arithmetic; the oracle's POA restatement for config 0, the
  library's host POA code, equal to it bit for bit, for config 1).
- The FP32 pipeline decides near-tied consensus choices as exact arithmetic
  does: on config 0 it equals the same pipeline with exact posteriors, and its
  difference to the reference (whose polynomial logAdd moves posteriors by up
  to ~3e-3) is recorded: edits per 100 kb, bounded to catch regressions."""
import os
import sys
import time

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import PARAMS
from tests import ref_pipeline as RP

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _cfg():
    from marginpolish_amd import polish
    cfg = polish.PolishConfig(PARAMS)
    cfg.maxRealignmentPolishIterations = 0
    cfg.maxPoaConsensusIterations = 3
    return cfg


def _edits(a, b):
    from marginpolish_amd.sim import edit_distance
    return edit_distance(a, b)


def test_config0_fasta_identical_to_reference_pipeline():
    from marginpolish_amd import pairhmm, polish, sim
    params, op, cfg = pairhmm.PolishParams(PARAMS), O.PolishParams(PARAMS), _cfg()
    truth, draft, chunks = sim.polish_workload(seed=20, genome_len=5000, draft_err=0.03, read_err=0.10, depth=20,
                                               mean=5000, sd=1, chunk_size=5000, boundary=0, W=32, rle=True)
    refs, reads, anchors = [c.ref for c in chunks], [c.reads for c in chunks], [c.anchors for c in chunks]
    say("case: config0 FASTA (1 chunk, %d reads)" % len(reads[0]))
    stats = {}
    cpu = RP.realign_all(op, cfg, refs, reads, anchors, poa="oracle", stats=stats)
    cpu_exact = RP.realign_all(op, cfg, refs, reads, anchors, poa="oracle", mode=O.LOGADD_EXACT)
    g32 = [s for _, s in polish.realign_all(params, cfg, refs, reads, anchors)]
    gref = [s for _, s in polish.realign_all(params, cfg, refs, reads, anchors, reference_arithmetic=True)]
    chars = np.frombuffer(b"ACGT", np.uint8)
    t = chars[truth].tobytes().decode()
    d = chars[draft].tobytes().decode()
    say("config0: realignments after the first %d; edit distance to the truth: draft %d, reference pipeline %d, "
        "FP32 GPU %d; FP32 vs reference pipeline %d edits, vs the exact-arithmetic pipeline %d"
        % (stats["realigned"], _edits(d, t), _edits(cpu[0], t), _edits(g32[0], t), _edits(g32[0], cpu[0]),
           _edits(g32[0], cpu_exact[0])))
    assert stats["realigned"] >= 1
    assert gref == cpu
    assert g32 == cpu_exact
    assert _edits(g32[0], cpu[0]) <= 10


def test_config1_fasta_vs_reference_pipeline():
    from marginpolish_amd import pairhmm, polish, sim
    params, op, cfg = pairhmm.PolishParams(PARAMS), O.PolishParams(PARAMS), _cfg()
    truth, draft, chunks = sim.polish_workload(seed=21, genome_len=400_000, draft_err=0.03, read_err=0.09, depth=60,
                                               mean=8000, sd=6000, chunk_size=10_000, boundary=500, W=10, rle=True)
    refs, reads, anchors = [c.ref for c in chunks], [c.reads for c in chunks], [c.anchors for c in chunks]
    say("case: config1 FASTA (%d chunks, %d chunk reads)" % (len(chunks), sum(len(r) for r in reads)))
    cpu = RP.realign_all(op, cfg, refs, reads, anchors, poa="library", machines=params.machines)
    gref = [s for _, s in polish.realign_all(params, cfg, refs, reads, anchors, reference_arithmetic=True)]
    g32 = [s for _, s in polish.realign_all(params, cfg, refs, reads, anchors)]
    total = sum(len(s) for s in cpu)
    edits = sum(_edits(a, b) for a, b in zip(g32, cpu))
    same = sum(a == b for a, b in zip(g32, cpu))
    say("config1: %d of %d chunks identical in FP32; %d edits over %d bases = %.2f per 100 kb; reference "
        "arithmetic identical in %d of %d chunks" % (same, len(cpu), edits, total, edits * 1e5 / total,
                                                    sum(a == b for a, b in zip(gref, cpu)), len(cpu)))
    assert gref == cpu
    assert edits * 1e5 / total <= 60.0


CONFIG2 = dict(draft_err=0.02, read_err=0.08, depth=50, mean=20000, sd=15000, chunk_size=100_000, boundary=5000,
               W=10, rle=True, gc=0.5, p_geom=0.5, tandem=0.01, hi=100_000)
FULL = pytest.mark.skipif(not os.environ.get("MP_FULL_CONFIG2"),
                          reason="the full 5 Mb contig takes ~5 minutes of CPU reference work on 16 threads; "
                                 "set MP_FULL_CONFIG2=1")


@pytest.mark.parametrize("genome_len", [1_000_000, pytest.param(5_000_000, marks=FULL)])
def test_config2_fasta_vs_reference_pipeline(genome_len):
    """Config 2's generator (GC 50%, homopolymers Geometric(0.5), 1% 2-6-mer
    tandem repeats; 50x reads of log-normal length, mean 20 kb, sd 15 kb, max
    100 kb, 8% error; a draft at 2%; chunks of 100 kb with 10 kb overlaps;
    up to 3 consensus iterations), the contig merged with mergeContigChunks:
    the reference-arithmetic GPU pipeline's FASTA is identical to the CPU
    reference pipeline's (posteriors and cuts from the oracle in the
    reference's arithmetic, POA from the library's host code); the FP32
    pipeline's edits against it are reported per 100 kb."""
    from marginpolish_amd import pairhmm, polish, stitch, sim
    params, op, cfg = pairhmm.PolishParams(PARAMS), O.PolishParams(PARAMS), _cfg()
    t0 = time.time()
    truth, draft, chunks = sim.polish_workload(seed=22, genome_len=genome_len, **CONFIG2)
    refs, reads, anchors = [c.ref for c in chunks], [c.reads for c in chunks], [c.anchors for c in chunks]
    overlap = 2 * CONFIG2["boundary"]
    say("case: config2 FASTA, %d bp contig, %d chunks, %d chunk reads (built in %.1f s)"
        % (genome_len, len(chunks), sum(len(r) for r in reads), time.time() - t0))
    t0 = time.time()
    stats = {}
    cpu = RP.realign_all(op, cfg, refs, reads, anchors, poa="library", machines=params.machines, stats=stats)
    t_cpu = time.time() - t0
    cpu_fasta = RP.merge_contig_chunks(cpu, overlap, op)
    t_merge = time.time() - t0 - t_cpu
    say("CPU reference pipeline: %.1f s (%d chunk realignments after the first), merge %.1f s"
        % (t_cpu, stats["realigned"], t_merge))
    t0 = time.time()
    gref = [s for _, s in polish.realign_all(params, cfg, refs, reads, anchors, reference_arithmetic=True)]
    gref_fasta = stitch.mergeContigChunks(gref, overlap, params)
    t_ref = time.time() - t0
    t0 = time.time()
    g32 = [s for _, s in polish.realign_all(params, cfg, refs, reads, anchors)]
    g32_fasta = stitch.mergeContigChunks(g32, overlap, params)
    t_32 = time.time() - t0
    total = sum(len(s) for s in cpu)
    edits = sum(_edits(a, b) for a, b in zip(g32, cpu))
    say("config2 %d bp: GPU reference arithmetic %.1f s, FASTA identical %s (%d of %d chunks); FP32 %.1f s, "
        "%d of %d chunks identical, %d edits over %d bases = %.2f per 100 kb; FASTA %d bp"
        % (genome_len, t_ref, gref_fasta == cpu_fasta, sum(a == b for a, b in zip(gref, cpu)), len(cpu), t_32,
           sum(a == b for a, b in zip(g32, cpu)), len(cpu), edits, total, edits * 1e5 / total, len(cpu_fasta)))
    assert stats["realigned"] >= 1
    assert gref == cpu
    assert gref_fasta == cpu_fasta
    assert edits * 1e5 / total <= 60.0
