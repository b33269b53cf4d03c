"""Chunk stitching (include/mp_stitch.h) against the oracle: removeOverlap's
alignment and cut, and mergeContigChunks' sequential rules."""
import random

import numpy as np
import pytest

from marginpolish_amd import pairhmm as ph
from marginpolish_amd import stitch
from oracle import oracle as O
from tests.conftest import PARAMS

pytestmark = pytest.mark.gpu


def _mutate(rng, s, rate):
    out = []
    for c in s:
        r = rng.random()
        if r < rate / 3:
            out.append(rng.choice("ACGT"))
        elif r < 2 * rate / 3:
            out.append(c + rng.choice("ACGT"))
        elif r < rate:
            continue
        else:
            out.append(c)
    return "".join(out)


def _oracle_tuples(prefix, suffix, approx, op):
    i = max(len(prefix) - approx, 0)
    j = min(approx, len(suffix))
    if len(prefix) - i == 0 or j == 0:
        return i, []
    mm, _, _, _, _ = O.posteriors(O.default_nucleotide_model(), O.symbols(prefix[i:]), O.symbols(suffix[:j]), [],
                                  op.p, True, True, with_indels=False, mode=O.LOGADD_EXACT)
    return i, [(int(w), int(x), int(y)) for w, x, y in mm]


def test_remove_overlaps_match_oracle():
    rng = random.Random(8)
    pp = ph.PolishParams(PARAMS)
    op = O.PolishParams(PARAMS)
    pre, suf, ov = [], [], []
    for k in range(40):
        core = "".join(rng.choice("ACGT") for _ in range(rng.randint(20, 300)))
        a = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 200)))
        b = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 200)))
        pre.append(a + core)
        suf.append(_mutate(rng, core, 0.02) + b)
        ov.append(len(core) + rng.randint(-10, 40))
    ce, cs, w = stitch.remove_overlaps(pp.p, pre, suf, ov)
    for k in range(len(pre)):
        i, tup = _oracle_tuples(pre[k], suf[k], ov[k], op)
        assert tup, k
        best = max(t[0] for t in tup)
        got = {(x + i, y): wt for wt, x, y in tup}
        assert (ce[k], cs[k]) in got, k
        # the chosen tuple is within the posterior tolerance (1e-4) of the oracle's best, and of its own weight
        assert got[(ce[k], cs[k])] >= best - 1000, (k, got[(ce[k], cs[k])], best)
        assert abs(got[(ce[k], cs[k])] - w[k]) <= 1000


def test_remove_overlap_known_example():
    """polisherTest.c:1001-1025 expects (6, 2): only the reference's polynomial
    logAdd gives it (a near tie); exact FP64 and this FP32 engine pick (7, 3)."""
    pp = ph.PolishParams(PARAMS)
    w, ce, cs = stitch.removeOverlap("ACGTGATTTCA", "GATTTCAACGT", 10, pp)
    assert (ce, cs) == (7, 3) and w > 0
    # nothing to align: keep both whole
    w, ce, cs = stitch.removeOverlap("ACGT", "", 10, pp)
    assert (w, ce, cs) == (0, 4, 0)


def _merge_reference(chunks, overlap, pp, spacer):
    """chunker.c:55-118 restated, one removeOverlap call per boundary on the
    actual (already cropped) previous string."""
    pieces = []
    for c, cur in enumerate(chunks):
        if not pieces:
            pieces.append(cur)
            continue
        prev = pieces[-1]
        w, ce, cs = stitch.removeOverlap(prev, cur, overlap, pp)
        if w > 0:
            pieces[-1] = prev[:ce]
            cur = cur[cs:]
        elif len(cur) == 0:
            cur = spacer
        elif overlap != 0:
            pieces.append("NNNNNNNNNN")
        pieces.append(cur)
    return "".join(pieces)


def test_merge_chunks_equals_sequential_merge():
    rng = random.Random(12)
    pp = ph.PolishParams(PARAMS)
    genome = "".join(rng.choice("ACGT") for _ in range(30000))
    size, bnd = 2500, 150
    chunks = []
    for s in range(0, len(genome), size):
        chunks.append(_mutate(rng, genome[max(0, s - bnd):s + size + bnd], 0.005))
    chunks[5] = ""                      # a missing chunk -> spacer (which the next boundary may crop)
    chunks[8] = chunks[8][:200]         # shorter than the overlap: its crops reach into the next window
    merged = stitch.mergeContigChunks(chunks, 2 * bnd, pp, "NNNNN")
    assert merged == _merge_reference(chunks, 2 * bnd, pp, "NNNNN")
    # first chunk empty (no overlap to find: a "NNNNNNNNNN" stitch), overlap 0
    odd = ["", chunks[1], chunks[2]]
    assert stitch.mergeContigChunks(odd, 2 * bnd, pp, "") == _merge_reference(odd, 2 * bnd, pp, "")
    assert stitch.mergeContigChunks(chunks[:3], 0, pp, "") == _merge_reference(chunks[:3], 0, pp, "")
    assert stitch.mergeContigChunks([], 10, pp, "") == ""


def test_merge_chunks_recovers_genome():
    """Chunks cut from one sequence without edits merge back to it."""
    rng = random.Random(2)
    pp = ph.PolishParams(PARAMS)
    genome = "".join(rng.choice("ACGT") for _ in range(20000))
    size, bnd = 3000, 200
    chunks = [genome[max(0, s - bnd):s + size + bnd] for s in range(0, len(genome), size)]
    assert stitch.mergeContigChunks(chunks, 2 * bnd, pp) == genome
