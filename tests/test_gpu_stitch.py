"""Chunk stitching (include/mp_stitch.h) against the oracle: removeOverlap's
alignment and cut, and mergeContigChunks' sequential rules."""
import random

import numpy as np
import pytest

from marginpolish_amd import pairhmm as ph
from marginpolish_amd import stitch
from oracle import oracle as O
from tests.conftest import PARAMS
from tests import ref_pipeline as RP

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


_oracle_cut = RP.remove_overlap


def test_remove_overlaps_match_oracle():
    """Every cut equals the reference's (the oracle in the reference's
    arithmetic): crop positions exactly, weights within one unit (the last bit
    of exp may differ between the GPU's and the host's FP64 exp)."""
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
        wo, ceo, cso = _oracle_cut(pre[k], suf[k], ov[k], op)
        assert (ce[k], cs[k]) == (ceo, cso), (k, (ce[k], cs[k]), (ceo, cso))
        assert abs(w[k] - wo) <= 1, (k, w[k], wo)


def test_remove_overlap_known_example():
    """polisherTest.c:1001-1025 expects (6, 2) -- a near tie that only the
    reference's polynomial logAdd resolves this way (exact arithmetic picks
    (7, 3)); stitching runs in the reference's arithmetic and returns (6, 2)."""
    pp = ph.PolishParams(PARAMS)
    w, ce, cs = stitch.removeOverlap("ACGTGATTTCA", "GATTTCAACGT", 10, pp)
    assert (ce, cs) == (6, 2) and w > 0
    # nothing to align: keep both whole
    w, ce, cs = stitch.removeOverlap("ACGT", "", 10, pp)
    assert (w, ce, cs) == (0, 4, 0)


def test_remove_overlap_wider_than_the_fp32_kernels():
    """A 3000-base overlap (a full 3000 x 3000 rectangle, wider than the
    2048-cell FP32 kernels) stitches, with the reference's cut."""
    rng = random.Random(31)
    pp = ph.PolishParams(PARAMS)
    op = O.PolishParams(PARAMS)
    core = "".join(rng.choice("ACGT") for _ in range(3000))
    pre = "".join(rng.choice("ACGT") for _ in range(500)) + core
    suf = _mutate(rng, core, 0.01) + "".join(rng.choice("ACGT") for _ in range(500))
    print("case: 3000-base overlap", flush=True)
    w, ce, cs = stitch.removeOverlap(pre, suf, 3000, pp)
    wo, ceo, cso = _oracle_cut(pre, suf, 3000, op)
    assert (ce, cs) == (ceo, cso) and abs(w - wo) <= 1, ((ce, cs, w), (ceo, cso, wo))


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
    op = O.PolishParams(PARAMS)
    merged = stitch.mergeContigChunks(chunks, 2 * bnd, pp, "NNNNN")
    assert merged == _merge_reference(chunks, 2 * bnd, pp, "NNNNN")
    assert merged == RP.merge_contig_chunks(chunks, 2 * bnd, op, "NNNNN")
    # first chunk empty (no overlap to find: a "NNNNNNNNNN" stitch), overlap 0
    odd = ["", chunks[1], chunks[2]]
    assert stitch.mergeContigChunks(odd, 2 * bnd, pp, "") == _merge_reference(odd, 2 * bnd, pp, "")
    assert stitch.mergeContigChunks(odd, 2 * bnd, pp, "") == RP.merge_contig_chunks(odd, 2 * bnd, op, "")
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
