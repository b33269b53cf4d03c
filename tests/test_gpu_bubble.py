"""Bubble-graph per-read x per-allele forward likelihoods
(mp_bubble_allele_supports) against the FP64 oracle's computeForwardProbability."""
import numpy as np
import pytest

from marginpolish_amd import pairhmm as ph
from marginpolish_amd import sim
from oracle import oracle as O
from tests.conftest import PARAMS, PARAMS_NO_RLE

pytestmark = pytest.mark.gpu

_CODE = {ord("A"): 0, ord("C"): 1, ord("G"): 2, ord("T"): 3, ord("a"): 0, ord("c"): 1, ord("g"): 2, ord("t"): 3}


def _sym(r, rl=True, cap=51):
    b, c = r
    b = np.frombuffer(b, np.uint8) if isinstance(b, bytes) else np.asarray(b, np.uint8)
    codes = np.array([_CODE.get(int(x), 4) for x in b], dtype=np.uint16)
    if not rl:
        return codes
    return ((np.minimum(np.asarray(c, np.uint64), cap - 1).astype(np.uint16) << 8) | codes).astype(np.uint16)


def _expected(op, opar, alleles, reads, rl=True):
    """bubbleGraph.c:905-928 restated: the first read with an equal string is
    computed (with its own strand), later equal ones copy it."""
    first = {}
    out = np.zeros((len(alleles), len(reads)))
    for k, (r, strand) in enumerate(reads):
        key = (bytes(r[0]), tuple(int(x) for x in r[1]))
        if key in first:
            out[:, k] = out[:, first[key]]
            continue
        first[key] = k
        m = op.forward if strand else op.reverse
        for j, a in enumerate(alleles):
            out[j, k] = O.forward_prob(m, _sym(a, rl), _sym(r, rl), [], opar)
    return out


def test_bubble_supports_match_oracle():
    pp = ph.PolishParams(PARAMS)
    op = O.PolishParams(PARAMS)
    bub = sim.bubbles(n_bubbles=120, depth=12, seed=21)
    print("bubbles: %d" % len(bub), flush=True)
    got, npairs = ph.bubble_allele_supports(pp.machines, pp.p, bub)
    want_pairs = 0
    worst = 0.0
    for (alleles, reads), g in zip(bub, got):
        w = _expected(op, op.p, alleles, reads)
        assert g.shape == w.shape
        worst = max(worst, float(np.max(np.abs(g - w) / np.maximum(1.0, np.abs(w)))))
        want_pairs += len(alleles) * len({(bytes(r[0]), tuple(int(x) for x in r[1])) for r, _ in reads})
    assert worst < 2e-6, worst
    assert npairs == want_pairs


def test_repeated_read_takes_first_strand():
    """Two reads with the same string but different strands: the second
    copies the first's values (computed with the first read's machine)."""
    pp = ph.PolishParams(PARAMS)
    op = O.PolishParams(PARAMS)
    alleles = [(b"ACGTAC", np.array([1, 2, 1, 3, 1, 1], np.uint64)), (b"ACGAC", np.array([1, 2, 1, 1, 1], np.uint64))]
    r = (b"ACGTAC", np.array([1, 2, 1, 2, 1, 1], np.uint64))
    reads = [(r, 0), ((b"CGTA", np.array([2, 1, 3, 1], np.uint64)), 1), ((r[0], r[1].copy()), 1)]
    (g,), npairs = ph.bubble_allele_supports(pp.machines, pp.p, [(alleles, reads)])
    assert npairs == 4
    assert np.array_equal(g[:, 2], g[:, 0])
    w = _expected(op, op.p, alleles, reads)
    assert np.max(np.abs(g - w) / np.abs(w)) < 2e-6
    # the forward machine would give a different value for read 2
    fwd = O.forward_prob(op.forward, _sym(alleles[0]), _sym(r), [], op.p)
    assert abs(fwd - g[0, 2]) > 1e-3


def test_no_rle_and_edge_cases():
    pp = ph.PolishParams(PARAMS_NO_RLE)
    op = O.PolishParams(PARAMS_NO_RLE)
    bub = sim.bubbles(n_bubbles=20, depth=6, seed=3)
    bub.append(([bub[0][0][0]], []))  # a bubble without reads
    got, _ = ph.bubble_allele_supports(pp.machines, pp.p, bub, include_repeat_counts=False)
    for (alleles, reads), g in zip(bub, got):
        w = _expected(op, op.p, alleles, reads, rl=False)
        assert g.shape == w.shape
        if g.size:
            assert np.max(np.abs(g - w) / np.maximum(1.0, np.abs(w))) < 2e-6
    got, n = ph.bubble_allele_supports(pp.machines, pp.p, [])
    assert got == [] and n == 0
    with pytest.raises(ph.MPError):
        ph.bubble_allele_supports(pp.machines, pp.p, bub, max_repeat_count_exclusive=0)
