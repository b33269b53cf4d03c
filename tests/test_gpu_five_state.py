"""GPU parity of the five-state machine (M, short X, short Y, long X, long Y;
no counterpart in the reference) on its per-cell-exponent kernel, through the C ABI,
against the FP64 oracle O_exact's five-state recursion (itself pinned by
path enumeration, the reduction to the 3-state machine and the per-x / per-y
posterior sums).

Bar: every posterior within 1e-4 absolute of O_exact; an entry present on one
side only must lie within 1e-4 of the threshold; reference output order for
the entries both sides hold; log-likelihoods within 1e-3 relative (1e-5 in
forward-only mode).
"""
import random
import sys

import numpy as np
import pytest

from oracle import brute
from oracle import oracle as O
from tests.conftest import PARAMS
from tests.parity import as_rows, compare, order_consistent

pytestmark = pytest.mark.gpu

TOL = 1e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ph():
    from marginpolish_amd import pairhmm
    return pairhmm


def _oracle_pair(om3_rev, om3_fwd, T5):
    mk = lambda m: O.Model(m.trans, m.match, m.gapX, m.gapY, m.rle, m.strand, m.repeat, trans5=T5)
    return mk(om3_rev), mk(om3_fwd)


@pytest.fixture(scope="module")
def five(ph):
    """(product machines, (oracle reverse, oracle forward)) for the RLE params
    with default long gaps, and for the default nucleotide machine with
    frequent long gaps (so that they carry posterior mass on short inputs)."""
    pp = ph.PolishParams(PARAMS)
    op = O.PolishParams(PARAMS)
    pp.machines.add_long_gaps()
    rle = (pp.machines, _oracle_pair(op.reverse, op.forward, pp.machines.transitions5()), pp.p)
    dm = ph.StateMachines.default_nucleotide()
    dm.add_long_gaps(0.05, 0.8)
    odm = O.default_nucleotide_model()
    dflt = (dm, _oracle_pair(odm, odm, dm.transitions5()), None)
    return rle, dflt


def _bp(ph, op):
    return ph.PairwiseAlignmentParameters(op.threshold, op.minDiagsBetweenTraceBack, op.traceBackDiagonals,
                                          op.diagonalExpansion, op.splitMatrixBiggerThanThis,
                                          op.dynamicAnchorExpansion)


def _rand_syms(rng, n, rle):
    bases = "".join(rng.choice("ACGT") for _ in range(n))
    if not rle:
        return O.symbols(bases)
    return O.rle_symbols(bases, [rng.choice([1, 1, 2, 3, 4, 7, 12]) for _ in range(n)])


def _check(gres, jobs, models, opar, check_order=True, lp_rel=1e-3):
    worst = 0.0
    for (sx, sy, an, strand, *rg), (gm, gx, gy, glp) in zip(jobs, gres):
        rl = rg[0] if rg else False
        rr = rg[1] if len(rg) > 1 else False
        om, ox, oy, olp, _ = O.posteriors(models[strand], sx, sy, an, opar, rl, rr, mode=O.LOGADD_EXACT)
        for g, o in ((gm, om), (gx, ox), (gy, oy)):
            g = as_rows(g)
            err, _, _ = compare(g, o, opar.threshold, TOL)
            assert err <= TOL + 1e-7, err
            worst = max(worst, err)
            if check_order:
                assert order_consistent(g, o)
        assert abs(glp - olp) <= lp_rel * max(1.0, abs(olp)), (glp, olp)
    return worst


def test_five_state_brute_force_tiny(ph, five):
    """All cell-states (threshold 0) vs exact path enumeration, both machines,
    both strands, four ragged combinations, one batched launch each."""
    rng = random.Random(202)
    p = ph.PairwiseAlignmentParameters(threshold=0.0, diagonalExpansion=20, dynamicAnchorExpansion=True)
    for mach, models, _ in five:
        cases = []
        for i in range(40):
            rl, rr, strand = bool(i & 1), bool(i & 2), (i >> 2) & 1
            lX, lY = rng.randint(1, 5), rng.randint(1, 5)
            om = models[strand]
            cases.append((_rand_syms(rng, lX, om.rle), _rand_syms(rng, lY, om.rle), None, strand, rl, rr))
        say("case: five-state brute force, rle=%s, %d pairs" % (models[1].rle, len(cases)))
        b = ph.Batch(mach, cases, p)
        b.run()
        for (sx, sy, _, strand, rl, rr), (gm, gx, gy, lp) in zip(cases, b.fetch()):
            total, post = brute.enumerate_paths(models[strand], sx, sy, rl, rr)
            assert abs(lp - total) < 1e-4 * max(1.0, abs(total))
            wants = (post[0], post[1] + post[3], post[2] + post[4])
            for st, lst in enumerate((gm, gx, gy)):
                got = np.zeros_like(post[0])
                for r in lst:
                    got[r["x"] + 1, r["y"] + 1] = r["weight"] / 1e7
                want = np.minimum(wants[st], 1.0)
                if st == 0:
                    want[0, :] = 0; want[:, 0] = 0
                elif st == 1:
                    want[0, :] = 0
                else:
                    want[:, 0] = 0
                assert np.max(np.abs(got - want)) <= TOL + 1e-7, (st, np.max(np.abs(got - want)))


def test_five_state_config0_like_parity(ph, five):
    """Reads with 10% errors against their reference, W = 32, both strands."""
    from marginpolish_amd import crop, sim
    mach, models, bp = five[0]
    w = sim.config0(n_reads=8, ref_len=3000, seed=41, strands=(1, 0))
    jobs = []
    for pr in w.pairs:
        sx, sy, an, _ = crop.crop_pair(pr.ref_sym, pr.read_sym, pr.anchors)
        jobs.append((sx, sy, an, pr.strand))
    opar = O.PolishParams(PARAMS).p
    say("case: five-state config0-like, 8 reads x 3 kb")
    b = ph.Batch(mach, jobs, bp)
    b.run()
    worst = _check(b.fetch(), jobs, models, opar)
    say("five-state max-abs posterior err vs O_exact: %.3g" % worst)


def test_five_state_windows_splits_ragged_wide(ph, five):
    """Several traceback windows, large-gap splits, ragged ends and band
    widths from 12 to ~520 cells (1 to 9 wavefront columns), posterior and
    match-only modes, batched and one launch per job."""
    from marginpolish_amd import crop, sim
    mach, models, _ = five[0]
    jobs = []
    for k, W in enumerate((10, 80, 200, 512)):
        w = sim.config0(n_reads=2, ref_len=1500, seed=50 + k, W=W, strands=(1, 0))
        for i, pr in enumerate(w.pairs):
            sx, sy, an, _ = crop.crop_pair(pr.ref_sym, pr.read_sym, pr.anchors)
            an = an[::5]
            if i == 1:  # an anchor-free stretch of ~600 x 600 cells: split into ragged-ended blocks
                an = an[(an[:, 0] < 400) | (an[:, 0] >= 1000)]
            jobs.append((sx, sy, an, pr.strand, bool(i & 1), bool(k & 1)))
    opar = O.Params(threshold=0.01, minDiagsBetweenTraceBack=400, traceBackDiagonals=40, diagonalExpansion=10,
                    splitMatrixBiggerThanThis=90000, dynamicAnchorExpansion=True)
    say("case: five-state windows/splits/ragged, widths up to ~520, %d pairs" % len(jobs))
    b = ph.Batch(mach, jobs, _bp(ph, opar))
    b.run()
    g = b.fetch()
    _check(g, jobs, models, opar)
    say("case: five-state one launch per job")
    b2 = ph.Batch(mach, jobs, _bp(ph, opar)).run_plain()
    for r1, r2 in zip(g, b2.fetch()):
        for l in range(3):
            assert np.array_equal(r1[l], r2[l])
        assert r1[3] == r2[3]
    say("case: five-state match-only")
    b3 = ph.Batch(mach, jobs, _bp(ph, opar), mode=1).run()
    for r1, r3 in zip(g, b3.fetch()):
        assert np.array_equal(r1[0], r3[0]) and len(r3[1]) == 0 and len(r3[2]) == 0


def test_five_state_structural_indels(ph, five):
    """A read with a 150-base deletion and a 120-base insertion against its
    reference (the events the long-gap states model), anchors away from the
    events, band 200: posteriors vs O_exact, and the long gaps raise the
    likelihood over the 3-state machine."""
    mach, models, _ = five[0]
    rng = np.random.default_rng(8)
    ref = "".join(rng.choice(list("ACGT"), 1400))
    ins = "".join(rng.choice(list("ACGT"), 120))
    read = ref[:400] + ref[550:900] + ins + ref[900:]
    sx, sy = O.rle_symbols(ref, [1] * len(ref)), O.rle_symbols(read, [1] * len(read))
    anchors = [(x, x, 200) for x in range(0, 400, 25)] + [(x, x - 150, 200) for x in range(560, 900, 25)] + \
        [(x, x - 30, 200) for x in range(910, 1400, 25)]
    opar = O.Params(threshold=0.01, minDiagsBetweenTraceBack=100000, traceBackDiagonals=40, diagonalExpansion=200,
                    splitMatrixBiggerThanThis=250_000_000, dynamicAnchorExpansion=True)
    jobs = [(sx, sy, np.array(anchors, np.int64), 1)]
    say("case: five-state structural indels, 1.4 kb, W=200")
    b = ph.Batch(mach, jobs, _bp(ph, opar))
    b.run()
    g = b.fetch()
    _check(g, jobs, models, opar)
    three = ph.StateMachines.from_json(PARAMS)
    b3 = ph.Batch(three, jobs, _bp(ph, opar)).run()
    assert g[0][3] > b3.fetch()[0][3] + 10.0


def test_five_state_forward_only(ph, five):
    """computeForwardProbability with the five-state machine: tiny pairs vs
    enumeration, config0-like pairs vs O_exact."""
    mach, models, _ = five[1]
    rng = random.Random(77)
    cases = []
    for i in range(24):
        lX, lY = rng.randint(1, 5), rng.randint(1, 5)
        cases.append((_rand_syms(rng, lX, False), _rand_syms(rng, lY, False), None, 1, bool(i & 1), bool(i & 2)))
    say("case: five-state forward-only tiny x24")
    b = ph.Batch(mach, cases, ph.PairwiseAlignmentParameters(diagonalExpansion=4), mode=2)
    b.run()
    _, lps = b.counts()
    for (sx, sy, _, _, rl, rr), lp in zip(cases, lps):
        total, _ = brute.enumerate_paths(models[1], sx, sy, rl, rr)
        assert abs(lp - total) < 1e-4 * max(1.0, abs(total)), (lp, total)
    from marginpolish_amd import crop, sim
    mach, models, _ = five[0]
    w = sim.config0(n_reads=6, ref_len=1500, seed=19, W=8, strands=(1, 0))
    jobs = []
    for pr in w.pairs:
        sx, sy, an, _ = crop.crop_pair(pr.ref_sym, pr.read_sym, pr.anchors)
        jobs.append((sx, sy, an, pr.strand))
    say("case: five-state forward-only config0-like x6")
    p2 = ph.PairwiseAlignmentParameters(diagonalExpansion=8, dynamicAnchorExpansion=True)
    o2 = O.Params(diagonalExpansion=8, dynamicAnchorExpansion=True)
    b = ph.Batch(mach, jobs, p2, mode=2)
    b.run()
    _, lps = b.counts()
    for (sx, sy, an, strand), lp in zip(jobs, lps):
        want = O.forward_prob(models[strand], sx, sy, an, o2)
        assert abs(lp - want) < 1e-5 * abs(want), (lp, want)


def test_five_state_refusals(ph, five):
    """Expected counts are 3-state only; bands wider than 2048 cells are
    refused; a batch keeps its state count across set_model."""
    mach = five[1][0]
    p = ph.PairwiseAlignmentParameters()
    rng = random.Random(4)
    sx = _rand_syms(rng, 50, False)
    with pytest.raises(ph.MPError) as ei:
        ph.Batch(mach, [(sx, sx, None, 1)], p, mode=3)
    assert ei.value.status == 6
    big = _rand_syms(rng, 2100, False)
    with pytest.raises(ph.MPError) as ei:
        ph.getAlignedPairs(mach, 1, big, big, p)
    assert ei.value.status == 6
    b = ph.Batch(mach, [(sx, sx, None, 1)], p)
    with pytest.raises(ph.MPError):
        b.set_model(ph.StateMachines.default_nucleotide())


def test_fp64_paths(ph, five):
    """FP64 batches (the per-cell-exponent kernel in double): the five-state
    machine and the 3-state machine (long gaps unreachable) against O_exact
    at 1e-6 per posterior (the emission tables stay FP32: ~6e-8 relative per
    factor) and 1e-5 relative per log-likelihood; bands wider than 1216 cells
    are refused."""
    from marginpolish_amd import crop, sim
    mach5, models5, _ = five[0]
    pp = ph.PolishParams(PARAMS)
    op = O.PolishParams(PARAMS)
    jobs = []
    for k, W in enumerate((10, 200, 512)):
        w = sim.config0(n_reads=2, ref_len=1200, seed=90 + k, W=W, strands=(1, 0))
        for i, pr in enumerate(w.pairs):
            sx, sy, an, _ = crop.crop_pair(pr.ref_sym, pr.read_sym, pr.anchors)
            jobs.append((sx, sy, an[::5], pr.strand, bool(i & 1), bool(k & 1)))
    opar = O.Params(threshold=0.01, minDiagsBetweenTraceBack=500, traceBackDiagonals=40, diagonalExpansion=10,
                    splitMatrixBiggerThanThis=250_000_000, dynamicAnchorExpansion=True)
    for tag, mach, models in (("five-state", mach5, models5), ("three-state", pp.machines, (op.reverse, op.forward))):
        say("case: FP64 %s, %d pairs" % (tag, len(jobs)))
        b = ph.Batch(mach, jobs, _bp(ph, opar), fp64=True)
        b.run()
        worst = 0.0
        for (sx, sy, an, strand, rl, rr), (gm, gx, gy, glp) in zip(jobs, b.fetch()):
            om, ox, oy, olp, _ = O.posteriors(models[strand], sx, sy, an, opar, rl, rr, mode=O.LOGADD_EXACT)
            for g, o in ((gm, om), (gx, ox), (gy, oy)):
                err, _, _ = compare(as_rows(g), o, opar.threshold, 1e-6)
                worst = max(worst, err)
                assert order_consistent(as_rows(g), o)
            assert abs(glp - olp) <= 1e-5 * abs(olp), (glp, olp)
        assert worst <= 1e-6, worst
        say("FP64 %s max-abs posterior err vs O_exact: %.3g" % (tag, worst))
    rng = random.Random(6)
    big = _rand_syms(rng, 1300, False)
    with pytest.raises(ph.MPError) as ei:
        ph.Batch(ph.StateMachines.default_nucleotide(), [(big, big, None, 1)], ph.PairwiseAlignmentParameters(),
                 fp64=True)
    assert ei.value.status == 6
