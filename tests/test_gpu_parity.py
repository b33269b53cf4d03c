"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle.

Bar (FP32 scaled arithmetic vs FP64): every posterior within 1e-4 absolute of
O_exact (exact log-sum-exp, same bands, windows and splits); an entry may be
present on one side only when its posterior is within 1e-4 of the threshold.
Output order equals the reference order for the entries both sides hold.
"""
import math
import os
import random
import sys

import numpy as np
import pytest

from oracle import brute
from oracle import oracle as O
from tests.conftest import PARAMS
from tests.parity import as_map, as_rows, compare, order_consistent

pytestmark = pytest.mark.gpu

TOL = 1e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ph():
    from marginpolish_amd import pairhmm
    return pairhmm


@pytest.fixture(scope="module")
def params(ph):
    return ph.PolishParams(PARAMS), O.PolishParams(PARAMS)


def _bp(ph, op):
    return ph.PairwiseAlignmentParameters(op.threshold, op.minDiagsBetweenTraceBack, op.traceBackDiagonals,
                                          op.diagonalExpansion, op.splitMatrixBiggerThanThis,
                                          op.dynamicAnchorExpansion)


def test_known_answer_agcg(ph):
    say("case: AGCG/AGTTCG default machine, threshold 0.2")
    m = ph.StateMachines.default_nucleotide()
    p = ph.PairwiseAlignmentParameters(threshold=0.2, diagonalExpansion=2)
    mm = ph.getAlignedPairsUsingAnchors(m, 1, O.symbols("AGCG"), O.symbols("AGTTCG"), None, p)
    assert {(int(r["x"]), int(r["y"])) for r in mm} == {(0, 0), (1, 1), (2, 4), (3, 5)}


def _rand_syms(rng, n, rle):
    bases = "".join(rng.choice("ACGT") for _ in range(n))
    if not rle:
        return O.symbols(bases)
    return O.rle_symbols(bases, [rng.choice([1, 1, 2, 3, 4, 7, 12]) for _ in range(n)])


def test_brute_force_tiny(ph, params):
    """All cell-states (threshold 0) vs exact path enumeration, 3 machines x 4
    ragged combinations, in one batched launch."""
    pp, op = params
    dm = ph.StateMachines.default_nucleotide()
    odm = O.default_nucleotide_model()
    rng = random.Random(101)
    p = ph.PairwiseAlignmentParameters(threshold=0.0, diagonalExpansion=20, dynamicAnchorExpansion=True)
    for mach, models in ((dm, (odm, odm)), (pp.machines, (op.reverse, op.forward))):
        cases = []
        for i in range(40):
            rl, rr, strand = bool(i & 1), bool(i & 2), (i >> 2) & 1
            lX, lY = rng.randint(1, 5), rng.randint(1, 5)
            om = models[strand]
            cases.append((_rand_syms(rng, lX, om.rle), _rand_syms(rng, lY, om.rle), None, strand, rl, rr))
        say("case: brute force, rle=%s, %d pairs" % (models[1].rle, len(cases)))
        b = ph.Batch(mach, cases, p)
        b.run()
        res = b.fetch()
        for (sx, sy, _, strand, rl, rr), (gm, gx, gy, lp) in zip(cases, res):
            total, post = brute.enumerate_paths(models[strand], sx, sy, rl, rr)
            assert abs(lp - total) < 1e-4 * max(1.0, abs(total))
            for st, lst in ((0, gm), (1, gx), (2, gy)):
                got = np.zeros_like(post[st])
                for r in lst:
                    got[r["x"] + 1, r["y"] + 1] = r["weight"] / 1e7
                want = np.minimum(post[st], 1.0)
                if st == 0:
                    want[0, :] = 0; want[:, 0] = 0
                elif st == 1:
                    want[0, :] = 0
                else:
                    want[:, 0] = 0
                assert np.max(np.abs(got - want)) <= TOL + 1e-7, (st, np.max(np.abs(got - want)))


def _oracle_jobs(op, jobs, oparams, mode=O.LOGADD_EXACT):
    out = []
    for (sx, sy, an, strand, *rg) in jobs:
        rl = rg[0] if rg else False
        rr = rg[1] if len(rg) > 1 else False
        m = op.forward if strand else op.reverse
        out.append(O.posteriors(m, sx, sy, an, oparams, rl, rr, mode=mode))
    return out


# Log-likelihood bar: FP32 rounding compounds along the anti-diagonals at
# about 1e-7 relative per step, i.e. ~1e-3 nats over 1e4 anti-diagonals; the
# bar is 1e-2 nats absolute.
LL_TOL = 1e-2
LL_DRIFT = {"max": 0.0}


def _check(gres, ores, thr, check_order=True):
    worst = 0.0
    for (gm, gx, gy, glp), (om, ox, oy, olp, _) in zip(gres, ores):
        for g, o in ((gm, om), (gx, ox), (gy, oy)):
            g = as_rows(g)
            err, ncom, _ = compare(g, o, thr, TOL)
            assert err <= TOL + 1e-7, err
            worst = max(worst, err)
            if check_order:
                assert order_consistent(g, o)
        assert abs(glp - olp) <= LL_TOL, (glp, olp)
        LL_DRIFT["max"] = max(LL_DRIFT["max"], abs(glp - olp))
    say("max log-likelihood drift vs O_exact so far: %.3g nats" % LL_DRIFT["max"])
    return worst


def test_config0_parity(ph, params):
    from marginpolish_amd import crop, sim
    pp, op = params
    w = sim.config0(strands=(1, 0))
    jobs = []
    for pr in w.pairs:
        sx, sy, an, _ = crop.crop_pair(pr.ref_sym, pr.read_sym, pr.anchors)
        jobs.append((sx, sy, an, pr.strand))
    say("case: config0 20 reads x 5 kb, W=32, both strands")
    b = ph.Batch(pp.machines, jobs, pp.p)
    b.run()
    g = b.fetch()
    worst = _check(g, _oracle_jobs(op, jobs, op.p), op.p.threshold)
    say("config0 max-abs posterior err vs O_exact: %.3g" % worst)
    # O_faithful differs from O_exact by the reference's own logAdd error
    of = _oracle_jobs(op, jobs[:3], op.p, O.LOGADD_POLY)
    errs = [compare(as_rows(gg[0]), oo[0], op.p.threshold, 1.0)[0] for gg, oo in zip(g[:3], of)]
    say("config0 max-abs posterior diff vs O_faithful: %.3g" % max(errs))


def test_multi_window_parity(ph, params):
    """Several traceback windows per pair (minDiagsBetweenTraceBack 300)."""
    from marginpolish_amd import crop, sim
    pp, op = params
    w = sim.config0(n_reads=6, ref_len=2000, seed=7, W=10, strands=(1, 0))
    jobs = []
    for pr in w.pairs:
        sx, sy, an, _ = crop.crop_pair(pr.ref_sym, pr.read_sym, pr.anchors)
        jobs.append((sx, sy, an, pr.strand))
    opar = O.Params(threshold=0.01, minDiagsBetweenTraceBack=300, traceBackDiagonals=40, diagonalExpansion=10,
                    splitMatrixBiggerThanThis=250_000_000, dynamicAnchorExpansion=True)
    say("case: multi-window, 6 pairs")
    b = ph.Batch(pp.machines, jobs, _bp(ph, opar))
    b.run()
    _check(b.fetch(), _oracle_jobs(op, jobs, opar), opar.threshold)


def test_split_and_ragged_parity(ph, params):
    """Large-gap splits (tiny splitMatrixBiggerThanThis) with ragged ends."""
    pp, op = params
    rng = random.Random(5)
    jobs = []
    for i in range(8):
        n = rng.randint(60, 200)
        sx = _rand_syms(rng, n, True)
        sy = sx.copy()
        keep = [j for j in range(n) if rng.random() > 0.05]
        sy = sx[keep]
        anchors = []
        for k in range(0, len(keep), rng.randint(15, 40)):
            if not anchors or (keep[k] > anchors[-1][0] and k > anchors[-1][1]):
                anchors.append((keep[k], k, 6))
        jobs.append((sx, sy, np.array(anchors, np.int64), i & 1, bool(i & 2), bool(i & 4)))
    opar = O.Params(threshold=0.01, minDiagsBetweenTraceBack=1000, traceBackDiagonals=40, diagonalExpansion=6,
                    splitMatrixBiggerThanThis=900, dynamicAnchorExpansion=True)
    say("case: splits + ragged, 8 pairs")
    b = ph.Batch(pp.machines, jobs, _bp(ph, opar))
    b.run()
    _check(b.fetch(), _oracle_jobs(op, jobs, opar), opar.threshold)


@pytest.fixture(scope="module")
def config1(params):
    from marginpolish_amd import sim
    return sim.config1()


def test_config1_subset_realign(ph, params, config1):
    """poa_realign caller semantics (reference cropping, gapX->deletes,
    gapY->inserts) on 300 chunk-level pairs of config 1, one batched launch."""
    from marginpolish_amd import crop
    pp, op = params
    rng = random.Random(3)
    sel = rng.sample(range(len(config1.pairs)), 300)
    pairs = [(config1.pairs[i].ref_sym, config1.pairs[i].read_sym, config1.pairs[i].anchors,
              config1.pairs[i].strand) for i in sel]
    say("case: config1 subset 300 pairs, realign_all")
    res = ph.realign_all(pp, pairs)
    worst = 0.0
    for (rs, qs, an, strand), (m, ins, dels) in zip(pairs, res):
        sx, sy, a2, off = crop.crop_pair(rs, qs, an)
        om, ox, oy, olp, _ = O.posteriors(op.forward if strand else op.reverse, sx, sy, a2, op.p)
        for g, o in ((m, om), (dels, ox), (ins, oy)):
            o = o.copy()
            o[:, 1] += off
            err, _, _ = compare(as_rows(g), o, op.p.threshold, TOL)
            worst = max(worst, err)
            assert err <= TOL + 1e-7
    say("config1 subset max-abs posterior err vs O_exact: %.3g" % worst)


def test_batch_equals_plain_and_match_only(ph, params, config1):
    from marginpolish_amd import crop
    pp, op = params
    jobs = []
    for pr in config1.pairs[:40]:
        sx, sy, an, _ = crop.crop_pair(pr.ref_sym, pr.read_sym, pr.anchors)
        jobs.append((sx, sy, an, pr.strand))
    say("case: batch vs plain, 40 pairs")
    b1 = ph.Batch(pp.machines, jobs, pp.p).run()
    b2 = ph.Batch(pp.machines, jobs, pp.p).run_plain()
    b3 = ph.Batch(pp.machines, jobs, pp.p, mode=1).run()
    for r1, r2, r3 in zip(b1.fetch(), b2.fetch(), b3.fetch()):
        for l in range(3):
            assert np.array_equal(r1[l], r2[l])
        assert np.array_equal(r1[0], r3[0]) and len(r3[1]) == 0 and len(r3[2]) == 0
        assert r1[3] == r2[3]


def test_config1_full_batch_properties(ph, params, config1):
    """Every pair of config 1 in one launch; size-independent checks."""
    pp, op = params
    pairs = [(p.ref_sym, p.read_sym, p.anchors, p.strand) for p in config1.pairs]
    say("case: config1 full batch %d pairs" % len(pairs))
    b = ph.Batch(pp.machines, pairs, pp.p, crop_reference=True)
    b.run()
    counts, lp = b.counts()
    assert np.all(np.isfinite(lp)) and np.all(lp < 0)
    res = b.fetch()
    for (rs, qs, an, _), (m, gx, gy, _) in zip(pairs[::50], res[::50]):
        mx = np.zeros(len(rs) + 1)
        for arr in (m, gx):
            assert np.all(arr["weight"] > 0) and np.all(arr["weight"] <= 10 ** 7)
            assert np.all(arr["x"] < len(rs)) and np.all(arr["y"] < len(qs))
            np.add.at(mx, arr["x"], arr["weight"])
        assert np.all(mx <= 1e7 * (1 + 1e-3) + 16)
        assert len(as_map(as_rows(m))) == len(m)


def test_long_pair_lists_in_wide_entries(ph, params):
    """A batch whose longest reference and read need more than the 40 bits of
    a 64-bit list entry keeps 16-byte entries: the config 0 pairs' lists in
    such a batch equal their lists in a batch of their own (64-bit entries),
    and the long pair's reference coordinates reach past 2^25."""
    from marginpolish_amd import crop, sim
    pp, op = params
    w = sim.config0(strands=(1, 0))
    jobs = []
    for pr in w.pairs[:4]:
        sx, sy, an, _ = crop.crop_pair(pr.ref_sym, pr.read_sym, pr.anchors)
        jobs.append((sx, sy, an, pr.strand))
    # the long pair: a 64 Mb reference, a 40 kb read copied from near its end, anchors on
    # that diagonal (27 + 16 bits of coordinates; one sub-problem of ~80 k anti-diagonals)
    rng = np.random.default_rng(3)
    lX, lY = (1 << 26) + 5, 40000
    xs = lX - lY - 1000
    ref = (rng.integers(0, 4, lX) | (1 << 8)).astype(np.uint16)
    read = ref[xs:xs + lY].copy()
    k = np.arange(50, lY - 50, 100, dtype=np.int64)
    anchors = np.stack([xs + k, k, np.full_like(k, 20)], axis=1).reshape(-1)
    say("case: 4 config0 pairs with a 64 Mb x 40 kb pair (16-byte list entries) vs alone")
    big = ph.Batch(pp.machines, jobs + [(ref, read, anchors, 1)], pp.p, crop_reference=True)
    big.run()
    gb = big.fetch()
    alone = ph.Batch(pp.machines, jobs, pp.p, crop_reference=True)
    alone.run()
    ga = alone.fetch()
    for (m1, x1, y1, l1), (m2, x2, y2, l2) in zip(gb[:4], ga):
        for a1, a2 in ((m1, m2), (x1, x2), (y1, y2)):
            assert np.array_equal(a1, a2)
        assert l1 == l2
    m, gx, gy, lp = gb[4]
    assert len(m) > lY // 2 and np.isfinite(lp)
    assert m["x"].min() >= xs - 1 and m["x"].max() >= (1 << 25) and m["x"].max() < lX and m["y"].max() < lY
    assert np.all(m["weight"] > 0) and np.all(m["weight"] <= 10 ** 7)


def test_too_wide_band_rejected_loudly(ph):
    m = ph.StateMachines.default_nucleotide()
    p = ph.PairwiseAlignmentParameters()
    rng = random.Random(1)
    sx = _rand_syms(rng, 2500, False)
    with pytest.raises(ph.MPError) as ei:
        ph.getAlignedPairs(m, 1, sx, sx, p)
    assert ei.value.status == 6


def _read_fasta(path):
    seqs, name = [], None
    for line in open(path):
        line = line.strip()
        if line.startswith(">"):
            name = line[1:]
            seqs.append([name, ""])
        elif line:
            seqs[-1][1] += line
    return seqs


def test_wide_band_anchor_free_ecoli(ph, params):
    """Anchor-free realignment (getAlignedPairsWithIndels, RL cap 49 as in
    poa_realign's anchor-free branch, polisher.c:676-683) of real ONT reads to
    their 100 bp truth windows: full-rectangle bands of 50-110 cells, run by
    the wide-band kernel, vs O_exact."""
    import glob
    pp, op = params
    root = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ecoli_100bp")
    jobs = []
    for ref in sorted(glob.glob(os.path.join(root, "*.ref.fasta"))):
        truth = _read_fasta(ref)[0][1]
        rb, rr = O.rle_encode(truth)
        sx = O.rle_symbols(rb, rr, 50)
        for name, read in _read_fasta(ref.replace(".ref.fasta", ".fasta"))[:12]:
            qb, qr = O.rle_encode(read)
            jobs.append((sx, O.rle_symbols(qb, qr, 50), None, 1 if name.endswith("_F") else 0))
    say("case: ecoli anchor-free x%d (wide kernel)" % len(jobs))
    b = ph.Batch(pp.machines, jobs, pp.p)
    b.run()
    worst = _check(b.fetch(), _oracle_jobs(op, jobs, op.p), op.p.threshold)
    say("ecoli anchor-free max-abs posterior err: %.3g" % worst)


def test_wide_band_windows_modes(ph, params):
    """Wide bands (W=80) with several traceback windows, splits and ragged
    ends, in posterior, forward-only and expectation modes, plus batched vs
    plain identity."""
    from marginpolish_amd import crop, sim
    pp, op = params
    w = sim.config0(n_reads=4, ref_len=1500, seed=17, W=80, strands=(1, 0))
    jobs = []
    for i, pr in enumerate(w.pairs):
        sx, sy, an, _ = crop.crop_pair(pr.ref_sym, pr.read_sym, pr.anchors)
        jobs.append((sx, sy, an[::7], pr.strand, bool(i & 1), bool(i & 2)))
    opar = O.Params(threshold=0.01, minDiagsBetweenTraceBack=500, traceBackDiagonals=40, diagonalExpansion=80,
                    splitMatrixBiggerThanThis=40000, dynamicAnchorExpansion=True)
    say("case: wide windows/splits x4")
    b = ph.Batch(pp.machines, jobs, _bp(ph, opar))
    b.run()
    g = b.fetch()
    _check(g, _oracle_jobs(op, jobs, opar), opar.threshold)
    b2 = ph.Batch(pp.machines, jobs, _bp(ph, opar)).run_plain()
    for r1, r2 in zip(g, b2.fetch()):
        for l in range(3):
            assert np.array_equal(r1[l], r2[l])
    be = ph.Batch(pp.machines, jobs, _bp(ph, opar), mode=3)
    be.run()
    t, lk = be.expectations()
    for (sx, sy, an, strand, rl, rr), tg, lg in zip(jobs, t, lk):
        te, le = O.expectations(op.forward if strand else op.reverse, sx, sy, an, opar, rl, rr)
        assert np.allclose(tg, te, rtol=1e-4, atol=1e-3), (tg, te)
        assert abs(lg - le) < 1e-6 * abs(le)
    bf = ph.Batch(pp.machines, jobs, _bp(ph, opar), mode=2)
    bf.run()
    _, lps = bf.counts()
    for (sx, sy, an, strand, rl, rr), lp in zip(jobs, lps):
        want = O.forward_prob(op.forward if strand else op.reverse, sx, sy, an, opar, rl, rr)
        assert abs(lp - want) < 1e-5 * abs(want), (lp, want)


def test_forward_log_prob_parity(ph, params):
    """computeForwardProbability: static band of diagonalExpansion, no split,
    vs O_exact and vs exact enumeration on tiny pairs."""
    pp, op = params
    rng = random.Random(31)
    dm = ph.StateMachines.default_nucleotide()
    odm = O.default_nucleotide_model()
    p = ph.PairwiseAlignmentParameters(diagonalExpansion=4)
    opar = O.Params(diagonalExpansion=4)
    cases = []
    for i in range(24):
        lX, lY = rng.randint(1, 5), rng.randint(1, 5)
        cases.append((_rand_syms(rng, lX, False), _rand_syms(rng, lY, False), None, 1, bool(i & 1), bool(i & 2)))
    say("case: forward-only tiny x24")
    b = ph.Batch(dm, cases, p, mode=2)
    b.run()
    _, lps = b.counts()
    for (sx, sy, _, _, rl, rr), lp in zip(cases, lps):
        total, _ = brute.enumerate_paths(odm, sx, sy, rl, rr)
        assert abs(lp - total) < 1e-4 * max(1.0, abs(total)), (lp, total)
    from marginpolish_amd import crop, sim
    w = sim.config0(n_reads=6, ref_len=1500, seed=9, W=8, strands=(1, 0))
    jobs = []
    for pr in w.pairs:
        sx, sy, an, _ = crop.crop_pair(pr.ref_sym, pr.read_sym, pr.anchors)
        jobs.append((sx, sy, an, pr.strand))
    say("case: forward-only config0-like x6")
    p2 = ph.PairwiseAlignmentParameters(diagonalExpansion=8, dynamicAnchorExpansion=True)
    o2 = O.Params(diagonalExpansion=8, dynamicAnchorExpansion=True)
    b = ph.Batch(pp.machines, jobs, p2, mode=2)
    b.run()
    _, lps = b.counts()
    for (sx, sy, an, strand), lp in zip(jobs, lps):
        want = O.forward_prob(op.forward if strand else op.reverse, sx, sy, an, o2)
        assert abs(lp - want) < 1e-5 * abs(want), (lp, want)
    one = ph.computeForwardProbability(pp.machines, jobs[0][3], jobs[0][0], jobs[0][1], jobs[0][2], p2)
    assert abs(one - lps[0]) < 1e-9 * abs(lps[0])


def test_expectations_parity(ph, params):
    """getExpectationsUsingAnchors: expected transition counts and the
    per-diagonal likelihood sum vs O_exact (multi-window, splits, ragged)."""
    pp, op = params
    from marginpolish_amd import crop, sim
    w = sim.config0(n_reads=6, ref_len=1200, seed=13, W=10, strands=(1, 0))
    jobs = []
    for i, pr in enumerate(w.pairs):
        sx, sy, an, _ = crop.crop_pair(pr.ref_sym, pr.read_sym, pr.anchors)
        jobs.append((sx, sy, an, pr.strand, bool(i & 1), bool(i & 2)))
    opar = O.Params(threshold=0.01, minDiagsBetweenTraceBack=400, traceBackDiagonals=40, diagonalExpansion=10,
                    splitMatrixBiggerThanThis=250_000_000, dynamicAnchorExpansion=True)
    say("case: expectations x6 (multi-window)")
    b = ph.Batch(pp.machines, jobs, _bp(ph, opar), mode=3)
    b.run()
    t, lk = b.expectations()
    for (sx, sy, an, strand, rl, rr), tg, lg in zip(jobs, t, lk):
        te, le = O.expectations(op.forward if strand else op.reverse, sx, sy, an, opar, rl, rr)
        assert np.allclose(tg, te, rtol=1e-4, atol=1e-3), (tg, te)
        assert abs(lg - le) < 1e-6 * abs(le), (lg, le)
    # tiny pairs vs the oracle (pinned by enumeration in the oracle tests)
    dm = ph.StateMachines.default_nucleotide()
    odm = O.default_nucleotide_model()
    rng = random.Random(8)
    cases = [(_rand_syms(rng, rng.randint(1, 6), False), _rand_syms(rng, rng.randint(1, 6), False), None, 1)
             for _ in range(20)]
    p = ph.PairwiseAlignmentParameters(threshold=0.0, diagonalExpansion=20)
    b = ph.Batch(dm, cases, p, mode=3)
    b.run()
    t, _ = b.expectations()
    for (sx, sy, _, _), tg in zip(cases, t):
        te, _ = O.expectations(odm, sx, sy, [], O.Params(threshold=0.0, diagonalExpansion=20))
        assert np.allclose(tg, te, rtol=1e-4, atol=1e-5), (tg, te)


def test_graph_replays_equal_eager_runs(ph, params, config1):
    """The batch's hipGraph (counter-clearing fill kernels + launches) replayed
    several times, with eager batches run and synchronised between the
    replays, gives the eager result every time; re-capture after set_model."""
    from marginpolish_amd import crop
    pp, op = params
    jobs = []
    for pr in config1.pairs[:200]:
        sx, sy, an, _ = crop.crop_pair(pr.ref_sym, pr.read_sym, pr.anchors)
        jobs.append((sx, sy, an, pr.strand))
    say("case: graph replays vs eager, 200 pairs")
    eager = ph.Batch(pp.machines, jobs, pp.p, graph=False).run().fetch()
    g = ph.Batch(pp.machines, jobs, pp.p)
    for rep in range(4):
        g.run()
        got = g.fetch()
        other = ph.Batch(pp.machines, jobs[rep::4], pp.p, graph=False).run()  # eager work between replays
        other.fetch()
        for r1, r2 in zip(got, eager):
            for l in range(3):
                assert np.array_equal(r1[l], r2[l]), (rep, l)
            assert r1[3] == r2[3]
    g.set_model(pp.machines)
    g.run()
    for r1, r2 in zip(g.fetch(), eager):
        assert np.array_equal(r1[0], r2[0])


def test_autotune_keeps_results(ph, params, config1):
    """The autotuner changes only the launch geometry: after it the batch's
    lists equal the untuned ones, and a new batch of the same shape starts
    from the cached choice."""
    from marginpolish_amd import crop
    pp, op = params
    jobs = []
    for pr in config1.pairs[:300]:
        sx, sy, an, _ = crop.crop_pair(pr.ref_sym, pr.read_sym, pr.anchors)
        jobs.append((sx, sy, an, pr.strand))
    say("case: autotune, 300 pairs")
    ref = ph.Batch(pp.machines, jobs, pp.p).run().fetch()
    b = ph.Batch(pp.machines, jobs, pp.p)
    threads, waves, ms = b.autotune(repeats=2)
    assert threads in (256, 512) and ms > 0
    for r1, r2 in zip(b.fetch(), ref):
        for l in range(3):
            assert np.array_equal(r1[l], r2[l])
    b2 = ph.Batch(pp.machines, jobs, pp.p).run()
    for r1, r2 in zip(b2.fetch(), ref):
        assert np.array_equal(r1[0], r2[0])


def test_windowed_brute_force_tiny(ph, params):
    """Multi-window posteriors of every cell-state (threshold 0) vs exact
    enumeration of the band paths that stop on each window's traceback
    anti-diagonal (2-4 windows per pair: traceBackDiagonals 1,
    minDiagsBetweenTraceBack 3-4), random anchors with dynamic and static
    bands, N symbols, all ragged combinations; the default machine and both
    run-length strands."""
    from tests.brute_cases import brute_posteriors, dense, emitted, windowed_cases
    pp, op = params
    dm = ph.StateMachines.default_nucleotide()
    odm = O.default_nucleotide_model()
    worst = 0.0
    for mach, models in ((dm, (odm, odm)), (pp.machines, (op.reverse, op.forward))):
        for strand in (0, 1):
            om = models[strand]
            cases = windowed_cases(200 + strand + 2 * om.rle, 24, om)
            groups = {}
            for c in cases:
                p = c[3]
                key = (p.minDiagsBetweenTraceBack, p.diagonalExpansion, p.dynamicAnchorExpansion)
                groups.setdefault(key, []).append(c)
            for key, grp in sorted(groups.items()):
                say("case: windowed brute force rle=%s strand=%d params=%s, %d pairs"
                    % (om.rle, strand, key, len(grp)))
                bp = _bp(ph, grp[0][3])
                jobs = [(sx, sy, np.array(an, np.int64).reshape(-1, 3) if an else None, strand, rl, rr)
                        for sx, sy, an, _, rl, rr in grp]
                b = ph.Batch(mach, jobs, bp)
                b.run()
                for case, (gm, gx, gy, lp) in zip(grp, b.fetch()):
                    _, totals, post = brute_posteriors(om, case)
                    assert abs(lp - totals[-1]) < 1e-4 * max(1.0, abs(totals[-1])), (lp, totals[-1])
                    for st, lst in ((0, gm), (1, gx), (2, gy)):
                        err = float(np.max(np.abs(dense(as_rows(lst), post[st].shape) - emitted(post, st))))
                        assert err <= TOL + 1e-7, (key, st, err)
                        worst = max(worst, err)
    say("windowed brute force max-abs posterior err: %.3g" % worst)


def _with_n(rng, syms, frac, rle):
    """Replace a fraction of the symbols' bases by N (4), keeping run lengths."""
    s = np.array(syms, dtype=np.uint16).copy()
    idx = rng.sample(range(len(s)), max(1, int(frac * len(s))))
    for i in idx:
        s[i] = ((int(s[i]) >> 8) << 8 | 4) if rle else 4
    return s


@pytest.mark.parametrize("rle", [True, False])
def test_n_symbols_parity(ph, params, rle):
    """Ambiguous bases (N = symbol 4) in both sequences: the N emissions
    (match -2.772588722, gap -1.386294361, stateMachine.c:360-380) and, with
    run lengths, N read through the repeat table's A row (polisher.c:1758),
    on both strands, vs O_exact; plus tiny all-N-heavy pairs vs exact
    enumeration."""
    from marginpolish_amd import crop, sim
    from tests.conftest import PARAMS_NO_RLE
    pp = ph.PolishParams(PARAMS if rle else PARAMS_NO_RLE)
    op = O.PolishParams(PARAMS if rle else PARAMS_NO_RLE)
    rng = random.Random(77 + rle)
    w = sim.config0(n_reads=8, ref_len=2000, seed=21, W=16, strands=(1, 0))
    jobs = []
    for pr in w.pairs:
        sx, sy, an, _ = crop.crop_pair(pr.ref_sym, pr.read_sym, pr.anchors)
        if not rle:
            sx, sy = (np.asarray(sx) & 255).astype(np.uint16), (np.asarray(sy) & 255).astype(np.uint16)
        jobs.append((_with_n(rng, sx, 0.03, rle), _with_n(rng, sy, 0.03, rle), an, pr.strand))
    say("case: N symbols config0-like rle=%s, 8 pairs" % rle)
    b = ph.Batch(pp.machines, jobs, pp.p)
    b.run()
    worst = _check(b.fetch(), _oracle_jobs(op, jobs, op.p), op.p.threshold)
    say("N symbols rle=%s max-abs posterior err vs O_exact: %.3g" % (rle, worst))
    cases = []
    for i in range(32):
        lX, lY = rng.randint(1, 5), rng.randint(1, 5)
        strand = i & 1
        from tests.brute_cases import rand_syms
        cases.append((rand_syms(rng, lX, rle, 0.4), rand_syms(rng, lY, rle, 0.4), None, strand,
                      bool(i & 2), bool(i & 4)))
    p = ph.PairwiseAlignmentParameters(threshold=0.0, diagonalExpansion=20, dynamicAnchorExpansion=True)
    say("case: N symbols tiny enumeration rle=%s, 32 pairs" % rle)
    b = ph.Batch(pp.machines, cases, p)
    b.run()
    for (sx, sy, _, strand, rl, rr), (gm, gx, gy, lp) in zip(cases, b.fetch()):
        om = op.forward if strand else op.reverse
        total, post = brute.enumerate_paths(om, sx, sy, rl, rr)
        assert abs(lp - total) < 1e-4 * max(1.0, abs(total))
        for st, lst in ((0, gm), (1, gx), (2, gy)):
            from tests.brute_cases import dense, emitted
            err = float(np.max(np.abs(dense(as_rows(lst), post[st].shape) - emitted(post, st))))
            assert err <= TOL + 1e-7, (st, err)


@pytest.mark.parametrize("W", [10, 48], ids=["narrow", "wide"])
def test_window_jobs_equal_exact_chain(ph, params, W):
    """Window jobs (each traceback window its own job, the forward restarted
    512 anti-diagonals early from a flat vector) against the reference's
    single forward chain (MP_BATCH_EXACT_CHAIN) and O_exact: 6 pairs of ~8 k
    anti-diagonals with a window every 1000 (restarts from the third window
    on), both strands, ragged ends; W = 10 runs on the split kernel, W = 48
    (bands of 97 cells) on the wide-band kernel.  The two runs round
    differently (their FP32 scale trajectories differ), so the bar is the
    FP32 level: posteriors within 2e-5 of each other (each within 1e-4 of
    O_exact), log-likelihoods (chained in FP64 across the windows) within
    1e-3 nats; the device's FP64 batch total equals the sum of the pairs'."""
    from marginpolish_amd import crop, sim
    pp, op = params
    w = sim.config0(n_reads=6, ref_len=6000, seed=23, W=W, strands=(1, 0))
    jobs = []
    for i, pr in enumerate(w.pairs):
        sx, sy, an, _ = crop.crop_pair(pr.ref_sym, pr.read_sym, pr.anchors)
        jobs.append((sx, sy, an, pr.strand, bool(i & 1), bool(i & 2)))
    opar = O.Params(threshold=0.01, minDiagsBetweenTraceBack=1000, traceBackDiagonals=40, diagonalExpansion=W,
                    splitMatrixBiggerThanThis=250_000_000, dynamicAnchorExpansion=True)
    say("case: window jobs vs exact chain, 6 pairs, W=%d" % W)
    bw = ph.Batch(pp.machines, jobs, _bp(ph, opar)).run()
    be = ph.Batch(pp.machines, jobs, _bp(ph, opar), exact_chain=True).run()
    assert bw.stats()[0] > 12 and be.stats()[0] == 6  # one job per window vs one per pair
    gw, ge = bw.fetch(), be.fetch()
    worst, drift = 0.0, 0.0
    for a, c in zip(gw, ge):
        for l in range(3):
            err, _, only = compare(as_rows(a[l]), as_rows(c[l]), opar.threshold, 2e-5)
            worst = max(worst, err)
        drift = max(drift, abs(a[3] - c[3]))
    say("window jobs vs exact chain: max posterior diff %.3g, max log-likelihood diff %.3g" % (worst, drift))
    assert worst <= 2e-5 and drift <= 1e-3
    _check(gw, _oracle_jobs(op, jobs, opar), opar.threshold)
    for b, g in ((bw, gw), (be, ge)):
        tot = b.total_log_prob()
        assert abs(tot - math.fsum(r[3] for r in g)) <= 1e-9 * abs(tot)


def _band_widths(an, lX, lY, e):
    """Cells per anti-diagonal of the library's band (mp_band_construct, dynamic expansion)."""
    import ctypes as C
    from marginpolish_amd import _lib
    a = np.ascontiguousarray(np.asarray(an, np.int64).reshape(-1, 3))
    L = np.zeros(lX + lY + 1, np.int32)
    R = np.zeros(lX + lY + 1, np.int32)
    _lib.check(_lib.lib().mp_band_construct(a.ctypes.data_as(C.POINTER(C.c_int64)), len(a), lX, lY, e, 1,
                                            L.ctypes.data_as(C.POINTER(C.c_int32)),
                                            R.ctypes.data_as(C.POINTER(C.c_int32))))
    return (R - L) // 2 + 1


def test_wide_pairs_with_narrow_windows(ph, params):
    """Pairs whose band is wider than 64 cells only where the anchors leave a
    gap (100-300 positions; the dynamic expansion spans it, as it spans the
    low-confidence stretches of anchors taken from a POA graph), with a
    traceback window every 1000 anti-diagonals: each window job takes its
    own class, so the windows that fit 64 cells run on the split kernel
    (through a narrow copy of the band) and the others on the wide-band
    kernel.  Against the single-chain run (MP_BATCH_EXACT_CHAIN, the whole
    pair on the wide-band kernel) and O_exact, at the bars of the window-job
    test above."""
    from marginpolish_amd import crop, sim
    pp, op = params
    w = sim.config0(n_reads=6, ref_len=8000, seed=31, W=10, strands=(1, 0))
    jobs = []
    for i, pr in enumerate(w.pairs):
        sx, sy, an, _ = crop.crop_pair(pr.ref_sym, pr.read_sym, pr.anchors)
        at, gap = 1500 + 700 * i, 100 + 40 * i
        an = an[(an[:, 0] < at) | (an[:, 0] >= at + gap)]
        wd = _band_widths(an, len(sx), len(sy), 10)
        assert wd.max() > 64 and np.median(wd) <= 32, (i, wd.max(), np.median(wd))
        jobs.append((sx, sy, an, pr.strand, bool(i & 1), bool(i & 2)))
    opar = O.Params(threshold=0.01, minDiagsBetweenTraceBack=1000, traceBackDiagonals=40, diagonalExpansion=10,
                    splitMatrixBiggerThanThis=250_000_000, dynamicAnchorExpansion=True)
    say("case: wide pairs with narrow windows, 6 pairs")
    bw = ph.Batch(pp.machines, jobs, _bp(ph, opar)).run()
    be = ph.Batch(pp.machines, jobs, _bp(ph, opar), exact_chain=True).run()
    assert bw.stats()[0] > 12 and be.stats()[0] == 6
    gw, ge = bw.fetch(), be.fetch()
    worst, drift = 0.0, 0.0
    for a, c in zip(gw, ge):
        for l in range(3):
            err, _, only = compare(as_rows(a[l]), as_rows(c[l]), opar.threshold, 2e-5)
            worst = max(worst, err)
        drift = max(drift, abs(a[3] - c[3]))
    say("wide pairs, window jobs vs exact chain: max posterior diff %.3g, max log-likelihood diff %.3g"
        % (worst, drift))
    assert worst <= 2e-5 and drift <= 1e-3
    _check(gw, _oracle_jobs(op, jobs, opar), opar.threshold)
    assert abs(bw.total_log_prob() - math.fsum(r[3] for r in gw)) <= 1e-9 * abs(bw.total_log_prob())


def test_overflow_reruns_only_the_jobs_over_capacity(ph, params):
    """A threshold of 0.001 makes jobs write more records than their regions
    (1.5 per posterior anti-diagonal) hold: the first run re-runs only those
    jobs, into regions of their exact size appended to the record buffer.
    Its lists and log-likelihoods equal the second run's (exact regions, no
    re-run) bit for bit, and the oracle's."""
    from marginpolish_amd import crop, sim
    pp, op = params
    w = sim.config0(n_reads=8, ref_len=3000, seed=41, W=10, strands=(1, 0))
    jobs = []
    for i, pr in enumerate(w.pairs):
        sx, sy, an, _ = crop.crop_pair(pr.ref_sym, pr.read_sym, pr.anchors)
        jobs.append((sx, sy, an, pr.strand, bool(i & 1), bool(i & 2)))
    opar = O.Params(threshold=0.001, minDiagsBetweenTraceBack=1000, traceBackDiagonals=40, diagonalExpansion=10,
                    splitMatrixBiggerThanThis=250_000_000, dynamicAnchorExpansion=True)
    say("case: record overflow, 8 pairs, threshold 0.001")
    b = ph.Batch(pp.machines, jobs, _bp(ph, opar)).run()
    r1 = b.rerun_stats()
    g1 = b.fetch()
    b.run()
    r2 = b.rerun_stats()
    g2 = b.fetch()
    say("first run: %d re-run(s), %d jobs over capacity, %d records beyond; second run %s" % (r1 + (r2,)))
    assert r1[0] == 1 and 1 <= r1[1] <= b.stats()[0] and r2 == (0, 0, 0)
    for a, c in zip(g1, g2):
        for l in range(3):
            assert np.array_equal(a[l], c[l])
        assert a[3] == c[3]
    _check(g1, _oracle_jobs(op, jobs, opar), opar.threshold)
