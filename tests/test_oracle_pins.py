"""Pins the CPU oracle against the reference's own known answers and against
what the mathematics fixes (brute-force enumeration, invariants).

Reference tests restated here (tests/pairwiseAlignerTest.c, polisherTest.c):
  test_bands                  64-127   exact band diagonals
  test_logAdd                 129-139  +-1e-3 over random inputs
  test_symbolRLE              152-166  symbol packing round trip
  test_cell / test_diagonalDPCalculations 168-196, 256-339
  test_getAlignedPairsWithBanding 418-455 invariants
  test_getSplitPoints         487-556  exact split tuples
  test_getAlignedPairsWithRaggedEnds 650-696
  test_removeOverlapExample   polisherTest.c:1001-1025 crop (6, 2)
"""
import math
import random

import numpy as np
import pytest

from oracle import brute
from oracle import oracle as O
from tests.conftest import PARAMS


def diag(xay, l, r):
    return (xay, l, r)


def test_bands_known_answer():
    # pairwiseAlignerTest.c:64-127: anchors (1,0),(2,1),(3,3), lX=6, lY=5, expansion 2
    xl, xr = O.band([(1, 0, 0), (2, 1, 0), (3, 3, 0)], 6, 5, expansion=2, dynamic=False)
    want = [(0, 0), (-1, 1), (-2, 2), (-1, 3), (-2, 4), (-1, 3), (-2, 4), (-3, 3), (-2, 2), (-1, 3), (0, 2), (1, 1)]
    assert list(zip(xl.tolist(), xr.tolist())) == want


def test_dynamic_band_equals_static_when_expansions_equal():
    rng = random.Random(5)
    for _ in range(50):
        lX, lY = rng.randint(1, 60), rng.randint(1, 60)
        anchors, x, y = [], -1, -1
        while True:
            x += rng.randint(1, 5); y += rng.randint(1, 5)
            if x >= lX or y >= lY:
                break
            anchors.append((x, y, 6))
        a = O.band(anchors, lX, lY, expansion=6, dynamic=False)
        b = O.band(anchors, lX, lY, expansion=0, dynamic=True)
        if anchors:
            assert (a[0] == b[0]).all() and (a[1] == b[1]).all()


def test_band_no_anchors_is_full_rectangle():
    for lX, lY in [(1, 1), (4, 6), (7, 3), (10, 10)]:
        xl, xr = O.band([], lX, lY, expansion=0, dynamic=True)
        for d in range(lX + lY + 1):
            xs = [x for x in range(lX + 1) if 0 <= d - x <= lY]
            assert xl[d] == 2 * min(xs) - d and xr[d] == 2 * max(xs) - d


def test_logadd_tolerance_and_exact():
    rng = random.Random(7)
    for _ in range(20000):
        i, j = rng.random() + 1e-300, rng.random() + 1e-300
        l = math.exp(O.logadd(math.log(i), math.log(j), O.LOGADD_POLY))
        assert abs(l - (i + j)) < 1e-3
        e = math.exp(O.logadd(math.log(i), math.log(j), O.LOGADD_EXACT))
        assert abs(e - (i + j)) < 1e-12
    assert O.logadd(-math.inf, -1.0) == -1.0
    assert O.logadd(-2.0, -math.inf) == -2.0
    # the polynomial returns max beyond the 7.5-nat cut
    assert O.logadd(0.0, -7.5) == 0.0


def test_symbol_rle_packing():
    for ch in "ACGTacgtnN":
        b = int(O.symbols(ch)[0])
        assert b == ("ACGT".index(ch.upper()) if ch.upper() in "ACGT" else 4)
        for rc in range(0, 255):
            s = int(O.rle_symbols(ch, [rc], max_exclusive=256)[0])
            assert s & 255 == b and s >> 8 == rc
    # cap at maxRepeatCountExclusive - 1 (stateMachine.c:124-131)
    assert int(O.rle_symbols("A", [77])[0]) >> 8 == 50


def test_split_points_known_answer():
    # pairwiseAlignerTest.c:487-556
    area = 2000 * 2000
    assert O.split_points([], 3000, 1000, area) == [(0, 0, 3000, 1000)]
    assert O.split_points([], 20000, 25000, area, True, True) == []
    assert O.split_points([], 20000, 25000, area, True, False) == [(18000, 23000, 20000, 25000)]
    assert O.split_points([], 20000, 25000, area, False, True) == [(0, 0, 2000, 2000)]
    assert O.split_points([], 20000, 25000, area, False, False) == [(0, 0, 2000, 2000), (18000, 23000, 20000, 25000)]
    anchors = [(2000, 2000, 0), (4002, 4001, 0), (5000, 5000, 0), (8000, 6000, 0), (9000, 9000, 0),
               (10000, 14000, 0), (15000, 15000, 0), (16000, 16000, 0)]
    assert O.split_points(anchors, 20000, 25000, area) == [
        (0, 0, 3001, 3001), (3002, 3001, 9500, 11001), (9501, 12000, 12001, 14500),
        (13000, 14501, 18000, 18001), (18001, 23000, 20000, 25000)]


def test_diagonal_dp_known_answer():
    # pairwiseAlignerTest.c:256-339: AGCG vs AGTTCG, default machine, band expansion 2 (no anchors)
    m = O.default_nucleotide_model()
    sx, sy = O.symbols("AGCG"), O.symbols("AGTTCG")
    for mode in (O.LOGADD_POLY, O.LOGADD_EXACT):
        f, b, cuts, post = O.full_dp(m, sx, sy, expansion=2, mode=mode, want_post=True)
        assert abs(f - b) < 1e-3
        assert np.all(np.abs(cuts - f) < 1e-2)
        if mode == O.LOGADD_EXACT:
            assert abs(f - b) < 1e-10 and np.all(np.abs(cuts - f) < 1e-10)
        pairs = {(x - 1, y - 1) for x in range(1, 5) for y in range(1, 7) if post[x, y] >= 0.2}
        assert pairs == {(0, 0), (1, 1), (2, 4), (3, 5)}


def _models():
    pp = O.PolishParams(PARAMS)
    return [("default", O.default_nucleotide_model()), ("rle_fwd", pp.forward), ("rle_rev", pp.reverse)]


def _rand_rle(rng, n):
    bases = "".join(rng.choice("ACGT") for _ in range(n))
    rl = [rng.choice([1, 1, 1, 2, 2, 3, 5, 9]) for _ in range(n)]
    return O.rle_symbols(bases, rl)


@pytest.mark.parametrize("ragged", [(False, False), (True, True), (True, False), (False, True)])
def test_brute_force_posteriors_and_total(ragged):
    """Exact posteriors of every state at every cell, by path enumeration, vs
    the windowed oracle (one window: D < minDiags) with threshold 0 so every
    cell-state is emitted."""
    rng = random.Random(11 + ragged[0] * 2 + ragged[1])
    p = O.Params(threshold=0.0, diagonalExpansion=20, minDiagsBetweenTraceBack=1000, dynamicAnchorExpansion=True)
    for name, m in _models():
        for _ in range(6):
            lX, lY = rng.randint(1, 5), rng.randint(1, 5)
            sx = _rand_rle(rng, lX) if m.rle else O.symbols("".join(rng.choice("ACGT") for _ in range(lX)))
            sy = _rand_rle(rng, lY) if m.rle else O.symbols("".join(rng.choice("ACGT") for _ in range(lY)))
            total, post = brute.enumerate_paths(m, sx, sy, *ragged)
            mm, gx, gy, tot, _ = O.posteriors(m, sx, sy, [], p, *ragged, mode=O.LOGADD_EXACT)
            assert abs(tot - total) < 1e-9, name
            fp = O.forward_prob(m, sx, sy, [], p, *ragged, mode=O.LOGADD_EXACT)
            assert abs(fp - total) < 1e-9, name
            for st, lst in ((0, mm), (1, gx), (2, gy)):
                got = np.zeros((lX + 1, lY + 1))
                for w, x, y in lst:
                    got[x + 1, y + 1] = w / 1e7
                want = np.floor(np.minimum(post[st], 1.0) * 1e7) / 1e7
                # M needs x,y>0, X needs x>0, Y needs y>0 (pairwiseAligner.c:634-660)
                if st == 0:
                    want[0, :] = 0; want[:, 0] = 0
                elif st == 1:
                    want[0, :] = 0
                else:
                    want[:, 0] = 0
                assert np.max(np.abs(got - want)) <= 1.01e-7, (name, st)


def test_expectations_vs_brute_force():
    """Expected transition counts (getExpectations) vs enumeration; the
    reference's likelihood field is the per-diagonal sum (pairwiseAligner.c:674)."""
    rng = random.Random(3)
    p = O.Params(threshold=0.0, diagonalExpansion=20)
    m = O.default_nucleotide_model()
    for _ in range(5):
        lX, lY = rng.randint(1, 4), rng.randint(1, 4)
        sx = O.symbols("".join(rng.choice("ACGT") for _ in range(lX)))
        sy = O.symbols("".join(rng.choice("ACGT") for _ in range(lY)))
        e, like = O.expectations(m, sx, sy, [], p, mode=O.LOGADD_EXACT)
        # brute force: count transitions along paths (excluding the start->first and last->end terms)
        T = brute._trans(m)
        paths = []

        def rec(x, y, s, lp, trs):
            if x == lX and y == lY:
                paths.append((lp + m.trans[["matchContinue", "matchFromGapX", "matchFromGapY"][s]], list(trs)))
                return
            for ns in (0, 1, 2):
                dx, dy = brute.STEP[ns]
                nx, ny = x + dx, y + dy
                if nx > lX or ny > lY:
                    continue
                em = brute._emit(m, ns, int(sx[nx - 1]) if dx else 4, int(sy[ny - 1]) if dy else 4)
                trs.append((s, ns))
                rec(nx, ny, ns, lp + T[s, ns] + em, trs)
                trs.pop()

        rec(0, 0, 0, 0.0, [])
        tot = brute._lse([q for q, _ in paths])
        want = np.zeros(9)
        for lp, trs in paths:
            w = math.exp(lp - tot)
            for a, b in trs:
                want[a * 3 + b] += w
        assert np.max(np.abs(e - want)) < 1e-9
        assert abs(like - tot * (lX + lY)) < 1e-8


def test_windowed_invariants():
    """test_getAlignedPairsWithBanding (418-455) over random anchors, small
    windows, and the posterior-mass invariant per x of the last window."""
    rng = random.Random(17)
    m = O.default_nucleotide_model()
    for t in range(40):
        n = rng.randint(0, 100)
        sx = "".join(rng.choice("ACGT") for _ in range(n))
        sy = "".join(c for c in sx if rng.random() > 0.1)
        lX, lY = len(sx), len(sy)
        tbd = rng.randint(1, 10)
        p = O.Params(traceBackDiagonals=tbd, minDiagsBetweenTraceBack=tbd + rng.randint(2, 10),
                     diagonalExpansion=2 * rng.randint(0, 10), dynamicAnchorExpansion=rng.random() > 0.5)
        anchors, x, y = [], -1, -1
        while True:
            x += rng.randint(1, 20); y += rng.randint(1, 20)
            if x >= lX or y >= lY:
                break
            anchors.append((x, y, 2 * rng.randint(0, 5)))
        for mode in (O.LOGADD_POLY, O.LOGADD_EXACT):
            mm, gx, gy, _, _ = O.posteriors(m, O.symbols(sx), O.symbols(sy), anchors, p, mode=mode)
            seen = set()
            for w, x_, y_ in mm:
                assert 0 < w <= 10 ** 7 and 0 <= x_ < lX and 0 <= y_ < lY
                assert (x_, y_) not in seen
                seen.add((x_, y_))
            for w, x_, y_ in gx:
                assert 0 < w <= 10 ** 7 and 0 <= x_ < lX and -1 <= y_ < lY
            for w, x_, y_ in gy:
                assert 0 < w <= 10 ** 7 and -1 <= x_ < lX and 0 <= y_ < lY


def test_single_window_mass_per_x():
    """One window, threshold 0: every x>=1 is emitted exactly once, in M or X,
    so sum_y P_M(x,y) + P_X(x,y) = 1 (and the same for y with M and Y)."""
    rng = random.Random(23)
    pp = O.PolishParams(PARAMS)
    p = O.Params(threshold=0.0, diagonalExpansion=10, minDiagsBetweenTraceBack=10000, dynamicAnchorExpansion=True)
    for _ in range(5):
        n = rng.randint(20, 60)
        sx = _rand_rle(rng, n)
        sy = sx.copy()
        anchors = [(i, i, 10) for i in range(0, n, 3)]
        mm, gx, gy, _, _ = O.posteriors(pp.forward, sx, sy, anchors, p, mode=O.LOGADD_EXACT)
        mx = np.zeros(n); my = np.zeros(n)
        for w, x, y in mm:
            mx[x] += w; my[y] += w
        for w, x, y in gx:
            mx[x] += w
        for w, x, y in gy:
            my[y] += w
        # floor() drops < 1 unit per entry
        assert np.all(np.abs(mx - 1e7) < 3 * (2 * 10 + 4)) and np.all(np.abs(my - 1e7) < 3 * (2 * 10 + 4))


def test_ragged_ends_core_recovered():
    # test_getAlignedPairsWithRaggedEnds (650-696), checked on the strong pairs
    rng = random.Random(29)
    m = O.default_nucleotide_model()
    p = O.Params()
    for _ in range(5):
        core = "".join(rng.choice("ACGT") for _ in range(100))
        pre = "".join(rng.choice("ACGT") for _ in range(100))
        suf = "".join(rng.choice("ACGT") for _ in range(100))
        mm, _, _, _, _ = O.posteriors(m, O.symbols(core), O.symbols(pre + core + suf), [], p, True, True,
                                      with_indels=False, mode=O.LOGADD_POLY)
        strong = [(x, y) for w, x, y in mm if w > 5e6]
        assert len(strong) >= 95
        assert all(y == x + 100 for x, y in strong)


def remove_overlap(prefix, suffix, approx_overlap, pp, mode=O.LOGADD_POLY):
    """removeOverlap (polisher.c:1937-2006) on the oracle."""
    i = max(len(prefix) - approx_overlap, 0)
    j = min(approx_overlap, len(suffix))
    mm, _, _, _, _ = O.posteriors(O.default_nucleotide_model(), O.symbols(prefix[i:]), O.symbols(suffix[:j]), [],
                                  pp.p, True, True, with_indels=False, mode=mode)
    best = None
    for w, x, y in mm:
        if best is None or w > best[0]:
            best = (w, x, y)
    if best is None:
        return len(prefix), 0
    return best[1] + i, best[2]


def test_remove_overlap_known_answer():
    # polisherTest.c:1001-1025 expects crop (6, 2).  Only the polynomial
    # logAdd reproduces it: the top two match posteriors along the overlap are
    # 0.998457 (x=5,y=2) vs 0.998214 under the reference arithmetic but
    # 0.998032 vs 0.998376 (x=6,y=3) in exact FP64, so the exact oracle picks
    # (7, 3).  This pins O_faithful to the reference's own numbers.
    pp = O.PolishParams(PARAMS)
    assert remove_overlap("ACGTGATTTCA", "GATTTCAACGT", 10, pp, O.LOGADD_POLY) == (6, 2)
    assert remove_overlap("ACGTGATTTCA", "GATTTCAACGT", 10, pp, O.LOGADD_EXACT) == (7, 3)


def test_params_loader():
    pp = O.PolishParams(PARAMS)
    t = pp.forward.trans
    assert math.isclose(math.exp(t["matchContinue"]), 0.8)
    assert math.isclose(math.exp(t["gapOpenX"]), 0.1) and math.isclose(math.exp(t["matchFromGapY"]), 0.5)
    assert t["gapSwitchToX"] == -math.inf
    assert pp.forward.rle and pp.forward.strand == 1 and pp.reverse.strand == 0
    # reverse strand: M'[i][j] = M[3-i][3-j]
    assert pp.reverse.match[0] == pp.forward.match[15] and pp.reverse.match[1] == pp.forward.match[14]
    # repeat rows sum to one in base 10 (allParams.np.json)
    tab = pp.repeat.reshape(5, 51, 51)
    s = (10.0 ** tab[:4, 1:, :]).sum(axis=2)
    assert np.all(np.abs(s - 1.0) < 1e-3)
    assert pp.p.minDiagsBetweenTraceBack == 10000 and pp.p.diagonalExpansion == 10 and pp.p.dynamicAnchorExpansion


def test_faithful_vs_exact_gap_is_bounded():
    """O_faithful (polynomial logAdd) vs O_exact on a realistic pair: the
    difference is the reference's own approximation error, O(1e-3)."""
    from marginpolish_amd import crop, sim
    pp = O.PolishParams(PARAMS)
    w = sim.config0(n_reads=2, ref_len=600, seed=0)
    for pr in w.pairs:
        sx, sy, an, _ = crop.crop_pair(pr.ref_sym, pr.read_sym, pr.anchors)
        m = pp.forward if pr.strand else pp.reverse
        a = O.posteriors(m, sx, sy, an, pp.p, mode=O.LOGADD_POLY)
        b = O.posteriors(m, sx, sy, an, pp.p, mode=O.LOGADD_EXACT)
        da = {(x, y): w_ for w_, x, y in a[0]}
        db = {(x, y): w_ for w_, x, y in b[0]}
        common = set(da) & set(db)
        assert len(common) > 0.9 * max(len(da), len(db))
        assert max(abs(da[k] - db[k]) for k in common) < 2e-2 * 1e7


def _brute_emission_counts(m, sx, sy, post):
    """Expected emission counts from exact state posteriors, on the
    forward-strand parameters (reverse strand: reverse complement)."""
    match, gx, gy = np.zeros((4, 4)), np.zeros(4), np.zeros(4)
    rl = np.zeros((4, 51, 51))
    st = m.strand
    for x in range(len(sx) + 1):
        for y in range(len(sy) + 1):
            cx = int(sx[x - 1]) if x > 0 else 4
            cy = int(sy[y - 1]) if y > 0 else 4
            bx = (cx & 255) if m.rle else cx
            by = (cy & 255) if m.rle else cy
            if x > 0 and y > 0:
                if bx < 4 and by < 4:
                    match[(bx, by) if st else (3 - bx, 3 - by)] += post[0, x, y]
                if m.rle:
                    b = 0 if bx >= 4 else bx
                    rl[b if st else 3 - b, cx >> 8, cy >> 8] += post[0, x, y]
            if x > 0 and bx < 4:
                gx[bx if st else 3 - bx] += post[1, x, y]
            if y > 0 and by < 4:
                gy[by if st else 3 - by] += post[2, x, y]
    return match, gx, gy, rl


@pytest.mark.parametrize("ragged", [(False, False), (True, True)])
def test_emission_counts_vs_brute_force(ragged):
    """Expected emission counts (one window) equal the exact state
    posteriors of path enumeration summed per parameter, for the default
    machine and both RLE strands (the reverse strand maps through the reverse
    complement), and every x (y) is emitted once: match + gap-X counts sum to
    the number of non-N reference bases."""
    rng = random.Random(23)
    p = O.Params(threshold=0.0, diagonalExpansion=20, minDiagsBetweenTraceBack=1000)
    for name, m in _models():
        for _ in range(4):
            lX, lY = rng.randint(1, 5), rng.randint(1, 5)
            sx = _rand_rle(rng, lX) if m.rle else O.symbols("".join(rng.choice("ACGT") for _ in range(lX)))
            sy = _rand_rle(rng, lY) if m.rle else O.symbols("".join(rng.choice("ACGT") for _ in range(lY)))
            _, post = brute.enumerate_paths(m, sx, sy, *ragged)
            want = _brute_emission_counts(m, sx, sy, post)
            got = O.emission_counts(m, sx, sy, [], p, *ragged)
            for g, w in zip(got, want):
                assert np.allclose(g, w, atol=1e-12, rtol=1e-10), (name, g, w)
            assert abs(got[0].sum() + got[1].sum() - lX) < 1e-9
            assert abs(got[0].sum() + got[2].sum() - lY) < 1e-9
            if m.rle:
                assert abs(got[3].sum() - got[0].sum()) < 1e-9


@pytest.mark.parametrize("name", ["default", "rle_fwd", "rle_rev"])
def test_windowed_posteriors_vs_brute_force(name):
    """Multi-window posteriors (pairwiseAligner.c:722-792: backward
    re-initialised with endStateProb at each traceback anti-diagonal, only the
    anti-diagonals up to tb - (traceBackDiagonals + 1) emitted per window) vs
    exact enumeration of the band paths that stop on each window's traceback
    anti-diagonal.  2-4 windows per pair, random anchors, N symbols, all four
    ragged combinations; every cell-state (threshold 0).  O_exact within the
    1e-7 weight quantum; O_faithful (polynomial logAdd, the same control flow)
    within 5e-3, its own approximation error on these tiny DPs."""
    from tests.brute_cases import brute_posteriors, dense, emitted, windowed_cases
    model = dict(_models())[name]
    cases = windowed_cases(41 + len(name), 24, model)
    nwin = []
    for case in cases:
        sx, sy, anchors, p, rl, rr = case
        wins, totals, post = brute_posteriors(model, case)
        nwin.append(len(wins))
        for mode, tol in ((O.LOGADD_EXACT, 1.01e-7), (O.LOGADD_POLY, 5e-3)):
            mm, gx, gy, tot, _ = O.posteriors(model, sx, sy, anchors, p, rl, rr, mode=mode)
            # the pair's reported total is the final window's
            assert abs(tot - totals[-1]) < (1e-9 if mode == O.LOGADD_EXACT else 1e-2), (name, tot, totals[-1])
            for st, lst in ((0, mm), (1, gx), (2, gy)):
                got = dense(lst, post[st].shape)
                want = emitted(post, st)
                if mode == O.LOGADD_EXACT:
                    want = np.floor(want * 1e7) / 1e7
                err = np.max(np.abs(got - want))
                assert err <= tol, (name, mode, st, err, wins)
    assert max(nwin) >= 3 and min(nwin) >= 2


def test_windowed_brute_force_catches_wrong_end_vector():
    """The windowed enumeration distinguishes the reference's backward
    initialisation from plausible mistakes: a flat end vector (log 0 for
    every state) at intermediate traceback points, or one global window over
    the whole matrix, changes some posterior by more than 1e-2."""
    from tests.brute_cases import brute_posteriors, windowed_cases
    model = O.default_nucleotide_model()
    T, start, end = brute.machine(model)
    worst_flat, worst_global = 0.0, 0.0
    for case in windowed_cases(7, 12, model, p_n=0.0):
        sx, sy, anchors, p, rl, rr = case
        _, _, post = brute_posteriors(model, case)
        saved = brute.machine
        try:
            brute.machine = lambda m: (T, start, [[0.0, 0.0, 0.0], end[1]])
            _, _, post_flat = brute_posteriors(model, case)
        finally:
            brute.machine = saved
        worst_flat = max(worst_flat, float(np.max(np.abs(post_flat - post))))
        if not anchors:
            _, post_global = brute.enumerate_paths(model, sx, sy, rl, rr)
            worst_global = max(worst_global, float(np.max(np.abs(post_global - post))))
    assert worst_flat > 1e-2 and worst_global > 1e-2, (worst_flat, worst_global)
