"""Precision and long-band stress at its own shape: 256 pairs of a 100 kb read
against its own 100 kb reference (run-length symbols, ~65 kb; 12% error;
structural indels of 50-500 bp over 0.5% of the length), the five-state
machine (long gaps open 0.0018, extend 0.9965), traceback windows every 10000
anti-diagonals, band half-widths W = 128, 256, 512 and 1024.

Per width, every tuple of every pair: FP32 (per-cell exponents) against the
FP64 GPU path (the same kernel in double; the emission tables stay FP32 in
both), and both against the CPU FP64 oracle (O_exact, five states) on the
first ~6000 x 6000 RLE positions of one pair per strand.  Bars: FP32 max-abs
posterior error 1e-4 and log-likelihood drift 1e-2 nats (relative 1e-7 for
the ~1e5-nat totals); FP64 against the oracle 1e-6 and 1e-4 nats (the FP32
rounding of the emission tables).  A tuple held by one side only must sit
within the bar of the 0.01 threshold.  GCUPS of both paths are printed per
width."""
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import PARAMS
from tests.parity import batch_list_err

pytestmark = pytest.mark.gpu

WIDTHS = (128, 256, 512, 1024)
SAMPLE = 6000  # RLE positions of the oracle's sample per strand


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def workload():
    from marginpolish_amd import pairhmm, sim
    t0 = time.time()
    w = sim.config4(n_pairs=256, W=128)
    say("config4 workload: %d pairs, %d RLE reference symbols, built in %.1f s"
        % (len(w.pairs), sum(p.ref_sym.size for p in w.pairs), time.time() - t0))
    pp = pairhmm.PolishParams(PARAMS)
    mach = pp.machines
    mach.add_long_gaps()
    return w, mach, pairhmm.default_engine()


def _params(ph, W):
    return ph.PairwiseAlignmentParameters(threshold=0.01, minDiagsBetweenTraceBack=10000, traceBackDiagonals=40,
                                          diagonalExpansion=W, splitMatrixBiggerThanThis=250_000_000,
                                          dynamicAnchorExpansion=True)


def _pairs(w, W):
    out = []
    for p in w.pairs:
        an = p.anchors.copy()
        an[:, 2] = W  # the workload's per-anchor expansion (sim.config4(W=W) draws the same sequences)
        out.append((p.ref_sym, p.read_sym, an, p.strand))
    return out


def _sample(pair, n):
    """The first ~n x n RLE positions of a pair, cut at an anchor."""
    sx, sy, an, strand = pair
    an = an[(an[:, 0] < n) & (an[:, 1] < n)]
    cut = int(an[-1, 0]) + 1, int(an[-1, 1]) + 1
    return sx[: cut[0]], sy[: cut[1]], an[:-1], strand


def _rows(lst):
    return {"weight": lst[:, 0], "x": lst[:, 1], "y": lst[:, 2]}


@pytest.mark.parametrize("W", WIDTHS)
def test_config4_precision_per_band_width(workload, W):
    from marginpolish_amd import pairhmm as ph
    w, mach, eng = workload
    pairs = _pairs(w, W)
    bp = _params(ph, W)
    say("case: config4 W=%d, %d pairs, FP32 vs FP64 GPU" % (W, len(pairs)))
    res, lps, gcups = {}, {}, {}
    for tag, fp64 in (("fp32", False), ("fp64", True)):
        b = ph.Batch(mach, pairs, bp, engine=eng, fp64=fp64)
        b.run()
        res[tag] = b.fetch()
        b.run()
        gcups[tag] = b.cells / (b.kernel_ms * 1e-3) / 1e9
        lps[tag] = np.array([r[3] for r in res[tag]])
        say("W=%d %s: %d cells, kernel %.1f ms, %.2f GCUPS" % (W, tag, b.cells, b.kernel_ms, gcups[tag]))
        b.close()
    err, edge, dups, both = batch_list_err(res["fp32"], res["fp64"])
    drift = np.abs(lps["fp32"] - lps["fp64"])
    say("W=%d FP32 vs FP64 GPU: %d tuples in both, max-abs posterior err %.3g, one-sided max %.4f, LL drift max "
        "%.3g nats (%.3g relative)" % (W, both, err, edge, drift.max(), (drift / np.abs(lps["fp64"])).max()))
    assert dups == 0
    assert err <= 1e-4
    assert edge <= 0.01 + 1e-4
    assert drift.max() <= 1e-2
    assert (drift / np.abs(lps["fp64"])).max() <= 1e-7
    # both precisions against the CPU FP64 oracle on one sample per strand
    T5 = mach.transitions5()
    op = O.PolishParams(PARAMS)
    opar = O.Params(threshold=0.01, minDiagsBetweenTraceBack=10000, traceBackDiagonals=40, diagonalExpansion=W,
                    splitMatrixBiggerThanThis=250_000_000, dynamicAnchorExpansion=True)
    samples = [_sample(pairs[0], SAMPLE), _sample(pairs[1], SAMPLE)]
    assert {s[3] for s in samples} == {0, 1}

    def oracle(s):
        sx, sy, an, strand = s
        base = op.forward if strand else op.reverse
        om5 = O.Model(base.trans, base.match, base.gapX, base.gapY, base.rle, base.strand, base.repeat, trans5=T5)
        return O.posteriors(om5, sx, sy, an, opar, mode=O.LOGADD_EXACT)

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=2) as pool:
        ora = list(pool.map(oracle, samples))
    say("W=%d oracle samples (%d and %d cells) in %.1f s" % (W, ora[0][4], ora[1][4], time.perf_counter() - t0))
    want = [tuple(_rows(np.asarray(l, np.int64).reshape(-1, 3)) for l in o[:3]) + (o[3],) for o in ora]
    for tag, fp64, bar, ll_bar in (("fp32", False, 1e-4, 1e-2), ("fp64", True, 1e-6, 1e-4)):
        b = ph.Batch(mach, samples, bp, engine=eng, fp64=fp64)
        b.run()
        got = b.fetch()
        b.close()
        e, ed, du, bo = batch_list_err(got, want)
        d = max(abs(g[3] - o[3]) for g, o in zip(got, want))
        say("W=%d %s vs CPU FP64 oracle: %d tuples in both, max-abs posterior err %.3g, one-sided max %.4f, "
            "LL drift %.3g nats" % (W, tag, bo, e, ed, d))
        assert du == 0
        assert e <= bar
        assert ed <= 0.01 + bar
        assert d <= ll_bar
