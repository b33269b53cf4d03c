"""Config 2 on the GPU: one POA realignment iteration of the 5 Mb workload
(50 chunks of 100 kb with 10 kb overlaps, 50x reads of mean 20 kb up to
100 kb, 8% error, RLE, anchors from the true alignment) as ONE batch, the
shape the bench times.

Checks:
- every pair: log-likelihood finite and negative; list entries inside the
  pair's matrix (-1 for a gap before the other sequence's first base);
  weights in (0, 1e7]; no repeated (x, y) in a list; on pairs with a single
  traceback window, the match and gap-X mass of every reference position at
  most 1 (with several windows each window's posteriors are conditioned on
  paths that end on its own traceback anti-diagonal, so a position near a
  window boundary can collect more: the oracle gives up to 1.3 here);
- the device FP64 batch sum of the pair log-likelihoods equals the sum of
  the per-pair FP64 values;
- a graph replay of the same batch gives bit-identical lists;
- a seeded sample of pairs (both strands, every length class): every
  posterior within 1e-4 of the exact FP64 oracle (O_exact) and the
  log-likelihood within 1e-2 nats, following getPosteriorProbsWithBanding
  (pairwiseAligner.c:687-808) through poa_realign's cropping
  (polisher.c:605-654).
"""
import os
import random
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import PARAMS
from tests.parity import as_rows, compare

pytestmark = pytest.mark.gpu

TOL = 1e-4
LL_TOL = 1e-2


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def workload():
    from marginpolish_amd import sim
    t = time.perf_counter()
    say("case: config2 workload generation (5 Mb reference, 50 chunks)")
    w = sim.config2()
    say("config2: %d pairs generated in %.1f s" % (len(w.pairs), time.perf_counter() - t))
    return w


@pytest.fixture(scope="module")
def batch(workload):
    from marginpolish_amd import pairhmm
    pp = pairhmm.PolishParams(PARAMS)
    pairs = [(p.ref_sym, p.read_sym, p.anchors, p.strand) for p in workload.pairs]
    say("case: config2 full batch, %d pairs, one launch" % len(pairs))
    t = time.perf_counter()
    b = pairhmm.Batch(pp.machines, pairs, pp.p, crop_reference=True)
    b.run()
    say("config2 batch: created and run in %.1f s, %d cells, kernel %.2f ms, lists %.2f ms"
        % (time.perf_counter() - t, b.cells, b.kernel_ms, b.list_ms))
    res = b.fetch()
    return pairs, b, res


EXTRA = []  # pairs the property test hands to the oracle sample


def test_config2_full_batch_properties(batch):
    from marginpolish_amd import crop
    pairs, b, res = batch
    counts, lp = b.counts()
    assert len(res) == len(pairs)
    assert np.all(np.isfinite(lp)) and np.all(lp < 0)
    # the device's FP64 batch sum is the sum of the per-pair FP64 log-likelihoods
    tot = b.total_log_prob()
    assert abs(tot - float(np.sum(lp))) <= 1e-9 * abs(tot)
    jobs, cap, recs = b.stats()
    minDiags = O.PolishParams(PARAMS).p.minDiagsBetweenTraceBack
    worst = (0.0, -1)
    # a record (one cell) gives up to three list entries: match, gap X, gap Y
    assert recs <= cap and recs <= int(counts.sum()) <= 3 * recs
    for i, ((rs, qs, an, _), (m, gx, gy, _)) in enumerate(zip(pairs, res)):
        first, end = crop.crop_bounds(an, len(rs), len(qs))
        col = np.zeros(len(rs) + 1)
        for arr in (m, gx, gy):
            if len(arr) == 0:
                continue
            assert arr["weight"].min() > 0 and arr["weight"].max() <= 10 ** 7, i
            # (a gap tuple before the other sequence's first base has -1 there)
            assert arr["x"].min() >= -1 and arr["x"].max() < len(rs), i
            assert arr["y"].min() >= -1 and arr["y"].max() < len(qs), i
            keys = (arr["x"].astype(np.int64) + 1) << 32 | (arr["y"].astype(np.int64) + 1)
            assert np.unique(keys).size == keys.size, i
        for arr in (m, gx):
            np.add.at(col, arr["x"], arr["weight"])
        if (end - first) + len(qs) < minDiags:
            # FP32 rounding of the posteriors
            assert col.max() <= 1e7 * (1 + 1e-3) + 16, i
        else:
            worst = max(worst, (col.max() / 1e7, i))
    # the multi-window pair with the largest column mass joins the oracle sample
    EXTRA.append(worst[1])
    say("config2 properties: %d pairs, %d records, %d device jobs; largest multi-window column mass %.3f (pair %d)"
        % (len(pairs), recs, jobs, worst[0], worst[1]))


def test_config2_graph_replay_identical(batch):
    pairs, b, res = batch
    rng = random.Random(11)
    pick = rng.sample(range(len(pairs)), 200)
    say("case: config2 graph replay, 200 sampled pairs compared")
    b.run()
    res2 = b.fetch()
    for i in pick:
        for li in range(3):
            assert np.array_equal(res[i][li], res2[i][li]), (i, li)
        assert res[i][3] == res2[i][3]


def test_config2_sample_vs_exact_oracle(batch):
    from marginpolish_amd import crop
    pairs, b, res = batch
    op = O.PolishParams(PARAMS)
    # a seeded sample covering both strands and the whole length range
    order = sorted(range(len(pairs)), key=lambda i: len(pairs[i][1]))
    rng = random.Random(5)
    n = 40
    pick = sorted({order[(q * len(order)) // n + rng.randrange(len(order) // n)] for q in range(n)})
    pick.append(order[-1])  # the longest read
    pick += [i for i in EXTRA if i >= 0 and i not in pick]
    assert {pairs[i][3] for i in pick} == {0, 1}
    say("case: config2 sample of %d pairs vs O_exact (RLE lengths %d..%d)"
        % (len(pick), min(len(pairs[i][1]) for i in pick), max(len(pairs[i][1]) for i in pick)))

    def one(i):
        rs, qs, an, strand = pairs[i]
        sx, sy, a2, off = crop.crop_pair(rs, qs, an)
        om, ox, oy, olp, _ = O.posteriors(op.forward if strand else op.reverse, sx, sy, a2, op.p)
        m, gx, gy, glp = res[i]
        worst = 0.0
        for g, o in ((m, om), (gx, ox), (gy, oy)):
            o = o.copy()
            o[:, 1] += off
            err, _, _ = compare(as_rows(g), o, op.p.threshold, TOL)
            worst = max(worst, err)
        return worst, abs(glp - olp), i

    t = time.perf_counter()
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as pool:
        got = list(pool.map(one, pick))
    perr = max(g[0] for g in got)
    lld = max(g[1] for g in got)
    say("config2 sample: %d pairs in %.1f s, max-abs posterior err %.3g, max LL drift %.3g nats"
        % (len(got), time.perf_counter() - t, perr, lld))
    for err, d, i in got:
        assert err <= TOL + 1e-7, (i, err)
        assert d <= LL_TOL, (i, d)
