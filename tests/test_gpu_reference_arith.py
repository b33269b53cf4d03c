"""Reference-arithmetic batches (MP_BATCH_REFERENCE_ARITHMETIC): the
pair-HMM in the reference's own FP64 log space with its polynomial logAdd,
operation order, windows and total cadence, against the oracle in the same
arithmetic (O_faithful).  Lists must be the oracle's tuple for tuple, in
order; a weight may differ by one unit only where the GPU's and the host's
FP64 exp round p * 1e7 on either side of an integer (measured: none), and the
log-likelihood agrees to 1e-9 relative."""
import random
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tests.brute_cases import windowed_cases
from tests.conftest import PARAMS
from tests.parity import as_rows

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ph():
    from marginpolish_amd import pairhmm
    return pairhmm


def _bp(ph, op):
    return ph.PairwiseAlignmentParameters(op.threshold, op.minDiagsBetweenTraceBack, op.traceBackDiagonals,
                                          op.diagonalExpansion, op.splitMatrixBiggerThanThis,
                                          op.dynamicAnchorExpansion)


STATS = {"tuples": 0, "weight_off_by_one": 0, "off_weights": []}


def _same_lists(g, o, what):
    """GPU rows and oracle rows: same (x, y) in the same order, weights within one unit."""
    g = as_rows(g)
    o = np.asarray(o, np.int64).reshape(-1, 3)
    assert g.shape == o.shape, (what, g.shape, o.shape)
    if g.size:
        assert np.array_equal(g[:, 1:], o[:, 1:]), what
        d = np.abs(g[:, 0] - o[:, 0])
        assert d.max() <= 1, (what, d.max())
        STATS["weight_off_by_one"] += int((d > 0).sum())
        STATS["off_weights"] += o[d > 0, 0].tolist()
    STATS["tuples"] += len(g)


def _run(ph, machines, models, jobs, opar, mode=0):
    b = ph.Batch(machines, jobs, _bp(ph, opar), mode=mode, reference_arithmetic=True)
    b.run()
    res = b.fetch()
    for k, ((sx, sy, an, strand, *rg), (gm, gx, gy, lp)) in enumerate(zip(jobs, res)):
        rl = rg[0] if rg else False
        rr = rg[1] if len(rg) > 1 else False
        om, ox, oy, olp, _ = O.posteriors(models[strand], sx, sy, an, opar, rl, rr, with_indels=mode == 0,
                                          mode=O.LOGADD_POLY)
        _same_lists(gm, om, ("match", k))
        if mode == 0:
            _same_lists(gx, ox, ("gapX", k))
            _same_lists(gy, oy, ("gapY", k))
        assert abs(lp - olp) <= 1e-9 * max(1.0, abs(olp)), (k, lp, olp)
    return b


def test_windowed_tiny_cases(ph):
    """2-4 traceback windows per pair, N symbols, ragged ends, threshold 0."""
    pp, op = ph.PolishParams(PARAMS), O.PolishParams(PARAMS)
    dm = ph.StateMachines.default_nucleotide()
    odm = O.default_nucleotide_model()
    for mach, models in ((dm, (odm, odm)), (pp.machines, (op.reverse, op.forward))):
        for strand in (0, 1):
            cases = windowed_cases(300 + strand + 2 * models[strand].rle, 16, models[strand])
            groups = {}
            for c in cases:
                q = c[3]
                groups.setdefault((q.minDiagsBetweenTraceBack, q.diagonalExpansion, q.dynamicAnchorExpansion),
                                  []).append(c)
            for key, grp in sorted(groups.items()):
                say("case: reference arithmetic tiny rle=%s strand=%d %s x%d" % (models[1].rle, strand, key,
                                                                                len(grp)))
                jobs = [(sx, sy, np.array(an, np.int64).reshape(-1, 3) if an else None, strand, rl, rr)
                        for sx, sy, an, _, rl, rr in grp]
                _run(ph, mach, models, jobs, grp[0][3])


def test_config0_like_and_splits(ph):
    """Config0-like reads (both strands, W = 32, one window), several windows
    (minDiagsBetweenTraceBack 300), and large-gap splits with ragged ends;
    posterior and match-only modes."""
    from marginpolish_amd import crop, sim
    pp, op = ph.PolishParams(PARAMS), O.PolishParams(PARAMS)
    models = (op.reverse, op.forward)
    w = sim.config0(n_reads=6, ref_len=2000, seed=3, strands=(1, 0))
    jobs = []
    for i, pr in enumerate(w.pairs):
        sx, sy, an, _ = crop.crop_pair(pr.ref_sym, pr.read_sym, pr.anchors)
        jobs.append((sx, sy, an, pr.strand, bool(i & 1), bool(i & 2)))
    say("case: reference arithmetic config0-like x6")
    _run(ph, pp.machines, models, jobs, op.p)
    _run(ph, pp.machines, models, jobs, op.p, mode=1)
    win = O.Params(threshold=0.01, minDiagsBetweenTraceBack=300, traceBackDiagonals=40, diagonalExpansion=10,
                   splitMatrixBiggerThanThis=250_000_000, dynamicAnchorExpansion=True)
    w2 = sim.config0(n_reads=4, ref_len=2000, seed=7, W=10, strands=(1, 0))
    jobs2 = []
    for pr in w2.pairs:
        sx, sy, an, _ = crop.crop_pair(pr.ref_sym, pr.read_sym, pr.anchors)
        jobs2.append((sx, sy, an, pr.strand))
    say("case: reference arithmetic multi-window x4")
    _run(ph, pp.machines, models, jobs2, win)
    rng = random.Random(5)
    jobs3 = []
    for i in range(6):
        n = rng.randint(60, 200)
        bases = "".join(rng.choice("ACGT") for _ in range(n))
        sx = O.rle_symbols(bases, [rng.choice([1, 1, 2, 3]) for _ in range(n)])
        keep = [j for j in range(n) if rng.random() > 0.05]
        sy = sx[keep]
        anchors = []
        for k in range(0, len(keep), rng.randint(15, 40)):
            if not anchors or (keep[k] > anchors[-1][0] and k > anchors[-1][1]):
                anchors.append((keep[k], k, 6))
        jobs3.append((sx, sy, np.array(anchors, np.int64), i & 1, bool(i & 2), bool(i & 4)))
    split = O.Params(threshold=0.01, minDiagsBetweenTraceBack=1000, traceBackDiagonals=40, diagonalExpansion=6,
                     splitMatrixBiggerThanThis=900, dynamicAnchorExpansion=True)
    say("case: reference arithmetic splits + ragged x6")
    _run(ph, pp.machines, models, jobs3, split)
    say("reference arithmetic: %d tuples, %d weights one unit off" % (STATS["tuples"], STATS["weight_off_by_one"]))
    # the arithmetic is the oracle's operation for operation; only the last bit of a
    # non-trivial exp may differ, which moves a weight in about 1e-9 of the tuples
    assert STATS["weight_off_by_one"] <= max(1, STATS["tuples"] // 10000), STATS["off_weights"][:20]


def test_refusals(ph):
    pp = ph.PolishParams(PARAMS)
    sx = O.symbols("ACGTACGT")
    with pytest.raises(ph.MPError):
        ph.Batch(pp.machines, [(sx, sx, None, 1)], pp.p, mode=2, reference_arithmetic=True)
    with pytest.raises(ph.MPError):
        ph.Batch(pp.machines, [(sx, sx, None, 1)], pp.p, fp64=True, reference_arithmetic=True)
