"""The whole realignment pipeline on the GPU: initial batch, POA graphs,
consensus iterations, run lengths, contig merge — a simulated draft gets
closer to the truth it was made from."""
import numpy as np
import pytest

from marginpolish_amd import pairhmm as ph
from marginpolish_amd import polish, sim
from marginpolish_amd.sim import banded_edit_distance
from tests.conftest import PARAMS

pytestmark = pytest.mark.gpu


def _strings(truth, draft):
    chars = np.frombuffer(b"ACGT", np.uint8)
    return chars[truth].tobytes().decode(), chars[draft].tobytes().decode()


def test_banded_edit_distance():
    rng = np.random.default_rng(0)
    for _ in range(20):
        a = "".join(rng.choice(list("ACGT"), size=rng.integers(0, 40)))
        b = "".join(rng.choice(list("ACGT"), size=rng.integers(0, 40)))
        # full DP for reference
        d = list(range(len(b) + 1))
        for i, ca in enumerate(a, 1):
            nd = [i] + [0] * len(b)
            for j, cb in enumerate(b, 1):
                nd[j] = min(d[j] + 1, nd[j - 1] + 1, d[j - 1] + (ca != cb))
            d = nd
        assert banded_edit_distance(a, b, band=100) == d[-1]


def _run(params_path, rle, iters, seed=7):
    truth, draft, chunks = sim.polish_workload(seed=seed, rle=rle)
    params = ph.PolishParams(params_path)
    cfg = polish.PolishConfig(params_path)
    cfg.maxRealignmentPolishIterations = 0
    cfg.maxPoaConsensusIterations = iters
    cfg.chunkBoundary = 250
    refs, reads, anchors = [c.ref for c in chunks], [c.reads for c in chunks], [c.anchors for c in chunks]
    return truth, draft, polish.polish_contig(params, cfg, refs, reads, anchors), (params, cfg, refs, reads, anchors)


def test_polish_draft_towards_truth():
    """Plain-string params (no run-length encoding): the consensus rounds move
    a 3%-error draft most of the way to the truth its 8%-error reads came from."""
    from tests.conftest import PARAMS_NO_RLE
    truth, draft, merged, args = _run(PARAMS_NO_RLE, False, 3)
    t_str, d_str = _strings(truth, draft)
    d0 = banded_edit_distance(d_str, t_str)
    d1 = banded_edit_distance(merged, t_str)
    print("draft edit distance %d, polished %d" % (d0, d1), flush=True)
    assert d1 < 0.4 * d0  # measured: 808 -> 248
    assert polish.polish_contig(*args) == merged  # deterministic


def test_polish_rle_pipeline_runs():
    """Run-length params: the whole pipeline (consensus rounds with the
    maximum-likelihood run lengths) runs and is deterministic.  Accuracy is
    not asserted: these simulated reads do not follow the ONT run-length
    error model the params' table was trained on (the estimates lengthen
    runs that the simulator does not shorten)."""
    truth, draft, merged, args = _run(PARAMS, True, 2)
    assert abs(len(merged) - len(truth)) < 0.05 * len(truth)
    assert polish.polish_contig(*args) == merged


def test_default_params_refuse_the_unbuilt_polish_path():
    cfg = polish.PolishConfig(PARAMS)  # maxRealignmentPolishIterations = 1 in these params
    truth, draft, chunks = sim.polish_workload(seed=1, genome_len=3000, chunk_size=3000)
    with pytest.raises(NotImplementedError):
        polish.realign_all(ph.PolishParams(PARAMS), cfg, [chunks[0].ref], [chunks[0].reads], [chunks[0].anchors])


def test_polish_rle_fixes_run_lengths():
    """Reads drawn from the params' own run-length error model: on covered
    short runs (true length <= 3, where P(observed | length) separates the
    lengths) the maximum-likelihood estimate repairs most of the draft's
    off-by-one runs.  Longer runs stay hard for the reference's estimator:
    its candidates only span the observed lengths, and the model under-calls
    long runs (measured on a CPU-built graph: 96 wrong of 2891 covered runs,
    mostly of length >= 5, against 51 in the draft)."""
    from oracle import oracle as O
    truth, draft, chunks = sim.polish_workload_rle(O.PolishParams(PARAMS).repeat)
    params = ph.PolishParams(PARAMS)
    cfg = polish.PolishConfig(PARAMS)
    cfg.maxRealignmentPolishIterations = 0
    cfg.chunkBoundary = 0
    cfg.maxPoaConsensusIterations = 0
    refs, reads, anchors = [c.ref for c in chunks], [c.reads for c in chunks], [c.anchors for c in chunks]
    res = polish.realign_all(params, cfg, refs, reads, anchors)
    est_wrong = draft_wrong = short = 0
    for c, (g, s) in zip(chunks, res):
        b = np.frombuffer(s.encode(), np.uint8)
        lens = np.diff(np.r_[np.flatnonzero(np.r_[True, b[1:] != b[:-1]]), b.size])
        st, _ = g.observations(0)
        covered = np.diff(st)[1:] > 0
        tl = truth[1][c.start:c.end].astype(np.int64)
        dl = draft[1][c.start:c.end].astype(np.int64)
        assert lens.size == tl.size  # the bases are kept: one polished run per draft run
        sel = covered & (tl <= 3)
        short += int(sel.sum())
        est_wrong += int((lens[sel] != tl[sel]).sum())
        draft_wrong += int((dl[sel] != tl[sel]).sum())
    print("short covered runs %d: draft wrong %d, estimate wrong %d" % (short, draft_wrong, est_wrong), flush=True)
    assert short > 0.8 * int((truth[1] <= 3).sum())
    assert est_wrong < draft_wrong / 2
