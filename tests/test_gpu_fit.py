"""Parameter-fitting harness on the GPU: the gradient of the summed
log-likelihood in the nine log transitions (expected transition counts from
the expectation-mode kernel) against central finite differences of the
batched likelihood, against the FP64 oracle's expected counts, and recovery of
the generating transitions by Adam.

Bars: gradient vs central differences (h = 1e-3 in log space) within 1e-3
relative, with an absolute floor of 0.1 counts for the FP32 rounding of the
likelihood; expected counts vs O_exact within 1e-4 relative (+1e-3 absolute);
fitted transitions within 5% relative of the generating ones, with the gap
states tied as in the reference's symmetric machine (untied, the two switch
transitions X->Y and Y->X trade off along a nearly flat direction of the
likelihood: a gap-X/gap-Y pair between matches reads the same in either
order)."""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import PARAMS

pytestmark = pytest.mark.gpu

T_STAR = np.array([[0.90, 0.05, 0.05], [0.70, 0.25, 0.05], [0.70, 0.05, 0.25]])


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _emissions():
    sub = np.full((4, 4), 0.02 / 3)
    np.fill_diagonal(sub, 0.98)
    return (sub / 4).reshape(16), np.full(4, 0.25), np.full(4, 0.25)


@pytest.fixture(scope="module")
def fit():
    from marginpolish_amd import fit
    return fit


@pytest.fixture(scope="module")
def setup(fit, tmp_path_factory):
    from marginpolish_amd import pairhmm
    m16, gx, gy = _emissions()
    path = str(tmp_path_factory.mktemp("fit") / "generating_model.json")
    machines = fit.generating_model(PARAMS, T_STAR, m16, gx, gy, path)
    return pairhmm, machines, (m16, gx, gy)


def test_gradient_matches_central_differences(fit, setup):
    import torch
    ph, machines, (m16, gx, gy) = setup
    pairs = fit.simulate_pairs(48, 300, T_STAR, m16, gx, gy, 16, seed=11)
    p = ph.PairwiseAlignmentParameters(threshold=0.01, diagonalExpansion=16, dynamicAnchorExpansion=True)
    say("case: gradient vs central differences, 48 pairs x 300 bp")
    like = fit.TransitionLikelihood(machines, pairs, p)
    logT = torch.tensor(np.log(T_STAR), requires_grad=True)
    ll = fit.log_likelihood(logT, like)
    ll.backward()
    g = logT.grad.numpy()
    base = np.log(T_STAR)
    h = 1e-3
    for a, b in fit.TRANS_FROM_TO:
        e = np.zeros((3, 3))
        e[a, b] = h
        fd = (like.log_likelihood(base + e).sum() - like.log_likelihood(base - e).sum()) / (2 * h)
        assert abs(fd - g[a, b]) <= 1e-3 * abs(g[a, b]) + 0.1, ((a, b), fd, g[a, b])
    like.close()


def test_counts_match_cpu_oracle(fit, setup):
    ph, machines, (m16, gx, gy) = setup
    pairs = fit.simulate_pairs(6, 150, T_STAR, m16, gx, gy, 12, seed=12)
    p = ph.PairwiseAlignmentParameters(threshold=0.01, diagonalExpansion=12, dynamicAnchorExpansion=True)
    say("case: harness counts vs O_exact, 6 pairs x 150 bp")
    like = fit.TransitionLikelihood(machines, pairs, p)
    like.log_likelihood(np.log(T_STAR))
    like._expect.run()
    t, _ = like._expect.expectations()
    like.close()
    tab = machines.tables(1)
    om = O.Model(dict(zip(fit.TRANS_NAMES, tab["trans"])), tab["match"], tab["gapX"], tab["gapY"])
    op = O.Params(threshold=0.01, minDiagsBetweenTraceBack=1 << 30, traceBackDiagonals=40, diagonalExpansion=12,
                  splitMatrixBiggerThanThis=250_000_000, dynamicAnchorExpansion=True)
    for (sx, sy, an, _), tg in zip(pairs, t):
        te, _ = O.expectations(om, sx, sy, an, op)
        assert np.allclose(tg, te, rtol=1e-4, atol=1e-3), (tg, te)


def test_fit_recovers_generating_transitions(fit, setup):
    ph, machines, (m16, gx, gy) = setup
    pairs = fit.simulate_pairs(2048, 1000, T_STAR, m16, gx, gy, 16, seed=13)
    p = ph.PairwiseAlignmentParameters(threshold=0.01, diagonalExpansion=16, dynamicAnchorExpansion=True)
    say("case: Adam from a 20% perturbation, 2048 pairs x 1000 bp, gap states tied")
    like = fit.TransitionLikelihood(machines, pairs, p)
    rng = np.random.default_rng(5)
    T0 = T_STAR * (1.0 + 0.2 * rng.choice([-1.0, 1.0], size=(3, 3)))
    That, hist = fit.fit_transitions(like, T0, steps=200, lr=0.05, symmetric=True)
    like.close()
    rel = np.abs(That - T_STAR) / T_STAR
    say("fitted", That.round(4).tolist(), "max rel err %.4f" % rel.max(), "LL/pair %.3f -> %.3f" % (hist[0], hist[-1]))
    assert hist[-1] > hist[0]
    assert rel.max() < 0.05, That


# ---------------------------------------------------------------------------
# emission parameters (substitution, gaps, run-length confusion)
# ---------------------------------------------------------------------------

def _rle_setup(fit, tmp_path_factory, name):
    m16, gx, gy = _emissions()
    path = str(tmp_path_factory.mktemp(name) / "generating_model.json")
    machines = fit.generating_model(PARAMS, T_STAR, m16, gx, gy, path)
    R = fit.rl_confusion()
    machines.set_repeat_log10(np.log(np.where(R > 0, R, 1e-300)) / fit.RL_SCALE, True)
    return machines, (m16, gx, gy, R)


def _true_params(fit, m16, gx, gy, R):
    with np.errstate(divide="ignore"):
        return fit.ModelParams(np.log(T_STAR), np.log(m16).reshape(4, 4), np.log(gx), np.log(gy),
                               np.log(np.where(R > 0, R, 1e-300)))


def test_emission_counts_match_cpu_oracle(fit, tmp_path_factory):
    """Expected emission counts on the GPU (expectation-mode kernels, both
    strands, run-length symbols) against the CPU oracle's, which path
    enumeration pins."""
    machines, (m16, gx, gy, R) = _rle_setup(fit, tmp_path_factory, "emit")
    pairs = fit.simulate_rle_pairs(4, 150, T_STAR, m16, gx, gy, R, 12, seed=21, strand=1) + \
        fit.simulate_rle_pairs(4, 150, T_STAR, m16, gx, gy, R, 12, seed=22, strand=0)
    p = __import__("marginpolish_amd.pairhmm", fromlist=["x"]).PairwiseAlignmentParameters(
        threshold=0.01, diagonalExpansion=12, dynamicAnchorExpansion=True)
    say("case: emission counts vs O_exact, 8 RLE pairs x 150")
    like = fit.ModelLikelihood(machines, pairs, p, rle=True)
    _, grads = like.evaluate(_true_params(fit, m16, gx, gy, R))
    like.close()
    want = [np.zeros((4, 4)), np.zeros(4), np.zeros(4), np.zeros((4, 51, 51))]
    for strand in (0, 1):
        tab = machines.tables(strand)
        om = O.Model(dict(zip(fit.TRANS_NAMES, tab["trans"])), tab["match"], tab["gapX"], tab["gapY"], rle=True,
                     strand=strand, repeat=np.concatenate([np.log(np.where(R > 0, R, 1e-300)).reshape(-1) /
                                                           fit.RL_SCALE, np.zeros(51 * 51)]))
        op = O.Params(threshold=0.01, minDiagsBetweenTraceBack=1 << 30, traceBackDiagonals=40, diagonalExpansion=12,
                      splitMatrixBiggerThanThis=250_000_000, dynamicAnchorExpansion=True)
        for sx, sy, an, st in pairs:
            if st != strand:
                continue
            for acc, v in zip(want, O.emission_counts(om, sx, sy, an, op)):
                acc += v
    for g, w in zip(grads[1:], want):
        assert np.allclose(g, w, rtol=1e-4, atol=1e-3), (np.abs(g - w).max(), w.max())


def test_emission_gradient_matches_central_differences(fit, tmp_path_factory):
    """d(sum LL)/d(log emission parameter) from the counts against central
    differences of the batched likelihood (h = 1e-3): substitution entries,
    gap entries and run-length entries."""
    machines, (m16, gx, gy, R) = _rle_setup(fit, tmp_path_factory, "emitfd")
    pairs = fit.simulate_rle_pairs(48, 300, T_STAR, m16, gx, gy, R, 16, seed=31)
    p = __import__("marginpolish_amd.pairhmm", fromlist=["x"]).PairwiseAlignmentParameters(
        threshold=0.01, diagonalExpansion=16, dynamicAnchorExpansion=True)
    say("case: emission gradient vs central differences, 48 RLE pairs x 300")
    like = fit.ModelLikelihood(machines, pairs, p, rle=True)
    base = _true_params(fit, m16, gx, gy, R)
    _, g = like.evaluate(base)
    h = 1e-3
    probes = [(1, (0, 0)), (1, (1, 2)), (1, (3, 3)), (2, (1,)), (3, (2,)), (4, (0, 1, 1)), (4, (2, 2, 3)),
              (4, (3, 1, 2))]
    for k, idx in probes:
        arr = base.arrays()
        up, dn = [a.copy() for a in arr], [a.copy() for a in arr]
        up[k][idx] += h
        dn[k][idx] -= h
        fd = (like.log_likelihood(fit.ModelParams(*up)).sum() - like.log_likelihood(fit.ModelParams(*dn)).sum()) / (2 * h)
        assert abs(fd - g[k][idx]) <= 1e-3 * abs(g[k][idx]) + 0.1, ((k, idx), fd, g[k][idx])
    like.close()


def test_fit_recovers_generating_emissions(fit, tmp_path_factory):
    """Adam on the transitions (gap states tied), the joint substitution
    matrix and both gap emissions from a 20% perturbation, 8192 pairs x
    1000 bp: every parameter within 5% of the generating one.  (With 4096
    pairs the maximum-likelihood switch transition X<->Y itself sits 5.3% from
    the generating 0.05 -- Adam started at the generating parameters lands on
    the same point: a gap-X/gap-Y switch and a substitution explain the same
    columns, so the switch is the least determined parameter.)"""
    from marginpolish_amd import pairhmm
    m16, gx, gy = _emissions()
    gx = np.array([0.3, 0.2, 0.2, 0.3])
    gy = np.array([0.2, 0.3, 0.3, 0.2])
    path = str(tmp_path_factory.mktemp("emitfit") / "generating_model.json")
    machines = fit.generating_model(PARAMS, T_STAR, m16, gx, gy, path)
    pairs = fit.simulate_pairs(8192, 1000, T_STAR, m16, gx, gy, 16, seed=41)
    p = pairhmm.PairwiseAlignmentParameters(threshold=0.01, diagonalExpansion=16, dynamicAnchorExpansion=True)
    say("case: Adam on transitions + substitution + gap emissions, 8192 pairs x 1000 bp")
    like = fit.ModelLikelihood(machines, pairs, p, rle=False)
    truth = fit.ModelParams(np.log(T_STAR), np.log(m16).reshape(4, 4), np.log(gx), np.log(gy))
    rng = np.random.default_rng(8)
    pert = lambda a: a + np.log1p(0.2 * rng.choice([-1.0, 1.0], size=np.shape(a)))
    init = fit.ModelParams(*(pert(a) for a in truth.arrays()[:4]))
    fitted, hist = fit.fit_model(like, init, steps=200, lr=0.05, symmetric=True, fit=("T", "match", "gapX", "gapY"))
    like.close()
    rel = [np.abs(np.exp(fitted.logT) - T_STAR) / T_STAR, np.abs(np.exp(fitted.log_match).reshape(16) - m16) / m16,
           np.abs(np.exp(fitted.log_gap_x) - gx) / gx, np.abs(np.exp(fitted.log_gap_y) - gy) / gy]
    say("fit: T %.4f, match %.4f, gapX %.4f, gapY %.4f; LL/pair %.3f -> %.3f"
        % tuple([r.max() for r in rel] + [hist[0], hist[-1]]))
    assert hist[-1] > hist[0]
    assert max(r.max() for r in rel) < 0.05


def test_fit_recovers_run_length_confusion(fit, tmp_path_factory):
    """The 51 x 51 run-length confusion with the substitution matrix and the
    gap-X emissions, on run-length symbols from a generating model without
    insertions (the run-length term scores matched symbols only, so only then
    is the model a proper distribution of the read's run lengths),
    transitions held at the generating values; 4096 pairs x 1000 bp, 20%
    perturbation: substitution and gap-X within 5%, run-length entries with at
    least 5000 expected observations within 5%."""
    from marginpolish_amd import pairhmm
    T_noins = np.array([[0.95, 0.05, 0.0], [0.70, 0.30, 0.0], [0.70, 0.0, 0.30]])
    m16, gx, gy = _emissions()
    path = str(tmp_path_factory.mktemp("rlfit") / "generating_model.json")
    machines = fit.generating_model(PARAMS, T_noins, m16, gx, gy, path)
    R = fit.rl_confusion()
    with np.errstate(divide="ignore"):
        logR = np.log(R)
    machines.set_repeat_log10(np.maximum(logR, -690.0) / fit.RL_SCALE, True)
    pairs = fit.simulate_rle_pairs(4096, 1000, T_noins, m16, gx, gy, R, 16, seed=43)
    p = pairhmm.PairwiseAlignmentParameters(threshold=0.01, diagonalExpansion=16, dynamicAnchorExpansion=True)
    say("case: Adam on run-length confusion + substitution + gap X, 4096 RLE pairs x 1000 bp, no insertions")
    like = fit.ModelLikelihood(machines, pairs, p, rle=True)
    with np.errstate(divide="ignore"):
        truth = fit.ModelParams(np.log(T_noins), np.log(m16).reshape(4, 4), np.log(gx), np.log(gy),
                                np.maximum(logR, -690.0))
    _, g_true = like.evaluate(truth)
    rng = np.random.default_rng(9)
    pert = lambda a: a + np.log1p(0.2 * rng.choice([-1.0, 1.0], size=np.shape(a)))
    init = fit.ModelParams(truth.logT, pert(truth.log_match), pert(truth.log_gap_x), truth.log_gap_y,
                           pert(truth.rho))
    fitted, hist = fit.fit_model(like, init, steps=200, lr=0.05, symmetric=False, fit=("match", "gapX", "rl"))
    like.close()
    relM = np.abs(np.exp(fitted.log_match).reshape(16) - m16) / m16
    relX = np.abs(np.exp(fitted.log_gap_x) - gx) / gx
    sel = g_true[4] >= 5000
    relR = np.abs(np.exp(fitted.rho[sel]) - R[sel]) / R[sel]
    say("fit: match %.4f, gapX %.4f, run-length (%d entries) %.4f; LL/pair %.3f -> %.3f"
        % (relM.max(), relX.max(), int(sel.sum()), relR.max(), hist[0], hist[-1]))
    assert hist[-1] > hist[0]
    assert relM.max() < 0.05 and relX.max() < 0.05
    assert sel.sum() >= 16 and relR.max() < 0.05
