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
