"""Parameter-fitting harness on the GPU.

Small cases: the gradient of the summed log-likelihood (expected transition
and emission counts from the expectation-mode kernel) against central finite
differences of the FP32 batched likelihood (h = 1e-3 in log space, within
1e-3 relative with an absolute floor of 0.1 counts for the FP32 rounding of
the likelihood) and against the FP64 oracle's expected counts (O_exact,
within 1e-4 relative + 1e-3 absolute).

The differentiation workload at its own shape (4096 pairs, 2 kb references,
band half-width 48, reads from a generating model with transitions, 4 x 4
substitution, gap emissions and a 51 x 51 run-length confusion): the FP32
gradient against central differences of the FP64 likelihood (h = 1e-4,
within 1e-3 relative), counts against O_exact on 16 of the pairs, and 200
Adam steps from a 20% perturbation recovering the generating parameters
within 5% (gap states tied as in the reference's symmetric machine: untied,
the two switch transitions X->Y and Y->X trade off along a nearly flat
direction of the likelihood, a gap-X/gap-Y pair between matches reading the
same in either order)."""
import math
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


# ---------------------------------------------------------------------------
# the differentiation workload at its own shape: 4096 pairs, 2 kb references,
# band half-width 48
# ---------------------------------------------------------------------------

N_PAIRS, REF_LEN, W = 4096, 2000, 48


def _spec_rle_pairs(fit, T, m16, gx, gy, R, seed):
    """4096 run-length pairs of 2 kb references, half on each strand."""
    half = N_PAIRS // 2
    return fit.simulate_rle_pairs(half, REF_LEN, T, m16, gx, gy, R, W, seed=seed, strand=1) + \
        fit.simulate_rle_pairs(half, REF_LEN, T, m16, gx, gy, R, W, seed=seed + 1, strand=0)


def _spec_params(fit):
    from marginpolish_amd import pairhmm
    return pairhmm.PairwiseAlignmentParameters(threshold=0.01, diagonalExpansion=W, dynamicAnchorExpansion=True)


def _table_step(v, h):
    """The step in log space that an FP32 table entry exp(v) really takes when
    v moves by +-h: the engine stores (float)exp(v) (emission and run-length
    tables; FP64 batches keep the transitions in double)."""
    up = np.log(np.float64(np.float32(np.exp(v + h))))
    dn = np.log(np.float64(np.float32(np.exp(v - h))))
    return up - dn


@pytest.fixture(scope="module")
def spec_rle(fit, tmp_path_factory):
    """The generating model (3-state transitions, 4 x 4 substitution, gap
    emissions, 51 x 51 run-length confusion) and 4096 pairs drawn from it."""
    machines, (m16, gx, gy, R) = _rle_setup(fit, tmp_path_factory, "spec")
    gx = np.array([0.3, 0.2, 0.2, 0.3])
    gy = np.array([0.2, 0.3, 0.3, 0.2])
    machines.set_emissions(np.log(m16), np.log(gx), np.log(gy))
    pairs = _spec_rle_pairs(fit, T_STAR, m16, gx, gy, R, seed=61)
    return machines, pairs, (m16, gx, gy, R)


def test_spec_shape_gradient_vs_fp64_central_differences(fit, spec_rle):
    """d(sum LL)/d(natural-log parameter) from the FP32 forward and the
    hand-written backward (expected counts) on 4096 run-length pairs x 2 kb
    at W = 48, against central differences (h = 1e-4) of the FP64
    likelihood, within 1e-3 relative: all nine transitions, all 16
    substitution entries, all gap emissions and run-length entries.  The
    divisor is the step each parameter's table entries really take (FP32
    emission tables: log fl32(e^(v+h)) - log fl32(e^(v-h)); FP64 transitions:
    2h)."""
    machines, pairs, (m16, gx, gy, R) = spec_rle
    say("case: spec-shape gradient vs FP64 central differences, %d RLE pairs x %d, W=%d" % (len(pairs), REF_LEN, W))
    like = fit.ModelLikelihood(machines, pairs, _spec_params(fit), rle=True, fp64_likelihood=True)
    base = _true_params(fit, m16, gx, gy, R)
    _, g = like.evaluate(base)
    say("cells per evaluation: %d" % like._expect.cells)
    h = 1e-4
    probes = [(0, ab) for ab in fit.TRANS_FROM_TO] + [(1, ij) for ij in np.ndindex(4, 4)] + \
        [(2, (i,)) for i in range(4)] + [(3, (i,)) for i in range(4)] + \
        [(4, (0, 1, 1)), (4, (1, 2, 2)), (4, (2, 3, 4)), (4, (3, 1, 2)), (4, (0, 2, 1))]
    worst = 0.0
    for k, idx in probes:
        arr = base.arrays()
        up, dn = [a.copy() for a in arr], [a.copy() for a in arr]
        up[k][idx] += h
        dn[k][idx] -= h
        step = 2 * h if k == 0 else _table_step(arr[k][idx], h)
        fd = (math.fsum(like.log_likelihood(fit.ModelParams(*up))) -
              math.fsum(like.log_likelihood(fit.ModelParams(*dn)))) / step
        rel = abs(fd - g[k][idx]) / abs(fd)
        worst = max(worst, rel)
        assert rel <= 1e-3, ((k, idx), fd, g[k][idx])
    like.close()
    say("max relative gradient error vs FP64 central differences over %d parameters: %.3g" % (len(probes), worst))


def test_spec_shape_counts_match_cpu_oracle(fit, spec_rle):
    """Expected transition counts (per pair) and emission counts (summed) of
    16 of the 4096 pairs, 8 per strand, against the FP64 oracle's (O_exact)."""
    machines, pairs, (m16, gx, gy, R) = spec_rle
    sample = pairs[:8] + pairs[N_PAIRS // 2: N_PAIRS // 2 + 8]
    say("case: spec-shape counts vs O_exact, 16 RLE pairs x %d, W=%d" % (REF_LEN, W))
    like = fit.ModelLikelihood(machines, sample, _spec_params(fit), rle=True)
    _, grads = like.evaluate(_true_params(fit, m16, gx, gy, R))
    t_gpu, _ = like._expect.expectations()
    like.close()
    op = O.Params(threshold=0.01, minDiagsBetweenTraceBack=1 << 30, traceBackDiagonals=40, diagonalExpansion=W,
                  splitMatrixBiggerThanThis=250_000_000, dynamicAnchorExpansion=True)
    rep = np.concatenate([np.log(np.where(R > 0, R, 1e-300)).reshape(-1) / fit.RL_SCALE, np.zeros(51 * 51)])
    models = {}
    for st in (0, 1):
        tab = machines.tables(st)
        models[st] = O.Model(dict(zip(fit.TRANS_NAMES, tab["trans"])), tab["match"], tab["gapX"], tab["gapY"],
                             rle=True, strand=st, repeat=rep)
    want = [np.zeros((4, 4)), np.zeros(4), np.zeros(4), np.zeros((4, 51, 51))]
    for (sx, sy, an, st), tg in zip(sample, t_gpu):
        te, _ = O.expectations(models[st], sx, sy, an, op)
        assert np.allclose(tg, te, rtol=1e-4, atol=1e-3), (tg, te)
        for acc, v in zip(want, O.emission_counts(models[st], sx, sy, an, op)):
            acc += v
    for g, w in zip(grads[1:], want):
        assert np.allclose(g, w, rtol=1e-4, atol=1e-3), (np.abs(g - w).max(), w.max())


def test_spec_shape_fit_recovers_generating_parameters(fit, tmp_path_factory):
    """200 Adam steps from a 20% perturbation of the transitions (gap states
    tied), the 4 x 4 substitution matrix and both gap emissions, on 4096
    pairs x 2 kb at W = 48: every parameter within 5% of the generating
    one."""
    from marginpolish_amd import pairhmm
    m16, gx, gy = _emissions()
    gx = np.array([0.3, 0.2, 0.2, 0.3])
    gy = np.array([0.2, 0.3, 0.3, 0.2])
    path = str(tmp_path_factory.mktemp("specfit") / "generating_model.json")
    machines = fit.generating_model(PARAMS, T_STAR, m16, gx, gy, path)
    pairs = fit.simulate_pairs(N_PAIRS, REF_LEN, T_STAR, m16, gx, gy, W, seed=41)
    say("case: Adam on transitions + substitution + gap emissions, %d pairs x %d, W=%d" % (N_PAIRS, REF_LEN, W))
    like = fit.ModelLikelihood(machines, pairs, _spec_params(fit), rle=False)
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


def test_spec_shape_fit_recovers_run_length_confusion(fit, tmp_path_factory):
    """The 51 x 51 run-length confusion with the substitution matrix and the
    gap-X emissions, on 4096 run-length pairs x 2 kb at W = 48 from a
    generating model without insertions (the run-length term scores matched
    symbols only, so only then is the model a proper distribution of the
    read's run lengths), transitions held at the generating values, 200 Adam
    steps from a 20% perturbation: substitution and gap-X within 5%,
    run-length entries with at least 5000 expected observations within 5%."""
    T_noins = np.array([[0.95, 0.05, 0.0], [0.70, 0.30, 0.0], [0.70, 0.0, 0.30]])
    m16, gx, gy = _emissions()
    path = str(tmp_path_factory.mktemp("specrl") / "generating_model.json")
    machines = fit.generating_model(PARAMS, T_noins, m16, gx, gy, path)
    R = fit.rl_confusion()
    with np.errstate(divide="ignore"):
        logR = np.log(R)
    machines.set_repeat_log10(np.maximum(logR, -690.0) / fit.RL_SCALE, True)
    pairs = _spec_rle_pairs(fit, T_noins, m16, gx, gy, R, seed=43)
    say("case: Adam on run-length confusion + substitution + gap X, %d RLE pairs x %d, W=%d, no insertions"
        % (N_PAIRS, REF_LEN, W))
    like = fit.ModelLikelihood(machines, pairs, _spec_params(fit), rle=True)
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
