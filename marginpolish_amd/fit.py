"""Parameter-fitting harness for the pair-HMM (outside the polishing path):
transitions, and with `ModelLikelihood` also the substitution (joint match),
gap and run-length emission parameters.

The gradient of a pair's log-likelihood with respect to each natural-log
transition probability is that transition's expected count over all alignment
paths, which the engine computes on the GPU in expectation mode
(getExpectationsUsingAnchors, pairwiseAligner.c:1151-1164, with
cell_calculateExpectation at 347-369).  The reference computes transition
counts only (emission counts are disabled at 359-362); this harness fits the
same nine transitions.  The end-state vector is pinned while fitting
(`StateMachines.set_transitions(..., pin_end=True)`), so the likelihood's
derivative is exactly the expected count, start transition included.

`PairHMMLogLikelihood` wraps one evaluation as a `torch.autograd.Function`;
`fit_transitions` runs Adam on per-state logits (log-softmax rows).  The
emission gradients are the expected emission counts (the posterior of each
state at each cell, binned by the parameter it reads), computed by the same
expectation-mode kernel; the reference has no counterpart (its emission
expectations are disabled), so they are checked against central differences,
the CPU oracle's counts and path enumeration.  `ModelLikelihood` /
`fit_model` cover every parameter; `simulate_rle_pairs` samples run-length
symbols through a 51 x 51 run-length confusion per base.  The
likelihood itself is the batched posterior kernel's total (records disabled by
a threshold above 1); the counts come from the expectation-mode kernel on the
same bands.  `simulate_pairs` samples read/reference pairs from a 3-state
pair-HMM (the generating model of the differentiation workload)."""
import dataclasses

import numpy as np

from . import _lib, pairhmm

M, X, Y = 0, 1, 2
# names of mp_model_tables' trans9 order and their (from, to) states
TRANS_NAMES = ("matchContinue", "matchFromGapX", "matchFromGapY", "gapOpenX", "gapOpenY", "gapExtendX",
               "gapExtendY", "gapSwitchToX", "gapSwitchToY")
TRANS_FROM_TO = ((M, M), (X, M), (Y, M), (M, X), (M, Y), (X, X), (Y, Y), (Y, X), (X, Y))


def named_from_matrix(logT):
    """3x3 natural-log transition matrix [from, to] -> trans9 order."""
    logT = np.asarray(logT, dtype=np.float64)
    return [float(logT[a, b]) for a, b in TRANS_FROM_TO]


def matrix_from_named(t9):
    out = np.full((3, 3), -np.inf)
    for v, (a, b) in zip(t9, TRANS_FROM_TO):
        out[a, b] = v
    return out


class TransitionLikelihood:
    """Sum of log-likelihoods of a fixed set of pairs as a function of the
    3x3 log-transition matrix, with its gradient (the expected counts)."""

    def __init__(self, machines, pairs, p, engine=None):
        self.machines = machines
        self.pairs = list(pairs)
        # one traceback window per pair: the expected counts are then those of
        # the full forward-backward, the exact gradient of the total
        p = dataclasses.replace(p, minDiagsBetweenTraceBack=1 << 30)
        like_p = dataclasses.replace(p, threshold=2.0)  # no posterior records
        self._like = pairhmm.Batch(machines, self.pairs, like_p, mode=_lib.MODE_MATCH_ONLY, engine=engine)
        self._expect = pairhmm.Batch(machines, self.pairs, p, mode=_lib.MODE_EXPECTATIONS, engine=engine)
        self.evaluations = 0

    def _set(self, logT):
        self.machines.set_transitions(named_from_matrix(logT), pin_end=True)
        self._like.set_model(self.machines)
        self._expect.set_model(self.machines)

    def log_likelihood(self, logT):
        """Per-pair log-likelihoods (FP64 host sum of the FP32 kernel totals)."""
        self._set(logT)
        self._like.run()
        _, lp = self._like.counts()
        self.evaluations += 1
        return lp

    def evaluate(self, logT):
        """(sum of log-likelihoods, 3x3 expected transition counts)."""
        lp = self.log_likelihood(logT)
        self._expect.run()
        t, _ = self._expect.expectations()
        return float(np.sum(lp)), t.sum(axis=0).reshape(3, 3)

    def close(self):
        self._like.close()
        self._expect.close()


def _torch_function():
    import torch

    class PairHMMLogLikelihood(torch.autograd.Function):
        """LL(log T) with dLL/dlog T = expected transition counts."""

        @staticmethod
        def forward(ctx, logT, like):
            ll, counts = like.evaluate(logT.detach().cpu().numpy())
            ctx.save_for_backward(torch.as_tensor(counts, dtype=logT.dtype, device=logT.device))
            return logT.new_tensor(ll)

        @staticmethod
        def backward(ctx, grad):
            (counts,) = ctx.saved_tensors
            return grad * counts, None

    return PairHMMLogLikelihood


def log_likelihood(logT, like):
    """Differentiable sum of log-likelihoods (torch tensor logT, 3x3)."""
    return _torch_function().apply(logT, like)


def fit_transitions(like, T0, steps=200, lr=0.05, symmetric=False):
    """Adam on per-state logits from the 3x3 probability matrix T0 (rows need
    not be normalised; they are renormalised).  symmetric ties the gap states
    as the reference's symmetric machine does (stateMachine3_loadSymmetric,
    stateMachine.c:649-668): M->X = M->Y, X->M = Y->M, X->X = Y->Y and
    X->Y = Y->X.  Returns (T, history of mean log-likelihood per pair)."""
    import torch
    L0 = np.log(np.asarray(T0, dtype=np.float64))
    if symmetric:
        theta = torch.tensor([L0[0, 0], (L0[0, 1] + L0[0, 2]) / 2, (L0[1, 0] + L0[2, 0]) / 2,
                              (L0[1, 1] + L0[2, 2]) / 2, (L0[1, 2] + L0[2, 1]) / 2], requires_grad=True)

        def logits():
            mm, op, fg, ex, sw = theta
            return torch.stack([torch.stack([mm, op, op]), torch.stack([fg, ex, sw]), torch.stack([fg, sw, ex])])
    else:
        theta = torch.tensor(L0, requires_grad=True)

        def logits():
            return theta
    opt = torch.optim.Adam([theta], lr=lr)
    n = max(1, len(like.pairs))
    hist = []
    for _ in range(steps):
        opt.zero_grad()
        loss = -log_likelihood(torch.log_softmax(logits(), dim=1), like) / n
        loss.backward()
        opt.step()
        hist.append(-float(loss.detach()))
    with torch.no_grad():
        return torch.softmax(logits(), dim=1).numpy(), hist


def generating_model(template_json, T, match16, gapX, gapY, path):
    """Write a copy of a params file (template_json) whose forward-strand hmm
    is the asymmetric 3-state model (type 3) with transition probabilities T
    ([from, to]) and emission probabilities match16 / gapX / gapY, without the
    run-length term, to `path`, and load it.  The reverse strand's tables are
    the reverse complement, as parser.c:241-243 builds them."""
    import json
    with open(template_json) as fh:
        doc = json.load(fh)
    pol = doc["polish"] if "polish" in doc else doc
    pol["hmmForwardStrandReadGivenReference"] = {
        "type": 3, "emissionsType": 0,
        "transitions": [float(v) for v in np.asarray(T, dtype=np.float64).reshape(9)],
        "emissions": [float(v) for v in list(match16) + list(gapX) + list(gapY)]}
    pol["useRunLengthEncoding"] = False
    pol["useRepeatCountsInAlignment"] = False
    with open(path, "w") as fh:
        json.dump(doc, fh)
    return pairhmm.StateMachines.from_json(path)


def _normalise(v):
    v = np.asarray(v, dtype=np.float64)
    return v / v.sum()


def simulate_pairs(n, ref_len, T, match16, gapX, gapY, W, seed=0):
    """Sample n read/reference pairs from a 3-state pair-HMM: a state path from
    the match state under transition matrix T (probabilities, [from, to]); M
    emits an (x, y) base pair from match16 (4x4, renormalised), X an x base from
    gapX, Y a y base from gapY; a path ends when its reference reaches ref_len
    bases.  Anchors = every match cell (0-based), expansion W.  Symbols carry no
    run length (a model without the run-length term).  All pairs advance
    together, one state per step."""
    rng = np.random.default_rng(seed)
    T = np.asarray(T, dtype=np.float64)
    cumT = np.cumsum(T / T.sum(axis=1, keepdims=True), axis=1)
    cm, cx, cy = (np.cumsum(_normalise(v)) for v in (match16, gapX, gapY))
    state = np.zeros(n, dtype=np.int64)
    xcount = np.zeros(n, dtype=np.int64)
    S, B1, B2 = [], [], []
    while (xcount < ref_len).any():
        live = xcount < ref_len
        ns = np.minimum((rng.random(n)[:, None] >= cumT[state]).sum(axis=1), 2)
        km = np.minimum(np.searchsorted(cm, rng.random(n), side="right"), 15)
        bx = np.minimum(np.searchsorted(cx, rng.random(n), side="right"), 3)
        by = np.minimum(np.searchsorted(cy, rng.random(n), side="right"), 3)
        S.append(np.where(live, ns, -1))
        B1.append(np.where(ns == M, km // 4, bx))
        B2.append(np.where(ns == M, km % 4, by))
        xcount += live & (ns != Y)
        state = np.where(live, ns, state)
    S, B1, B2 = np.stack(S, 1), np.stack(B1, 1), np.stack(B2, 1)
    out = []
    for i in range(n):
        s = S[i]
        s = s[s >= 0]
        b1, b2 = B1[i, : s.size], B2[i, : s.size]
        hx, hy = s != Y, s != X
        xs, ys = b1[hx], b2[hy]
        if ys.size == 0:
            ys = np.array([0])
        ix, iy = np.cumsum(hx) - 1, np.cumsum(hy) - 1
        mm = s == M
        an = np.stack([ix[mm], iy[mm], np.full(int(mm.sum()), W)], axis=1).astype(np.int64).reshape(-1, 3)
        out.append((xs.astype(np.uint16), ys.astype(np.uint16), an, 1))
    return out


RL_SCALE = 2.3025  # the match emission's run-length factor is exp(2.3025 * log10 R) (stateMachine.c:708-722)


@dataclasses.dataclass
class ModelParams:
    """Natural-log parameters: transitions T[from, to] (3x3), joint match
    emission [x base, y base] (4x4), gap X (4), gap Y (4), and the run-length
    term rho[base, underlying, observed] = 2.3025 * log10 R (4x51x51; the
    factor the match emission multiplies by is exp(rho))."""
    logT: np.ndarray
    log_match: np.ndarray
    log_gap_x: np.ndarray
    log_gap_y: np.ndarray
    rho: np.ndarray = None

    def arrays(self):
        return [np.asarray(v, dtype=np.float64) for v in (self.logT, self.log_match, self.log_gap_x, self.log_gap_y,
                                                         self.rho if self.rho is not None else np.zeros((4, 51, 51)))]


class ModelLikelihood:
    """Sum of log-likelihoods of a fixed set of pairs as a function of all
    parameters (ModelParams), with the gradient in each natural-log
    parameter: expected transition counts and expected emission counts."""

    def __init__(self, machines, pairs, p, rle=False, engine=None, fp64_likelihood=False):
        """fp64_likelihood: log_likelihood() runs the FP64 kernel (the
        emission tables stay FP32), e.g. for finite differences; the
        gradient always comes from the FP32 expectation kernel."""
        self.machines = machines
        self.pairs = list(pairs)
        self.rle = bool(rle)
        p = dataclasses.replace(p, minDiagsBetweenTraceBack=1 << 30)
        like_p = dataclasses.replace(p, threshold=2.0)
        self._like = pairhmm.Batch(machines, self.pairs, like_p, mode=_lib.MODE_MATCH_ONLY, engine=engine,
                                   fp64=fp64_likelihood)
        self._expect = pairhmm.Batch(machines, self.pairs, p, mode=_lib.MODE_EXPECTATIONS, engine=engine)
        self.evaluations = 0

    def _set(self, prm):
        logT, lm, lgx, lgy, rho = prm.arrays()
        self.machines.set_transitions(named_from_matrix(logT), pin_end=True)
        self.machines.set_emissions(lm.reshape(16), lgx, lgy)
        if self.rle:
            self.machines.set_repeat_log10(np.minimum(rho / RL_SCALE, 0.0), True)
        self._like.set_model(self.machines)
        self._expect.set_model(self.machines)

    def log_likelihood(self, prm):
        self._set(prm)
        self._like.run()
        _, lp = self._like.counts()
        self.evaluations += 1
        return lp

    def evaluate(self, prm):
        """(sum of log-likelihoods, [dT 3x3, dmatch 4x4, dgapX 4, dgapY 4, drho 4x51x51])."""
        lp = self.log_likelihood(prm)
        self._expect.run()
        t, _ = self._expect.expectations()
        m, gx, gy, rl = self._expect.emission_counts()
        return float(np.sum(lp)), [t.sum(axis=0).reshape(3, 3), m, gx, gy, rl if self.rle else np.zeros_like(rl)]

    def close(self):
        self._like.close()
        self._expect.close()


def _torch_model_function():
    import torch

    class PairHMMModelLogLikelihood(torch.autograd.Function):
        """LL(params) with dLL/d(natural-log parameter) = expected counts."""

        @staticmethod
        def forward(ctx, logT, lm, lgx, lgy, rho, like):
            prm = ModelParams(*(v.detach().cpu().numpy() for v in (logT, lm, lgx, lgy, rho)))
            ll, grads = like.evaluate(prm)
            ctx.save_for_backward(*(torch.as_tensor(g, dtype=logT.dtype) for g in grads))
            return logT.new_tensor(ll)

        @staticmethod
        def backward(ctx, grad):
            return tuple(grad * g for g in ctx.saved_tensors) + (None,)

    return PairHMMModelLogLikelihood


def model_log_likelihood(logT, lm, lgx, lgy, rho, like):
    """Differentiable sum of log-likelihoods (torch tensors)."""
    return _torch_model_function().apply(logT, lm, lgx, lgy, rho, like)


FIT_ALL = ("T", "match", "gapX", "gapY", "rl")


def fit_model(like, init, steps=200, lr=0.05, symmetric=True, fit=FIT_ALL):
    """Adam on logits of the distributions named in `fit`: transition rows
    "T" (gap states tied when symmetric), the joint match emission "match"
    (16), "gapX" (4), "gapY" (4) and, with run-length symbols, each
    run-length row "rl" rho[base, underlying, :] (51); the others stay at
    init.  init: ModelParams of probabilities' logs (rows need not be
    normalised).  Returns (ModelParams of the fit, history of mean
    log-likelihood per pair).

    The run-length term scores only matched symbols: an inserted read
    symbol's run length contributes a factor 1, so with run-length symbols
    the model is a proper distribution of the data (and its maximum
    likelihood consistent) only when no read symbol is inserted."""
    import torch
    L0, M0, GX0, GY0, R0 = init.arrays()
    if symmetric:
        th_t = torch.tensor([L0[0, 0], (L0[0, 1] + L0[0, 2]) / 2, (L0[1, 0] + L0[2, 0]) / 2,
                             (L0[1, 1] + L0[2, 2]) / 2, (L0[1, 2] + L0[2, 1]) / 2])

        def t_logits():
            mm, op, fg, ex, sw = th_t
            return torch.stack([torch.stack([mm, op, op]), torch.stack([fg, ex, sw]), torch.stack([fg, sw, ex])])
    else:
        th_t = torch.tensor(L0)

        def t_logits():
            return th_t
    th_t.requires_grad_("T" in fit)
    th_m = torch.tensor(M0.reshape(16), requires_grad="match" in fit)
    th_gx = torch.tensor(GX0, requires_grad="gapX" in fit)
    th_gy = torch.tensor(GY0, requires_grad="gapY" in fit)
    th_r = torch.tensor(np.maximum(R0, -30.0), requires_grad=bool("rl" in fit and like.rle))
    params = [t for t in (th_t, th_m, th_gx, th_gy, th_r) if t.requires_grad]
    opt = torch.optim.Adam(params, lr=lr)
    n = max(1, len(like.pairs))

    def current():
        return (torch.log_softmax(t_logits(), dim=1), torch.log_softmax(th_m, dim=0).reshape(4, 4),
                torch.log_softmax(th_gx, dim=0), torch.log_softmax(th_gy, dim=0), torch.log_softmax(th_r, dim=2))
    hist = []
    for _ in range(steps):
        opt.zero_grad()
        loss = -model_log_likelihood(*current(), like) / n
        loss.backward()
        opt.step()
        hist.append(-float(loss.detach()))
    with torch.no_grad():
        return ModelParams(*(v.numpy() for v in current())), hist


def rl_confusion(p_same=0.8, p1=0.08, p2=0.02):
    """A 51 x 51 run-length confusion per base: observed = underlying with
    p_same, +-1 with p1 each, +-2 with p2 each (observed >= 1), rows
    normalised; row 0 is uniform (no run has length 0)."""
    R = np.zeros((51, 51))
    R[0] = 1.0 / 51
    for u in range(1, 51):
        for o, pr in ((u, p_same), (u - 1, p1), (u + 1, p1), (u - 2, p2), (u + 2, p2)):
            if 1 <= o <= 50:
                R[u, o] += pr
        R[u] /= R[u].sum()
    return np.stack([R] * 4)


def simulate_rle_pairs(n, ref_len, T, match16, gapX, gapY, R, W, seed=0, p_run=0.6, strand=1):
    """simulate_pairs with run-length symbols: every emitted x symbol takes a
    run length ~ Geometric(p_run) (at most 50); a match's y run length is drawn
    from R[row, underlying x run length, :] (row = the x base on the forward
    strand, 3 - base on the reverse, as the repeat table is read); a gap-Y y
    symbol takes a Geometric(p_run) run length."""
    pairs = simulate_pairs(n, ref_len, T, match16, gapX, gapY, W, seed)
    rng = np.random.default_rng(seed + 7919)
    cumR = np.cumsum(R / R.sum(axis=2, keepdims=True), axis=2)
    out = []
    for sx, sy, an, _ in pairs:
        rx = np.minimum(rng.geometric(p_run, size=sx.size), 50)
        ry = np.minimum(rng.geometric(p_run, size=sy.size), 50)
        mx, my = an[:, 0], an[:, 1]
        b = sx[mx].astype(np.int64)
        row = b if strand else 3 - b
        u = rx[mx]
        ry[my] = np.minimum((rng.random(mx.size)[:, None] >= cumR[row, u]).sum(axis=1), 50)
        X = ((rx.astype(np.uint32) << 8) | sx.astype(np.uint32)).astype(np.uint16)
        Yv = ((ry.astype(np.uint32) << 8) | sy.astype(np.uint32)).astype(np.uint16)
        out.append((X, Yv, an, strand))
    return out


def main():
    """Differentiation workload: simulate from the generating model, check
    the gradient (transitions and emissions) against central differences,
    fit from a perturbed start; then the run-length confusion on run-length
    symbols without insertions.  Prints one JSON line."""
    import argparse
    import json
    import os
    import time
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=4096)
    ap.add_argument("--ref-len", type=int, default=2000)
    ap.add_argument("--band", type=int, default=48)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--lr", type=float, default=0.05)
    ap.add_argument("--perturb", type=float, default=0.2)
    ap.add_argument("--fd-pairs", type=int, default=64)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--asymmetric", action="store_true", help="fit all nine transitions (default: tied gap states)")
    ap.add_argument("--params", required=True, help="params JSON used as the template of the generating model")
    ap.add_argument("--workdir", default=".")
    a = ap.parse_args()
    Tstar = np.array([[0.90, 0.05, 0.05], [0.70, 0.25, 0.05], [0.70, 0.05, 0.25]])
    sub = np.full((4, 4), 0.02 / 3)
    np.fill_diagonal(sub, 0.98)
    match16 = (sub / 4).reshape(16)  # joint P(x, y): uniform reference base, 2% substitutions
    gx = np.array([0.3, 0.2, 0.2, 0.3])
    gy = np.array([0.2, 0.3, 0.3, 0.2])
    dm = generating_model(a.params, Tstar, match16, gx, gy, os.path.join(a.workdir, "generating_model.json"))
    t0 = time.time()
    pairs = simulate_pairs(a.pairs, a.ref_len, Tstar, match16, gx, gy, a.band, a.seed)
    t_sim = time.time() - t0
    p = pairhmm.PairwiseAlignmentParameters(threshold=0.01, diagonalExpansion=a.band, dynamicAnchorExpansion=True)
    truth = ModelParams(np.log(Tstar), np.log(match16).reshape(4, 4), np.log(gx), np.log(gy))
    # gradient vs central differences on a subset: every transition, every emission parameter
    fdl = ModelLikelihood(dm, pairs[: a.fd_pairs], p)
    _, g = fdl.evaluate(truth)
    h = 1e-3
    worst = {}
    for k, name, shape in ((0, "transitions", None), (1, "match", (4, 4)), (2, "gapX", (4,)), (3, "gapY", (4,))):
        idxs = [ab for ab in TRANS_FROM_TO] if k == 0 else list(np.ndindex(*shape))
        w = 0.0
        for idx in idxs:
            up, dn = [x.copy() for x in truth.arrays()], [x.copy() for x in truth.arrays()]
            up[k][idx] += h
            dn[k][idx] -= h
            fd = (fdl.log_likelihood(ModelParams(*up)).sum() - fdl.log_likelihood(ModelParams(*dn)).sum()) / (2 * h)
            w = max(w, abs(fd - g[k][idx]) / max(abs(g[k][idx]), 1.0))
        worst[name] = w
    fdl.close()
    like = ModelLikelihood(dm, pairs, p)
    rng = np.random.default_rng(a.seed + 1)
    pert = lambda v: v + np.log1p(a.perturb * rng.choice([-1.0, 1.0], size=np.shape(v)))
    init = ModelParams(*(pert(v) for v in truth.arrays()[:4]))
    t1 = time.time()
    fitted, hist = fit_model(like, init, a.steps, a.lr, symmetric=not a.asymmetric,
                             fit=("T", "match", "gapX", "gapY"))
    t_fit = time.time() - t1
    rel = {"transitions": np.abs(np.exp(fitted.logT) - Tstar) / Tstar,
           "match": np.abs(np.exp(fitted.log_match).reshape(16) - match16) / match16,
           "gapX": np.abs(np.exp(fitted.log_gap_x) - gx) / gx, "gapY": np.abs(np.exp(fitted.log_gap_y) - gy) / gy}
    cells = int(like._expect.cells)
    evals = like.evaluations
    like.close()
    print(json.dumps({
        "pairs": a.pairs, "ref_len": a.ref_len, "band_half_width": a.band, "cells": cells,
        "simulate_s": round(t_sim, 2), "fd_pairs": a.fd_pairs, "fd_h": h,
        "max_rel_grad_vs_central_difference": worst,
        "adam_steps": a.steps, "fit_s": round(t_fit, 2), "evaluations": evals,
        "T_generating": Tstar.round(4).tolist(), "T_fit": np.exp(fitted.logT).round(4).tolist(),
        "max_rel_err_fit_vs_generating": {k: float(v.max()) for k, v in rel.items()},
        "mean_ll_per_pair_first_last": [hist[0], hist[-1]],
    }), flush=True)


if __name__ == "__main__":
    main()
