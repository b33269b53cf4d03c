"""Parameter-fitting harness for the pair-HMM transitions (outside the
polishing path).

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


def main():
    """Differentiation workload: simulate, check the gradient against central
    differences, fit from a perturbed start, print one JSON line."""
    import argparse
    import json
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
    import os
    Tstar = np.array([[0.90, 0.05, 0.05], [0.70, 0.25, 0.05], [0.70, 0.05, 0.25]])
    sub = np.full((4, 4), 0.02 / 3)
    np.fill_diagonal(sub, 0.98)
    match16 = (sub / 4).reshape(16)  # joint P(x, y): uniform reference base, 2% substitutions
    gap = np.full(4, 0.25)
    dm = generating_model(a.params, Tstar, match16, gap, gap, os.path.join(a.workdir, "generating_model.json"))
    t0 = time.time()
    pairs = simulate_pairs(a.pairs, a.ref_len, Tstar, match16, gap, gap, a.band, a.seed)
    t_sim = time.time() - t0
    p = pairhmm.PairwiseAlignmentParameters(threshold=0.01, diagonalExpansion=a.band, dynamicAnchorExpansion=True)
    # gradient vs central differences on a subset
    fdl = TransitionLikelihood(dm, pairs[: a.fd_pairs], p)
    logT = np.log(Tstar)
    _, g = fdl.evaluate(logT)
    h = 1e-3
    worst = 0.0
    for a_, b_ in TRANS_FROM_TO:
        e = np.zeros((3, 3))
        e[a_, b_] = h
        fd = (fdl.log_likelihood(logT + e).sum() - fdl.log_likelihood(logT - e).sum()) / (2 * h)
        worst = max(worst, abs(fd - g[a_, b_]) / max(abs(g[a_, b_]), 1.0))
    fdl.close()
    like = TransitionLikelihood(dm, pairs, p)
    rng = np.random.default_rng(a.seed + 1)
    Tinit = Tstar * (1.0 + a.perturb * rng.choice([-1.0, 1.0], size=(3, 3)))
    t1 = time.time()
    That, hist = fit_transitions(like, Tinit, a.steps, a.lr, symmetric=not a.asymmetric)
    t_fit = time.time() - t1
    rel = np.abs(That - Tstar) / Tstar
    print(json.dumps({
        "pairs": a.pairs, "ref_len": a.ref_len, "band_half_width": a.band, "cells": int(like._expect.cells),
        "simulate_s": round(t_sim, 2), "fd_pairs": a.fd_pairs, "fd_h": h,
        "max_rel_grad_vs_central_difference": worst,
        "adam_steps": a.steps, "fit_s": round(t_fit, 2), "evaluations": like.evaluations,
        "T_generating": Tstar.round(4).tolist(), "T_start": Tinit.round(4).tolist(), "T_fit": That.round(4).tolist(),
        "max_rel_err_fit_vs_generating": float(rel.max()),
        "mean_ll_per_pair_first_last": [hist[0], hist[-1]],
    }), flush=True)
    like.close()


if __name__ == "__main__":
    main()
