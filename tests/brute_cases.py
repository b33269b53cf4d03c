"""Tiny multi-window cases for the windowed path enumeration (test
infrastructure shared by the oracle pins and the GPU parity tests).

Parameters are chosen so that a pair of at most 5 x 5 symbols has several
traceback windows: traceBackDiagonals = 1, minDiagsBetweenTraceBack 3-4
(the reference requires traceBackDiagonals + 1 < minDiagsBetweenTraceBack,
pairwiseAligner.c:697), diagonalExpansion 2 or 4, random anchors (dynamic or
static band) or none, and N symbols (base 4) in both sequences.
"""
import random

import numpy as np

from oracle import brute
from oracle import oracle as O


def rand_syms(rng, n, rle, p_n=0.0):
    bases = "".join("N" if rng.random() < p_n else rng.choice("ACGT") for _ in range(n))
    if not rle:
        return O.symbols(bases)
    return O.rle_symbols(bases, [rng.choice([1, 1, 2, 3, 4, 7, 12, 50]) for _ in range(n)])


def rand_anchors(rng, lX, lY):
    anchors, x, y = [], -1, -1
    while True:
        x += rng.randint(1, 3)
        y += rng.randint(1, 3)
        if x >= lX or y >= lY:
            return anchors
        anchors.append((x, y, 2 * rng.randint(0, 2)))


def windowed_cases(seed, count, model, p_n=0.2, min_windows=2):
    """[(sx, sy, anchors, O.Params, ragged_left, ragged_right)], each with at
    least `min_windows` traceback windows."""
    rng = random.Random(seed)
    out = []
    while len(out) < count:
        lX, lY = rng.randint(2, 5), rng.randint(2, 5)
        sx = rand_syms(rng, lX, model.rle, p_n)
        sy = rand_syms(rng, lY, model.rle, p_n)
        anchors = rand_anchors(rng, lX, lY) if rng.random() < 0.6 else []
        p = O.Params(threshold=0.0, minDiagsBetweenTraceBack=rng.choice([3, 4]), traceBackDiagonals=1,
                     diagonalExpansion=rng.choice([2, 4]), splitMatrixBiggerThanThis=10 ** 9,
                     dynamicAnchorExpansion=rng.random() < 0.5)
        L, R = O.band(anchors, lX, lY, p.diagonalExpansion, p.dynamicAnchorExpansion)
        wins = brute.traceback_windows(L, R, lX + lY, p.minDiagsBetweenTraceBack, p.traceBackDiagonals,
                                       p.diagonalExpansion)
        if len(wins) < min_windows:
            continue
        k = len(out)
        out.append((sx, sy, anchors, p, bool(k & 1), bool(k & 2)))
    return out


def brute_posteriors(model, case):
    """(windows, totals, post) of the windowed enumeration of one case."""
    sx, sy, anchors, p, rl, rr = case
    L, R = O.band(anchors, len(sx), len(sy), p.diagonalExpansion, p.dynamicAnchorExpansion)
    return brute.enumerate_windowed(model, sx, sy, L, R, p.minDiagsBetweenTraceBack, p.traceBackDiagonals,
                                    p.diagonalExpansion, rl, rr)


def emitted(post, st):
    """The posterior array of state st restricted to the cells the reference
    emits for it: M needs x, y > 0, X needs x > 0, Y needs y > 0
    (pairwiseAligner.c:634-660)."""
    want = np.minimum(post[st], 1.0).copy()
    if st == 0:
        want[0, :] = 0
        want[:, 0] = 0
    elif st == 1:
        want[0, :] = 0
    else:
        want[:, 0] = 0
    return want


def dense(rows, shape):
    """(weight, x, y) rows of one list -> dense posterior array [x+1][y+1]."""
    got = np.zeros(shape)
    for r in rows:
        w, x, y = int(r[0]), int(r[1]), int(r[2])
        assert got[x + 1, y + 1] == 0, ("duplicate", x, y)
        got[x + 1, y + 1] = w / 1e7
    return got
