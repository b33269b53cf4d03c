"""Brute-force path enumeration for tiny pair-HMM inputs (test infrastructure).

Enumerates every alignment path of the pair HMM over the full
(lX+1) x (lY+1) matrix and sums path probabilities exactly in FP64 (as
log-sum-exp over an explicit path list).  It is independent of the DP
recursions: a path is a sequence of states, each stepping by its kind --
M (1,1), gap X (1,0), gap Y (0,1); its probability is
start[s0] * prod T(s_{i-1}->s_i) * e_{s_i}(cell_i) * end[s_n] with the
boundary vectors of stateMachine.c:518-557 and the emission semantics of
stateMachine.c:559-581 (X emits sX[x-1], Y emits sY[y-1], M emits the pair).

The 3-state machine is the reference's StateMachine3.  The five-state machine
(a Model with trans5: M, short X, short Y, long X, long Y) has no reference
counterpart; its boundary vectors extend the 3-state ones: start in M, ragged
start in any gap state, end through each state's transition to M, ragged end
with M at the mean of the short-gap open logs and each gap state at its own
extension.  Use only for lX, lY <= ~6.
"""
from __future__ import annotations

import math

import numpy as np

M, X, Y = 0, 1, 2
STEP = {M: (1, 1), X: (1, 0), Y: (0, 1)}


def _trans(model):
    t = model.trans
    T = np.full((3, 3), -math.inf)
    T[M, M] = t["matchContinue"]; T[X, M] = t["matchFromGapX"]; T[Y, M] = t["matchFromGapY"]
    T[M, X] = t["gapOpenX"]; T[X, X] = t["gapExtendX"]; T[Y, X] = t["gapSwitchToX"]
    T[M, Y] = t["gapOpenY"]; T[Y, Y] = t["gapExtendY"]; T[X, Y] = t["gapSwitchToY"]
    return T


def kind(s):
    """Emission / step kind of state s: 0 match, 1 gap X, 2 gap Y (five states: 3 -> X, 4 -> Y)."""
    return 0 if s == 0 else (1 if s % 2 == 1 else 2)


def machine(model):
    """(T, start, end_vectors) of the model: T[from, to] in natural log,
    start[ragged][s], end[ragged][s]."""
    if getattr(model, "trans5", None) is not None:
        T = np.asarray(model.trans5, dtype=np.float64)
        ns = 5
        start = [[0.0] + [-math.inf] * 4, [-math.inf] + [0.0] * 4]
        end = [[T[s, 0] for s in range(ns)],
               [(T[0, 1] + T[0, 2]) / 2.0] + [T[s, s] for s in range(1, ns)]]
        return T, start, end
    T = _trans(model)
    t = model.trans
    start = [[0.0, -math.inf, -math.inf], [-math.inf, 0.0, 0.0]]
    end = [[t["matchContinue"], t["matchFromGapX"], t["matchFromGapY"]],
           [(t["gapOpenX"] + t["gapOpenY"]) / 2.0, t["gapExtendX"], t["gapExtendY"]]]
    return T, start, end


def _emit(model, state, sx, sy):
    """Log emission of a step of this kind into a cell; sx/sy are the consumed symbols."""
    if state == X:
        b = (sx & 255) if model.rle else sx
        return -1.386294361 if b >= 4 else model.gapX[b]
    if state == Y:
        b = (sy & 255) if model.rle else sy
        return -1.386294361 if b >= 4 else model.gapY[b]
    if not model.rle:
        if sx >= 4 or sy >= 4:
            return -2.772588722
        return model.match[sx * 4 + sy]
    xb, yb = sx & 255, sy & 255
    e = -2.772588722 if (xb >= 4 or yb >= 4) else model.match[xb * 4 + yb]
    b = 0 if xb == 4 else xb
    row = b if model.strand else 3 - b
    R = 51
    return e + 2.3025 * model.repeat[row * R * R + (sx >> 8) * R + (sy >> 8)]


def _lse(vals):
    vals = [v for v in vals if v != -math.inf]
    if not vals:
        return -math.inf
    m = max(vals)
    return m + math.log(sum(math.exp(v - m) for v in vals))


def enumerate_paths(model, sX, sY, ragged_left=False, ragged_right=False):
    """Return (logTotal, post) with post[s][x][y] the exact posterior of
    being in state s at matrix cell (x, y)."""
    lX, lY = len(sX), len(sY)
    T, start_v, end_v = machine(model)
    ns = T.shape[0]
    start = start_v[1 if ragged_left else 0]
    end = end_v[1 if ragged_right else 0]

    paths = []  # (logp, [(state, x, y), ...])

    def rec(x, y, s, lp, cells):
        if x == lX and y == lY:
            paths.append((lp + end[s], list(cells)))
            return
        for nxt in range(ns):
            k = kind(nxt)
            dx, dy = STEP[k]
            nx, ny = x + dx, y + dy
            if nx > lX or ny > lY:
                continue
            e = _emit(model, k, int(sX[nx - 1]) if dx else 4, int(sY[ny - 1]) if dy else 4)
            step = lp + T[s, nxt] + e
            if step == -math.inf:
                continue
            cells.append((nxt, nx, ny))
            rec(nx, ny, nxt, step, cells)
            cells.pop()

    for s0 in range(ns):
        if start[s0] != -math.inf:
            rec(0, 0, s0, start[s0], [])
    total = _lse([p for p, _ in paths])
    post = np.zeros((ns, lX + 1, lY + 1))
    for lp, cells in paths:
        w = math.exp(lp - total)
        for s, x, y in cells:
            post[s, x, y] += w
    return total, post


def traceback_windows(L, R, D, min_diags, traceback_diagonals, expansion):
    """The traceback points of getPosteriorProbsWithBanding
    (pairwiseAligner.c:715-783) as [(tb, lo, hi, at_end)]: a window fires at
    the first anti-diagonal xay >= lo + minDiagsBetweenTraceBack whose band is
    at most 2*diagonalExpansion+1 cells wide, or at the last one (D); it emits
    the posteriors of anti-diagonals in (lo, hi] with hi = tb - (traceBack + 1)
    (hi = D at the end), and the next window starts at lo = hi."""
    out, lo = [], 0
    for xay in range(1, D + 1):
        at_end = xay == D
        width = (R[xay] - L[xay]) // 2 + 1
        if at_end or (xay >= lo + min_diags and width <= 2 * expansion + 1):
            hi = D if at_end else xay - (traceback_diagonals + 1)
            out.append((xay, lo, hi, at_end))
            lo = hi
            if at_end:
                break
    return out


def enumerate_windowed(model, sX, sY, L, R, min_diags, traceback_diagonals, expansion,
                       ragged_left=False, ragged_right=False):
    """Exact windowed posteriors by path enumeration inside a band.

    L[d], R[d] bound x - y on anti-diagonal d = x + y (the band of
    pairwiseAligner.c:120-226; a path may only visit band cells).  For the
    window ending at traceback anti-diagonal tb (end vector = endStateProb, or
    raggedEndStateProb for the final window of a ragged-right alignment), the
    posterior of state s at cell (x, y) is

        sum of path weight x end[last state] over band paths from the start
            that visit (x, y) in state s and stop ON anti-diagonal tb
        / the same sum over all band paths that stop on tb

    -- a match stepping from tb-1 to tb+1 never reaches tb and is not counted,
    which is what initialising only B[tb] (and zeroing B[tb-1]) does in the
    reference (pairwiseAligner.c:729-750).  A cell gets the posterior of the
    window whose emitted range (lo, hi] holds its anti-diagonal.

    Returns (windows, totals, post) with post[s][x][y] and totals[k] the log
    of window k's path sum."""
    lX, lY = len(sX), len(sY)
    D = lX + lY
    T, start_v, end_v = machine(model)
    ns = T.shape[0]
    start = start_v[1 if ragged_left else 0]
    wins = traceback_windows(L, R, D, min_diags, traceback_diagonals, expansion)

    def in_band(x, y):
        d = x + y
        return 0 <= x <= lX and 0 <= y <= lY and L[d] <= x - y <= R[d]

    post = np.zeros((ns, lX + 1, lY + 1))
    totals = []
    for tb, lo, hi, at_end in wins:
        end = end_v[1 if (at_end and ragged_right) else 0]
        paths = []

        def rec(x, y, s, lp, cells):
            if x + y == tb:
                paths.append((lp + end[s], list(cells)))
                return
            for nxt in range(ns):
                k = kind(nxt)
                dx, dy = STEP[k]
                nx, ny = x + dx, y + dy
                if nx + ny > tb or not in_band(nx, ny):
                    continue
                e = _emit(model, k, int(sX[nx - 1]) if dx else 4, int(sY[ny - 1]) if dy else 4)
                step = lp + T[s, nxt] + e
                if step == -math.inf:
                    continue
                cells.append((nxt, nx, ny))
                rec(nx, ny, nxt, step, cells)
                cells.pop()

        for s0 in range(ns):
            if start[s0] != -math.inf:
                rec(0, 0, s0, start[s0], [])
        total = _lse([p for p, _ in paths])
        totals.append(total)
        for lp, cells in paths:
            w = math.exp(lp - total)
            for s, x, y in cells:
                if lo < x + y <= hi:
                    post[s, x, y] += w
    return wins, totals, post
