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
