"""Pins of the oracle's five-state machine (M, short X, short Y, long X,
long Y), which has no counterpart in the reference (its StateMachineType has
the 3-state types only, stateMachine.h:96-99):

* exact path enumeration on tiny inputs, every cell-state, four ragged
  combinations, the canonical long-gap pattern and a dense random 5x5 matrix;
* reduction to the reference-pinned 3-state oracle when the long gaps are
  unreachable (long-gap open 0), through windows, anchors, splits and ragged
  right ends;
* a single-window global alignment emits every x exactly once in M or a gap-X
  state and every y once in M or a gap-Y state (threshold 0: the per-x and
  per-y posterior sums are 1), with the forward-only total equal to the
  posterior run's total;
* the host C ABI's five-state model: JSON "type": 5, set/get of the 25
  transitions, the short-gap slots of mp_model_tables, and refusals.
"""
import json
import math
import os
import random

import numpy as np
import pytest

from oracle import brute
from oracle import oracle as O
from tests.conftest import PARAMS


def five(m3, long_open=0.02, long_extend=0.9):
    """A five-state oracle model: short gaps = m3's gaps, long gaps added."""
    from marginpolish_amd.pairhmm import with_long_gaps
    t = m3.trans
    t9 = [t["matchContinue"], t["matchFromGapX"], t["matchFromGapY"], t["gapOpenX"], t["gapOpenY"],
          t["gapExtendX"], t["gapExtendY"], t["gapSwitchToX"], t["gapSwitchToY"]]
    T = with_long_gaps(t9, long_open, long_extend)
    return O.Model(m3.trans, m3.match, m3.gapX, m3.gapY, m3.rle, m3.strand, m3.repeat, trans5=T)


def dense_random(m3, rng):
    """Every transition allowed, rows normalised (the recursions are generic)."""
    P = np.array([[rng.uniform(0.05, 1.0) for _ in range(5)] for _ in range(5)])
    P /= P.sum(1, keepdims=True)
    return O.Model(m3.trans, m3.match, m3.gapX, m3.gapY, m3.rle, m3.strand, m3.repeat, trans5=np.log(P))


def _models(rng):
    pp = O.PolishParams(PARAMS)
    d = O.default_nucleotide_model()
    return [("default+long", five(d, 0.05, 0.8)), ("rle_fwd+long", five(pp.forward, 0.05, 0.8)),
            ("rle_rev+long", five(pp.reverse, 0.05, 0.8)), ("dense", dense_random(d, rng))]


def _rand_syms(rng, n, rle):
    bases = "".join(rng.choice("ACGT") for _ in range(n))
    if not rle:
        return O.symbols(bases)
    return O.rle_symbols(bases, [rng.choice([1, 1, 1, 2, 2, 3, 5, 9]) for _ in range(n)])


def _dense(lst, lX, lY):
    got = np.zeros((lX + 1, lY + 1))
    for w, x, y in lst:
        got[x + 1, y + 1] = w / 1e7
    return got


@pytest.mark.parametrize("ragged", [(False, False), (True, True), (True, False), (False, True)])
def test_five_state_brute_force(ragged):
    rng = random.Random(71 + 2 * ragged[0] + ragged[1])
    p = O.Params(threshold=0.0, diagonalExpansion=20, minDiagsBetweenTraceBack=1000, dynamicAnchorExpansion=True)
    for name, m in _models(rng):
        for _ in range(5):
            lX, lY = rng.randint(1, 5), rng.randint(1, 5)
            sx, sy = _rand_syms(rng, lX, m.rle), _rand_syms(rng, lY, m.rle)
            total, post = brute.enumerate_paths(m, sx, sy, *ragged)
            assert post.shape[0] == 5
            mm, gx, gy, tot, _ = O.posteriors(m, sx, sy, [], p, *ragged, mode=O.LOGADD_EXACT)
            assert abs(tot - total) < 1e-9, name
            assert abs(O.forward_prob(m, sx, sy, [], p, *ragged, mode=O.LOGADD_EXACT) - total) < 1e-9, name
            wants = (post[0], post[1] + post[3], post[2] + post[4])
            for st, lst in enumerate((mm, gx, gy)):
                want = np.floor(np.minimum(wants[st], 1.0) * 1e7) / 1e7
                if st == 0:
                    want[0, :] = 0; want[:, 0] = 0
                elif st == 1:
                    want[0, :] = 0
                else:
                    want[:, 0] = 0
                assert np.max(np.abs(_dense(lst, lX, lY) - want)) <= 1.01e-7, (name, st)


def test_brute_force_enumeration_distinguishes_long_gaps():
    """The enumeration itself: with the long-gap open 0 the five-state total
    and posteriors equal the 3-state ones of the same short-gap machine;
    with long gaps reachable, the long-gap states carry mass and the total
    moves."""
    m3 = O.default_nucleotide_model()
    sx, sy = O.symbols("ACGTA"), O.symbols("ACTA")
    t3, _ = brute.enumerate_paths(m3, sx, sy)
    t5, post5 = brute.enumerate_paths(five(m3, 0.0, 0.9), sx, sy)
    assert abs(t5 - t3) < 1e-12 and post5[3:].max() == 0.0
    t5b, post5b = brute.enumerate_paths(five(m3, 0.05, 0.9), sx, sy)
    assert post5b[3:].max() > 0.0 and t5b != t3


def test_five_state_reduces_to_three_state_through_windows():
    """Long-gap open 0: the five-state oracle equals the 3-state oracle
    (itself pinned by the reference's tests) on multi-window, anchored,
    split, ragged-right problems."""
    rng = np.random.default_rng(5)
    pp = O.PolishParams(PARAMS)
    for m3 in (pp.forward, pp.reverse):
        m5 = five(m3, 0.0, 0.9)
        for trial in range(3):
            n = 300
            bases = "".join(rng.choice(list("ACGT"), n))
            runs = rng.choice([1, 1, 2, 3, 5], n)
            sx = O.rle_symbols(bases, runs)
            keep = rng.random(n) > 0.08
            sy = O.rle_symbols("".join(b for b, k in zip(bases, keep) if k), [r for r, k in zip(runs, keep) if k])
            ys = np.cumsum(keep) - 1
            anchors = [(x, int(ys[x]), 10) for x in range(0, n, 7) if keep[x]]
            p = O.Params(threshold=0.01, diagonalExpansion=10, minDiagsBetweenTraceBack=120, traceBackDiagonals=20,
                         splitMatrixBiggerThanThis=150 * 150 if trial == 2 else 250_000_000,
                         dynamicAnchorExpansion=True)
            for rr in (False, True):
                a = O.posteriors(m3, sx, sy, anchors, p, False, rr, mode=O.LOGADD_EXACT)
                b = O.posteriors(m5, sx, sy, anchors, p, False, rr, mode=O.LOGADD_EXACT)
                assert abs(a[3] - b[3]) < 1e-9 * max(1.0, abs(a[3]))
                assert a[4] == b[4]
                for la, lb in zip(a[:3], b[:3]):
                    assert la.shape == lb.shape
                    assert np.array_equal(la[:, 1:], lb[:, 1:])
                    assert np.max(np.abs(la[:, 0] - lb[:, 0]), initial=0) <= 1


def test_five_state_emission_sums_single_window():
    """Every x >= 1 is emitted once in M or a gap-X state, every y >= 1 once
    in M or a gap-Y state, so at threshold 0 the posterior sums are 1 (less
    the floor of each weight to 1e-7)."""
    rng = random.Random(9)
    d = O.default_nucleotide_model()
    for m in (five(d, 0.01, 0.95), dense_random(d, rng)):
        for _ in range(3):
            lX, lY = rng.randint(25, 45), rng.randint(25, 45)
            sx, sy = _rand_syms(rng, lX, False), _rand_syms(rng, lY, False)
            p = O.Params(threshold=0.0, diagonalExpansion=200, minDiagsBetweenTraceBack=100000)
            mm, gx, gy, tot, _ = O.posteriors(m, sx, sy, [], p, mode=O.LOGADD_EXACT)
            M, X, Y = _dense(mm, lX, lY), _dense(gx, lX, lY), _dense(gy, lX, lY)
            sx_sum = (M + X).sum(1)[1:]
            sy_sum = (M + Y).sum(0)[1:]
            slack = (lY + 1) * 2e-7
            assert np.all(np.abs(sx_sum - 1.0) <= slack), sx_sum
            assert np.all(np.abs(sy_sum - 1.0) <= (lX + 1) * 2e-7), sy_sum
            assert abs(O.forward_prob(m, sx, sy, [], p, mode=O.LOGADD_EXACT) - tot) < 1e-9


def test_five_state_expectations_refused():
    m = five(O.default_nucleotide_model())
    with pytest.raises(ValueError):
        O.expectations(m, O.symbols("ACGT"), O.symbols("ACG"), [], O.Params())


# ---------------------------------------------------------------------------
# host C ABI (no device needed)
# ---------------------------------------------------------------------------

def _five_state_json(tmp_path, P):
    with open(PARAMS) as fh:
        doc = json.load(fh)
    pol = doc["polish"] if "polish" in doc else doc
    h = pol["hmmForwardStrandReadGivenReference"]
    pol["hmmForwardStrandReadGivenReference"] = {"type": 5, "emissionsType": h["emissionsType"],
                                                 "transitions": [float(v) for v in P.reshape(25)],
                                                 "emissions": h["emissions"]}
    path = os.path.join(str(tmp_path), "five_state.json")
    with open(path, "w") as fh:
        json.dump(doc, fh)
    return path


def test_host_five_state_model(tmp_path):
    from marginpolish_amd import pairhmm as ph
    from marginpolish_amd._lib import MPError
    rng = np.random.default_rng(2)
    P = rng.uniform(0.01, 1.0, (5, 5))
    P[1, 3:] = P[2, 3:] = 0.0
    P /= P.sum(1, keepdims=True)
    m = ph.StateMachines.from_json(_five_state_json(tmp_path, P))
    assert m.num_states == 5
    assert np.allclose(m.transitions5(), np.where(P > 0, np.log(np.where(P > 0, P, 1.0)), -np.inf))
    assert np.isneginf(m.transitions5()[1, 3])
    t = m.tables(0)["trans"]  # short-gap slots, reverse strand too
    with np.errstate(divide="ignore"):
        L = np.log(P)
    assert np.allclose(t, [L[0, 0], L[1, 0], L[2, 0], L[0, 1], L[0, 2], L[1, 1], L[2, 2], L[2, 1], L[1, 2]])
    with pytest.raises(MPError):
        m.set_transitions([-1.0] * 9)
    three = ph.StateMachines.from_json(PARAMS)
    assert three.num_states == 3
    with pytest.raises(MPError):
        three.transitions5()
    with pytest.raises(MPError):
        three.set_transitions5(np.zeros((5, 5)) + 0.5)  # not log probabilities
    three.add_long_gaps()
    T = three.transitions5()
    assert three.num_states == 5 and np.allclose(np.exp(T).sum(1), 1.0)
    bad = P.copy()
    bad[0, 0] = -0.1
    with pytest.raises(MPError):
        ph.StateMachines.from_json(_five_state_json(tmp_path, bad))
