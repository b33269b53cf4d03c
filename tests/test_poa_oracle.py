"""Pins of the POA-augmentation oracle (oracle/poa.py) against the reference's
own known answers (tests/polisherTest.c) and invariants the definitions fix."""
import random

import pytest

from oracle import poa as P


def _node_inserts(node):
    return [(e[0][0], e[1] + e[2]) for e in node.inserts]


def _node_deletes(node):
    return [(e[0], e[1] + e[2]) for e in node.deletes]


def test_poa_augment_example():
    """test_poa_augment_example (polisherTest.c:160-241): GATTACA vs GATACGGT."""
    ref = P.rle_no_rle("GATTACA")
    read = P.rle_no_rle("GATACGGT")
    poa = P.Poa(ref, 51)
    matches = [(100, 0, 0), (100, 1, 1), (50, 2, 2), (50, 3, 2), (100, 4, 3), (100, 5, 4), (50, 6, 5),
               (25, 6, 6), (25, 6, 7)]
    inserts = [(50, 5, 5), (25, 5, 6), (50, 6, 6), (75, 6, 7)]
    deletes = [(50, 2, 1), (50, 3, 2)]
    P.augment(poa, read, True, 0, matches, inserts, deletes, True)
    assert len(poa.nodes) == 8
    want = [
        ("N", [0, 0, 0, 0, 0], [], []),
        ("G", [0, 0, 100, 0, 0], [], []),
        ("A", [100, 0, 0, 0, 0], [], [(1, 100.0)]),
        ("T", [0, 0, 0, 50, 0], [], []),
        ("T", [0, 0, 0, 50, 0], [], []),
        ("A", [100, 0, 0, 0, 0], [], []),
        ("C", [0, 100, 0, 0, 0], [("G", 50.0), ("GG", 25.0)], []),
        ("A", [0, 0, 75, 25, 0], [("GT", 50.0), ("T", 75.0)], []),
    ]
    for node, (base, bw, ins, dels) in zip(poa.nodes, want):
        assert node.base == base
        assert node.base_weights == [float(v) for v in bw]
        assert _node_inserts(node) == ins
        assert _node_deletes(node) == dels


def test_rotate_string():
    """test_rle_rotateString (polisherTest.c:869-898)."""
    assert P.rotate(P.rle("GATACA"), 2)[0] == "CAGATA"
    rng = random.Random(7)
    for _ in range(500):
        s = "".join(rng.choice("ACGTN") for _ in range(rng.randint(0, 20)))
        t = P.rle(s)
        k = rng.randint(0, 20)
        r = P.rotate(t, k)
        for j in range(len(t[0])):
            assert r[0][(j + k) % len(t[0])] == t[0][j]
            assert r[1][(j + k) % len(t[0])] == t[1][j]


def test_get_shift():
    """test_getShift (polisherTest.c:59-112): the insert placed at the shifted
    position spells the string with the insert appended."""
    rng = random.Random(11)
    for _ in range(10000):
        s = "".join(rng.choice("ACGT") for _ in range(rng.randint(1, 20)))
        ins = "".join(rng.choice("ACGT") for _ in range(rng.randint(1, 4)))
        sr, ir = P.rle(s), P.rle(ins)
        i = P.get_shift(sr, len(sr[0]), ir, True)
        k = sum(sr[1][:i])
        assert s[:k] + ins + s[k:] == s + ins


def test_get_shift_examples():
    """Hand-checked cases: minimal internal repeat, counts compared or not, the
    single-symbol extra step."""
    ref = P.rle("GATATAT")
    assert P.get_shift(ref, 7, P.rle("AT"), True) == 1
    assert P.get_shift(ref, 7, P.rle("ATAT"), True) == 1
    ref = P.rle("GAAC")  # G A2 C
    assert P.get_shift(ref, 2, P.rle("A"), True) == 1   # count differs: one extra step only
    assert P.get_shift(ref, 2, P.rle("A"), False) == 1
    assert P.get_shift(ref, 2, P.rle("AA"), True) == 1  # A2 equals the run
    ref = P.rle("CCGCCG")  # C2 G C2 G
    assert P.get_shift(ref, 4, P.rle("CCG"), True) == 0
    assert P.get_shift(ref, 4, P.rle("CG"), True) == 4   # C2G is not a copy of C1G with counts compared
    assert P.get_shift(ref, 4, P.rle("CG"), False) == 0
    assert P.common_suffix(P.rle("GATTACA"), 6, P.rle("TACA"), True) == 3  # G A T2 A C A: T1 stops it
    assert P.common_suffix(P.rle("GATTACA"), 6, P.rle("TTACA"), True) == 4
    assert P.common_suffix(P.rle("GATTACA"), 6, P.rle("TACA"), False) == 4


def _random_case(rng, ref_len=12, read_len=12, cmp=True):
    ref = P.rle("".join(rng.choice("ACGT") for _ in range(ref_len)))
    read = P.rle("".join(rng.choice("ACGT") for _ in range(read_len)))
    n, m = len(ref[0]), len(read[0])
    cells = [(x, y) for x in range(n) for y in range(m)]
    matches = [(rng.randint(1, 10 ** 7), x, y) for (x, y) in cells if rng.random() < 0.3]
    inserts = [(rng.randint(1, 10 ** 7), x, y) for x in range(-1, n) for y in range(m) if rng.random() < 0.25]
    deletes = [(rng.randint(1, 10 ** 7), x, y) for x in range(n) for y in range(-1, m) if rng.random() < 0.25]
    return ref, read, matches, inserts, deletes


def _spell(ref, pos, s=("", ()), drop=0, counts=True):
    """The sequence after placing s at pos (or dropping `drop` symbols there):
    expanded, or as a base string only when counts are not compared."""
    bases = ref[0][:pos] + s[0] + ref[0][pos + drop:]
    if not counts:
        return bases
    return P.expand((bases, ref[1][:pos] + s[1] + ref[1][pos + drop:]))


def test_shifted_indels_spell_the_same_sequence():
    """Every complete insert / delete, moved to its leftmost equivalent
    position (getShift then the common-suffix rotation), spells the same
    sequence as at the position where it was found (expanded when repeat
    counts are compared, as base strings when they are not), and never moves
    right."""
    rng = random.Random(3)
    seen_ins = seen_del = moved = 0
    for case in range(300):
        cmp = case % 2 == 0
        ref, read, matches, inserts, deletes = _random_case(rng)
        poa = P.Poa(ref, 51)
        mset = set((x, y) for _, x, y in matches)
        for found, pos, placed, s, w, off in P.insert_events(poa, read, mset, inserts, cmp):
            seen_ins += 1
            moved += pos != found
            assert 0 <= pos <= found
            assert _spell(ref, pos, placed, counts=cmp) == _spell(ref, found, s, counts=cmp)
            if cmp:  # counts compared: equal as run-length strings, not only expanded
                assert sorted(placed[1]) == sorted(s[1])
        for found, pos, length, w, off in P.delete_events(poa, mset, deletes, cmp):
            seen_del += 1
            moved += pos != found
            assert 0 <= pos <= found
            assert _spell(ref, pos, drop=length, counts=cmp) == _spell(ref, found, drop=length, counts=cmp)
    assert seen_ins > 100 and seen_del > 100 and moved > 50


def test_weights_are_run_minima_and_observations_follow():
    """Insert/delete weights are the minimum tuple weight over the complete
    run; each placement adds one observation carrying that weight."""
    rng = random.Random(5)
    for _ in range(100):
        ref, read, matches, inserts, deletes = _random_case(rng)
        poa = P.Poa(ref, 51)
        P.augment(poa, read, rng.random() < 0.5, 3, matches, inserts, deletes, True)
        for node in poa.nodes:
            for s, wf, wr, obs in node.inserts:
                assert wf + wr == pytest.approx(sum(o[2] for o in obs), rel=0, abs=0)
                assert all(o[0] == 3 for o in obs)
            for ln, wf, wr, obs in node.deletes:
                assert wf + wr == sum(o[2] for o in obs)
        tot = sum(sum(n.base_weights) for n in poa.nodes)
        assert tot == sum(float(w) for w, _, _ in matches)
        assert tot == sum(sum(n.rc_weights) for n in poa.nodes)


def _peaked_table():
    """log P(observed | underlying) = -|observed - underlying| for every base:
    the maximum-likelihood run length is then a weighted median."""
    import numpy as np
    u = np.arange(51)
    t = -np.abs(u[None, :] - u[:, None]).astype(float)  # [underlying][observed]
    return np.stack([t] * 5)


def test_repeat_count_estimate_known_answers():
    """repeatSubMatrix_getMLRepeatCount (polisher.c:1846-1865) on a table whose
    answer is a weighted median; no observation -> 1; an overlong observation
    is capped at 50; observations that are all overlong leave the smallest
    observed count at its initial bound, which the reference treats as "no
    valid observation" (-> 1)."""
    ref = ("ACGT", (3, 1, 2, 1))
    reads = [("A", (2,)), ("A", (3,)), ("A", (3,)), ("A", (5,)), ("C", (80,)), ("G", (60,)), ("G", (70,)),
             ("C", (3,))]
    poa = P.Poa(ref, 51)
    for rn in range(4):
        P.augment(poa, reads[rn], True, rn, [(10, 0, 0)], [], [])
    P.augment(poa, reads[4], False, 4, [(5, 1, 0)], [], [])      # node 2: overlong (-> 50) with weight 5
    P.augment(poa, reads[7], True, 7, [(1, 1, 0)], [], [])       # and 3 with weight 1
    P.augment(poa, reads[5], True, 5, [(7, 2, 0)], [], [])       # node 3: all overlong
    P.augment(poa, reads[6], True, 6, [(7, 2, 0)], [], [])
    strands = [True, True, True, True, False, True, True, True]
    got = P.estimate_repeat_counts(poa, reads, strands, _peaked_table())
    assert got == [3, 50, 1, 1]
    # weights move the median
    poa = P.Poa(ref, 51)
    for rn, w in enumerate([1, 1, 1, 100]):
        P.augment(poa, reads[rn], True, rn, [(w, 0, 0)], [], [])
    assert P.estimate_repeat_counts(poa, reads, strands, _peaked_table())[0] == 5


def test_anchor_alignments_known_answer():
    """poa_getAnchorAlignments (polisher.c:536-592) on a hand-built graph:
    thresholds 0.9 -> 10, 0.95 -> 4, 0.99 -> 0 (allParams.np.json); an
    observation at or below 0.9 is skipped, a non-increasing one is dropped,
    unaligned nodes (-1) are skipped."""
    poa = P.Poa(P.rle_no_rle("ACGTA"), 51)
    read = P.rle_no_rle("ACGTA")
    m = [(9_500_000, 0, 0), (9_000_000, 1, 1), (9_990_000, 2, 2), (9_999_999, 2, 1), (9_200_000, 3, 3),
         (9_600_000, 4, 4)]
    P.augment(poa, read, True, 0, m, [], [])
    params = [0.9, 10, 0.95, 4, 0.99, 0]
    assert P.anchor_alignments(poa, 1, params) == [[(0, 0, 4), (2, 2, 0), (3, 3, 10), (4, 4, 4)]]
    assert P.anchor_alignments(poa, 1, params, [0, 1, -1, 5, 6]) == [[(0, 0, 4), (5, 3, 10), (6, 4, 4)]]


def test_log_add_matches_log_sum_exp():
    """logAdd (pairwiseAligner.c:282-299): within the cubic's error of the
    exact log-sum-exp below the 7.5 cut, the larger argument above it."""
    import math
    for x in [0.0, -0.3, -1.0, -2.2, -4.0, -7.4]:
        for y in [0.0, -0.5, -3.0]:
            exact = max(x, y) + math.log1p(math.exp(-abs(x - y)))
            got = P.log_add(x, y)
            if abs(x - y) >= 7.5:
                assert got == max(x, y)
            else:
                assert abs(got - exact) < 2e-3, (x, y)
    assert P.log_add(float("-inf"), -2.0) == -2.0 and P.log_add(-2.0, -12.0) == -2.0


def _graph(ref, events):
    """A graph with hand-set weights: events = [(kind, node, payload, weight)]."""
    poa = P.Poa(P.rle_no_rle(ref), 51)
    for node in poa.nodes[1:]:
        node.base_weights[P.char_to_symbol(node.base)] = 100.0
        node.rc_weights[1] = 100.0
    for kind, i, payload, w in events:
        if kind == "sub":
            poa.nodes[i].base_weights = [0.0] * 5
            poa.nodes[i].base_weights[P.char_to_symbol(payload)] = w
        elif kind == "ins":
            P._add_insert(poa.nodes[i], P.rle_no_rle(payload), w, True, (0, 0, w))
        else:
            P._add_delete(poa.nodes[i], payload, w, True, (0, 0, w))
    return poa


def test_consensus_known_answers():
    """poa_getConsensus (polisher.c:1040-1275) on hand-weighted graphs: the
    reference when it is supported, a heavier substitution, an insert or a
    delete when the reads carry it, nothing when they do not."""
    assert P.consensus(_graph("GATTACA", []), rle=False)[0] == "GATTACA"
    assert P.consensus(_graph("GATTACA", []), rle=False)[1] == list(range(7))
    assert P.consensus(_graph("GATTACA", [("sub", 4, "G", 100.0)]), rle=False)[0] == "GATGACA"
    # the reference base keeps a position unless the alternative beats it after the 0.5 discount
    g = _graph("GATTACA", [])
    g.nodes[4].base_weights = [0.0, 0.0, 60.0, 100.0, 0.0]
    assert P.consensus(g, rle=False)[0] == "GATGACA"
    g.nodes[4].base_weights = [0.0, 0.0, 40.0, 100.0, 0.0]
    assert P.consensus(g, rle=False)[0] == "GATTACA"
    assert P.consensus(_graph("GATTACA", [("ins", 3, "GG", 90.0)]), rle=False)[0] == "GATGGTACA"
    assert P.consensus(_graph("GATTACA", [("ins", 3, "GG", 5.0)]), rle=False)[0] == "GATTACA"
    s, m = P.consensus(_graph("GATTACA", [("del", 2, 2, 90.0)]), rle=False)
    assert s == "GAACA" and m == [0, 1, -1, -1, 2, 3, 4]


def test_insert_strings_merge_whatever_made_them():
    """addToInserts (polisher.c:177-205) merges equal run-length strings: a
    substring of a read and the rotation that makes the same string are one
    entry (the oracle once kept them apart because one carried its run lengths
    as a list and the other as a tuple)."""
    read = ("GAT", [1, 3, 2])
    s = P.substring(read, 0, 2)
    r = P.rotate(("AG", (3, 1)), 1)
    assert s == r == ("GA", (1, 3))
    node = P.Node("A", 1, 51)
    P._add_insert(node, s, 5.0, True, (0, 0, 5.0))
    P._add_insert(node, r, 7.0, True, (1, 0, 7.0))
    assert len(node.inserts) == 1 and node.inserts[0][1] == 12.0
