"""The library's POA augmentation (host code, include/mp_poa.h) against the
oracle restatement of poa_augment, bit for bit: node weights, insert and
delete lists in order, every observation.  Runs on the CPU (no kernel)."""
import random

import numpy as np
import pytest

from marginpolish_amd import _lib
from marginpolish_amd import crop
from marginpolish_amd import poa as mpoa
from marginpolish_amd import sim
from oracle import oracle as O
from oracle import poa as P
from tests.conftest import PARAMS
from tests.poa_compare import assert_same_graph, native_as_oracle, oracle_lists


def _both(ref, max_rc=51):
    return mpoa.Poa(ref, max_rc), P.Poa(ref, max_rc)


def test_poa_augment_example_native():
    """test_poa_augment_example (polisherTest.c:160-241) through the library."""
    ref, read = P.rle_no_rle("GATTACA"), P.rle_no_rle("GATACGGT")
    matches = [(100, 0, 0), (100, 1, 1), (50, 2, 2), (50, 3, 2), (100, 4, 3), (100, 5, 4), (50, 6, 5),
               (25, 6, 6), (25, 6, 7)]
    inserts = [(50, 5, 5), (25, 5, 6), (50, 6, 6), (75, 6, 7)]
    deletes = [(50, 2, 1), (50, 3, 2)]
    g, o = _both(ref)
    g.augment(read, 1, 0, matches, inserts, deletes, True)
    P.augment(o, read, True, 0, matches, inserts, deletes, True)
    nodes = native_as_oracle(g)
    assert_same_graph(nodes, o)
    assert [e[0][0] for e in nodes[6].inserts] == ["G", "GG"]
    assert [e[1] for e in nodes[6].inserts] == [50.0, 25.0]
    assert [e[0][0] for e in nodes[7].inserts] == ["GT", "T"]
    assert [(e[0], e[1]) for e in nodes[2].deletes] == [(1, 100.0)]


def _rand_rle(rng, n, alphabet="ACGT", max_run=4):
    b = "".join(rng.choice(alphabet) for _ in range(n))
    runs = tuple(rng.choice([1, 1, 1, 2, 2, 3, max_run]) for _ in range(n))
    # adjacent equal bases are allowed in an RLE string built elsewhere; keep them distinct as rleString_construct does
    out_b, out_c = [], []
    for c, r in zip(b, runs):
        if out_b and out_b[-1] == c:
            continue
        out_b.append(c)
        out_c.append(r)
    return "".join(out_b), tuple(out_c)


def _rand_lists(rng, n, m, dens=(0.3, 0.25, 0.25)):
    matches = [(rng.randint(1, 10 ** 7), x, y) for x in range(n) for y in range(m) if rng.random() < dens[0]]
    inserts = [(rng.randint(1, 10 ** 7), x, y) for x in range(-1, n) for y in range(m) if rng.random() < dens[1]]
    deletes = [(rng.randint(1, 10 ** 7), x, y) for x in range(n) for y in range(-1, m) if rng.random() < dens[2]]
    for lst in (matches, inserts, deletes):
        rng.shuffle(lst)  # the reference sorts the gap lists itself; match order sets observation order
    return matches, inserts, deletes


@pytest.mark.parametrize("cmp", [True, False])
def test_random_reads_match_oracle(cmp):
    """Several reads per graph, random dense lists (long complete runs,
    repeats, both strands), counts compared or not."""
    rng = random.Random(17 + cmp)
    for case in range(40):
        ref = _rand_rle(rng, rng.randint(1, 25), "ACGT" if case % 5 else "AC")
        g, o = _both(ref, rng.choice([2, 5, 51]))
        for r in range(rng.randint(1, 4)):
            read = _rand_rle(rng, rng.randint(1, 25), "ACGT" if case % 5 else "AC", max_run=60)
            m, i, d = _rand_lists(rng, len(ref[0]), len(read[0]))
            strand = rng.random() < 0.5
            g.augment(read, strand, r, m, i, d, cmp)
            P.augment(o, read, strand, r, m, i, d, cmp)
        assert_same_graph(native_as_oracle(g), o)


def test_edge_cases():
    rng = random.Random(4)
    # empty lists, empty read
    g, o = _both(P.rle("GATTACA"))
    g.augment(("", ()), 1, 0, [], [], [], True)
    P.augment(o, ("", ()), True, 0, [], [], [], True)
    assert_same_graph(native_as_oracle(g), o)
    # empty reference: one prefix node; only inserts at x = -1 are legal
    g, o = _both(("", ()))
    read = P.rle("ACGT")
    ins = [(7, -1, 0), (8, -1, 1), (9, -1, 2), (6, -1, 3)]
    g.augment(read, 0, 5, [], ins, [], True)
    P.augment(o, read, False, 5, [], ins, [], True)
    nodes = native_as_oracle(g)
    assert_same_graph(nodes, o)
    assert nodes[0].inserts[0][0][0] == "ACGT" and nodes[0].inserts[0][2] == 6.0
    # lower-case and non-ACGT reference bases, read counts beyond the repeat-count bound
    ref = ("gaTNxC", (1, 2, 3, 1, 1, 70))
    read = ("GATNYC", (1, 2, 300, 1, 1, 55))
    g, o = _both(ref, 51)
    m = [(10, k, k) for k in range(6)]
    g.augment(read, 1, 0, m, [], [], True)
    P.augment(o, read, True, 0, m, [], [], True)
    nodes = native_as_oracle(g)
    assert_same_graph(nodes, o)
    assert "".join(n.base for n in nodes) == "NGATNNC"
    assert nodes[3].rc_weights[50] == 10.0 and nodes[5].base_weights[4] == 10.0
    # a homopolymer / tandem repeat reference where indels shift far left
    ref = P.rle("G" + "AT" * 30 + "C")
    read = P.rle("G" + "AT" * 31 + "C")
    n, mlen = len(ref[0]), len(read[0])
    m = [(9, x, x) for x in range(n - 1)] + [(9, n - 1, mlen - 1)]
    ins = [(5, n - 2, n - 1), (5, n - 2, n)]
    g, o = _both(ref)
    g.augment(read, 1, 0, m, ins, [(3, 10, 9), (3, 11, 9)], True)
    P.augment(o, read, True, 0, m, ins, [(3, 10, 9), (3, 11, 9)], True)
    nodes = native_as_oracle(g)
    assert_same_graph(nodes, o)
    assert nodes[1].inserts and nodes[1].inserts[0][0][0] in ("AT", "TA")
    del rng


def test_out_of_range_coordinates_are_rejected():
    g = mpoa.Poa(P.rle("ACGT"))
    read = P.rle("ACG")
    for m, i, d in (([(1, 4, 0)], [], []), ([(1, 0, 3)], [], []), ([], [(1, -2, 0)], []), ([], [(1, 0, 3)], []),
                    ([], [], [(1, -1, 0)]), ([], [], [(1, 0, -2)])):
        with pytest.raises(_lib.MPError) as e:
            g.augment(read, 1, 0, m, i, d, True)
        assert e.value.status == 1
    assert g.sizes()["nodeObservations"] == 0


def test_realistic_lists_match_oracle():
    """Lists from the exact FP64 pair-HMM oracle on simulated RLE reads
    (cropped reference, uncropped coordinates), several reads into one
    graph, forward and reverse strands."""
    pp = O.PolishParams(PARAMS)
    w = sim.config0(n_reads=4, ref_len=700, seed=9, strands=(1, 0))
    ref_sym = w.pairs[0].ref_sym
    ref = mpoa.rle_from_symbols(ref_sym)
    ref_o = (ref[0].tobytes().decode(), tuple(int(c) for c in ref[1]))
    g, o = mpoa.Poa(ref), P.Poa(ref_o, 51)
    total = 0
    for r, pr in enumerate(w.pairs):
        sx, sy, an, first = crop.crop_pair(pr.ref_sym, pr.read_sym, pr.anchors)
        model = pp.forward if pr.strand else pp.reverse
        m, gx, gy, _, _ = O.posteriors(model, sx, sy, an, pp.p)
        lists = []
        for t in (m, gy, gx):
            t = np.asarray(t, dtype=np.int64).reshape(-1, 3).copy()
            t[:, 1] += first
            lists.append([tuple(int(v) for v in row) for row in t])
        read = mpoa.rle_from_symbols(pr.read_sym)
        read_o = (read[0].tobytes().decode(), tuple(int(c) for c in read[1]))
        g.augment(read, pr.strand, r, *lists, True)
        P.augment(o, read_o, bool(pr.strand), r, *lists, True)
        total += sum(len(x) for x in lists)
    nodes = native_as_oracle(g)
    assert_same_graph(nodes, o)
    assert total > 3000
    assert sum(len(n.inserts) for n in nodes) > 10 and sum(len(n.deletes) for n in nodes) > 10


@pytest.mark.parametrize("table", ["params", "random"])
def test_repeat_count_estimates_match_oracle(table):
    """poa_estimateRepeatCountsUsingBayesianModel: library vs oracle, bit for
    bit, on graphs built from FP64-oracle posterior lists of simulated reads
    (uncapped run lengths, both strands)."""
    from marginpolish_amd import pairhmm as ph
    pp = O.PolishParams(PARAMS)
    mach = ph.PolishParams(PARAMS).machines
    rep = np.asarray(pp.repeat, dtype=np.float64).reshape(5, 51, 51).copy()
    if table == "random":
        rep[:4] = -np.abs(np.random.default_rng(3).normal(2.0, 1.0, size=(4, 51, 51)))
        mach.set_repeat_log10(rep[:4])
    w = sim.config0(n_reads=5, ref_len=800, seed=13, strands=(1, 0))
    ref = mpoa.rle_from_symbols(w.pairs[0].ref_sym, w.pairs[0].ref_runs)
    ref_o = (ref[0].tobytes().decode(), tuple(int(c) for c in ref[1]))
    g, o = mpoa.Poa(ref), P.Poa(ref_o, 51)
    reads, reads_o, strands = [], [], []
    for r, pr in enumerate(w.pairs):
        sx, sy, an, first = crop.crop_pair(pr.ref_sym, pr.read_sym, pr.anchors)
        m, gx, gy, _, _ = O.posteriors(pp.forward if pr.strand else pp.reverse, sx, sy, an, pp.p)
        lists = []
        for t in (m, gy, gx):
            t = np.asarray(t, dtype=np.int64).reshape(-1, 3).copy()
            t[:, 1] += first
            lists.append([tuple(int(v) for v in row) for row in t])
        runs = pr.read_runs.copy()
        runs[::37] += 60  # some overlong observations
        read = mpoa.rle_from_symbols(pr.read_sym, runs)
        reads.append(read)
        reads_o.append((read[0].tobytes().decode(), tuple(int(c) for c in read[1])))
        strands.append(pr.strand)
        g.augment(read, pr.strand, r, *lists, True)
        P.augment(o, reads_o[-1], bool(pr.strand), r, *lists, True)
    got = g.estimate_repeat_counts(mach, reads, strands)
    want = P.estimate_repeat_counts(o, reads_o, [bool(s) for s in strands], rep)
    assert got.tolist() == want
    # several graphs in one call (the same graph twice, and an empty one)
    empty = mpoa.Poa(ref)
    many = mpoa.estimate_repeat_counts_many([g, empty, g], mach, [reads, [], reads], [strands, [], strands])
    assert many[0].tolist() == want and many[2].tolist() == want and set(many[1].tolist()) == {1}
    assert len(set(want)) > 3


def test_anchor_alignments_match_oracle():
    """poa_getAnchorAlignments: library vs oracle on a graph of several
    reads built from FP64-oracle lists, identity map and a consensus map with
    unaligned nodes and shifts."""
    pp = O.PolishParams(PARAMS)
    w = sim.config0(n_reads=6, ref_len=900, seed=17, strands=(1, 0))
    ref = mpoa.rle_from_symbols(w.pairs[0].ref_sym, w.pairs[0].ref_runs)
    ref_o = (ref[0].tobytes().decode(), tuple(int(c) for c in ref[1]))
    g, o = mpoa.Poa(ref), P.Poa(ref_o, 51)
    for r, pr in enumerate(w.pairs):
        sx, sy, an, first = crop.crop_pair(pr.ref_sym, pr.read_sym, pr.anchors)
        m, gx, gy, _, _ = O.posteriors(pp.forward if pr.strand else pp.reverse, sx, sy, an, pp.p)
        lists = []
        for t in (m, gy, gx):
            t = np.asarray(t, dtype=np.int64).reshape(-1, 3).copy()
            t[:, 1] += first
            lists.append([tuple(int(v) for v in row) for row in t])
        read = mpoa.rle_from_symbols(pr.read_sym, pr.read_runs)
        read_o = (read[0].tobytes().decode(), tuple(int(c) for c in read[1]))
        g.augment(read, pr.strand, r, *lists, True)
        P.augment(o, read_o, bool(pr.strand), r, *lists, True)
    params = [0.9, 10, 0.95, 4, 0.99, 0]
    n = len(ref[0])
    rng = np.random.default_rng(1)
    cmap = np.cumsum(rng.integers(0, 3, size=n)) - 1
    cmap[rng.random(n) < 0.1] = -1
    for mp_ in (None, cmap):
        got = g.anchor_alignments(len(w.pairs), params, mp_)
        want = P.anchor_alignments(o, len(w.pairs), params, None if mp_ is None else [int(v) for v in mp_])
        assert [a.tolist() for a in got] == [[list(t) for t in lst] for lst in want]
        assert sum(len(a) for a in got) > 1000


def _edit_distance(a, b):
    prev = list(range(len(b) + 1))
    for i, ca in enumerate(a, 1):
        cur = [i] + [0] * len(b)
        for j, cb in enumerate(b, 1):
            cur[j] = min(prev[j] + 1, cur[j - 1] + 1, prev[j - 1] + (ca != cb))
        prev = cur
    return prev[-1]


@pytest.mark.parametrize("rle,penalty", [(True, 0.5), (False, 0.5), (True, 1.0)])
def test_consensus_matches_oracle_and_polishes(rle, penalty):
    """poa_getConsensus: library vs oracle bit for bit (string and map) on a
    graph of reads simulated from a truth sequence and aligned (FP64 oracle)
    to a draft with errors; the consensus is closer to the truth than the
    draft.  Without run-length encoding the graph is built from plain
    strings and the no-RLE params, as the reference requires."""
    from tests.conftest import PARAMS_NO_RLE
    pp = O.PolishParams(PARAMS if rle else PARAMS_NO_RLE)
    rng = np.random.default_rng(23)
    truth = sim._ref_runs(rng, 600, 0.45, 0.45)
    draft, dpos = sim._mutate(rng, truth, 0.04)
    reads = [sim._mutate(rng, truth, 0.06) for _ in range(16)]
    chars = np.frombuffer(b"ACGT", np.uint8)

    def plain(bases):  # symbols without run lengths (useRepeatCountsInAlignment off)
        return bases.astype(np.uint16), np.ones(bases.size, np.uint64)

    if rle:
        rb, rr, _ = sim._rle(draft)
        ref_rle = mpoa.rle_from_symbols(sim._symbols(rb, rr), rr)
    else:
        ref_rle = (chars[draft], np.ones(draft.size, np.uint64))
    ref_o = (bytes(ref_rle[0]).decode(), tuple(int(c) for c in ref_rle[1]))
    g, o = mpoa.Poa(ref_rle, 51 if rle else 2), P.Poa(ref_o, 51 if rle else 2)
    for r, (rd, pos) in enumerate(reads):
        t_idx = np.flatnonzero(pos >= 0)
        keep = dpos[t_idx] >= 0
        if rle:
            pr = sim._pair(draft, rd, dpos[t_idx][keep], pos[t_idx][keep], 10, r % 2, 0, r)
            ref_sym, read_sym, anchors, read = pr.ref_sym, pr.read_sym, pr.anchors, \
                mpoa.rle_from_symbols(pr.read_sym, pr.read_runs)
        else:
            X, Y = dpos[t_idx][keep], pos[t_idx][keep]
            ok = np.r_[True, (X[1:] > X[:-1]) & (Y[1:] > Y[:-1])]
            anchors = np.stack([X[ok], Y[ok], np.full(int(ok.sum()), 10)], axis=1).astype(np.int64)
            ref_sym, read_sym = plain(draft)[0], plain(rd)[0]
            read = (chars[rd], np.ones(rd.size, np.uint64))
        sx, sy, an, first = crop.crop_pair(ref_sym, read_sym, anchors)
        m, gx, gy, _, _ = O.posteriors(pp.forward if r % 2 else pp.reverse, sx, sy, an, pp.p)
        lists = []
        for t in (m, gy, gx):
            t = np.asarray(t, dtype=np.int64).reshape(-1, 3).copy()
            t[:, 1] += first
            lists.append([tuple(int(v) for v in row) for row in t])
        g.augment(read, r % 2, r, *lists, True)
        P.augment(o, (bytes(read[0]).decode(), tuple(int(c) for c in read[1])), bool(r % 2), r, *lists, True)
    s, cmap = g.consensus(rle, penalty)
    s_o, cmap_o = P.consensus(o, rle, penalty)
    assert s == s_o
    assert cmap.tolist() == cmap_o
    if penalty == 0.5:
        t_str = bytes(chars[truth]).decode()
        d_str = bytes(chars[draft]).decode()
        assert _edit_distance(s, t_str) < _edit_distance(d_str, t_str) / 2
