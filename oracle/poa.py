"""Oracle for POA graph augmentation from posterior lists -- test infrastructure.

A pure-Python restatement of the reference's reference-graph construction and
poa_augment (impl/polisher.c), used only by tests/ and bench.py's CPU leg to
check the library's native implementation.  Nothing in the product imports it.

RLE strings are (bases, counts): a str of base characters and a tuple of run
lengths (RleString, inc/margin.h).  Aligned pairs are (weight, x, y) tuples
with x a reference and y a read coordinate, as the posterior lists hold them.

Pins (tests/test_poa_oracle.py): the reference's known-answer tests
test_poa_augment_example, test_rle_rotateString and the property of
test_getShift (polisherTest.c:59-240, 869-898), plus brute-force invariants
(a shifted insert or delete spells the same sequence as the unshifted one,
complete-insert/delete enumeration against a direct definition).
"""
from __future__ import annotations

import numpy as np

NUC_SYMBOL = {"A": 0, "a": 0, "C": 1, "c": 1, "G": 2, "g": 2, "T": 3, "t": 3}
ALPHABET_SIZE = 5  # A C G T N (stateMachine.c:63-71)


def char_to_symbol(c):
    """convertNucleotideCharToSymbol (stateMachine.c:25-42)."""
    return NUC_SYMBOL.get(c, 4)


def rle(s):
    """rleString_construct (polisher.c:1357-1390)."""
    bases, counts = [], []
    for c in s:
        if bases and bases[-1] == c:
            counts[-1] += 1
        else:
            bases.append(c)
            counts.append(1)
    return "".join(bases), tuple(counts)


def rle_no_rle(s):
    """rleString_construct_no_rle: every character its own run of 1."""
    return s, tuple([1] * len(s))


def expand(r):
    return "".join(b * c for b, c in zip(*r))


def rotate(r, k):
    """rleString_rotateString (polisher.c:1508-1520): character i moves to (i + k) mod length."""
    b, c = r
    n = len(b)
    if n == 0:
        return r
    ob, oc = [None] * n, [None] * n
    for i in range(n):
        ob[(i + k) % n] = b[i]
        oc[(i + k) % n] = c[i]
    return "".join(ob), tuple(oc)


def substring(r, start, length):
    """rleString_copySubstring (polisher.c:1433-1453).  The run lengths come
    back as a tuple, the form rotate() returns, so that equal strings compare
    equal in addToInserts whichever of the two made them."""
    return r[0][start:start + length], tuple(r[1][start:start + length])


def _matches_ref(ref, start, s, length, cmp_counts):
    """matchesReferenceSubstring (polisher.c:237-249)."""
    for l in range(length):
        if ref[0][start + l] != s[0][l] or (cmp_counts and ref[1][start + l] != s[1][l]):
            return False
    return True


def _has_internal_repeat(s, k, cmp_counts):
    """hasInternalRepeat (polisher.c:251-267)."""
    n = len(s[0])
    if n % k:
        return False
    for i in range(k, n, k):
        for j in range(k):
            if s[0][j] != s[0][j + i] or (cmp_counts and s[1][j] != s[1][j + i]):
                return False
    return True


def get_shift(ref, ref_start, s, cmp_counts):
    """getShift (polisher.c:269-297): walk an indel left over whole copies of
    its minimal internal repeat, then one more symbol for a single-symbol
    string when repeat counts are compared."""
    n = len(s[0])
    m = 1
    while m < n and not _has_internal_repeat(s, m, cmp_counts):
        m += 1
    k = ref_start - m
    while k >= 0:
        if not _matches_ref(ref, k, s, m, cmp_counts):
            break
        ref_start = k
        k -= m
    if n == 1 and cmp_counts and ref_start > 0 and ref[0][ref_start - 1] == s[0][0]:
        ref_start -= 1
    return ref_start


def common_suffix(ref, length1, s, cmp_counts):
    """getMaxCommonSuffixLength (polisher.c:299-314): longest suffix of
    ref[:length1] equal to a suffix of s."""
    i = 0
    n = len(s[0])
    while length1 - i - 1 >= 0 and n - i - 1 >= 0:
        if ref[0][length1 - 1 - i] != s[0][n - 1 - i] or (cmp_counts and ref[1][length1 - 1 - i] != s[1][n - 1 - i]):
            break
        i += 1
    return i


class Node:
    __slots__ = ("base", "repeat_count", "base_weights", "rc_weights", "inserts", "deletes", "observations")

    def __init__(self, base, rc, max_rc):
        self.base = base
        self.repeat_count = rc
        self.base_weights = [0.0] * ALPHABET_SIZE
        self.rc_weights = [0.0] * max_rc
        self.inserts = []   # [rle string, forward weight, reverse weight, observations]
        self.deletes = []   # [length, forward weight, reverse weight, observations]
        self.observations = []  # (readNo, offset, weight)


class Poa:
    """poa_getReferenceGraph (polisher.c:118-131): a prefix node 'N'/1 then one
    node per reference symbol, base upper-cased, non-ACGT -> 'N'
    (poaNode_construct, polisher.c:90-107)."""

    def __init__(self, ref, max_rc):
        self.ref = ref
        self.max_rc = max_rc
        self.nodes = [Node("N", 1, max_rc)]
        for b, c in zip(*ref):
            u = b.upper()
            self.nodes.append(Node(u if char_to_symbol(u) < 4 else "N", c, max_rc))


def _add_insert(node, s, weight, strand, obs):
    """addToInserts (polisher.c:177-205): first equal RLE string or a new entry at the end."""
    for e in node.inserts:
        if e[0] == s:
            break
    else:
        e = [s, 0.0, 0.0, []]
        node.inserts.append(e)
    e[1 if strand else 2] += weight
    e[3].append(obs)


def _add_delete(node, length, weight, strand, obs):
    """addToDeletes (polisher.c:207-235): first entry of the same length or a new one."""
    for e in node.deletes:
        if e[0] == length:
            break
    else:
        e = [length, 0.0, 0.0, []]
        node.deletes.append(e)
    e[1 if strand else 2] += weight
    e[3].append(obs)


UINT_MAX = 4294967295.0


def insert_events(poa, read, mset, inserts, cmp_counts=True):
    """Complete inserts of one read (polisher.c:351-443): runs of inserts with
    equal x and consecutive y; every sub-run flanked by a match (or the read's
    ends) on both sides.  Yields (found_at, placed_at, placed_string,
    found_string, weight, read_offset) in the reference's order."""
    read_len = len(read[0])
    ins = sorted(inserts, key=lambda t: (t[1], t[2]))
    i = 0
    while i < len(ins):
        x0, y0 = ins[i][1], ins[i][2]
        j = i + 1
        while j < len(ins) and ins[j][1] == x0 and ins[j][2] == y0 + j - i:
            j += 1
        for k in range(i, j):
            yp = y0 + k - i - 1
            if (x0, yp) not in mset and yp > -1:
                continue
            for l in range(k, j):
                yn = y0 + l - i + 1
                if (x0 + 1, yn) not in mset and yn < read_len:
                    continue
                s = substring(read, ins[k][2], l + 1 - k)
                w = UINT_MAX
                for m in range(k, l + 1):
                    w = min(w, float(ins[m][0]))
                pos = get_shift(poa.ref, x0 + 1, s, cmp_counts)
                cs = common_suffix(poa.ref, pos, s, cmp_counts)
                placed = rotate(s, cs) if cs > 0 else s
                yield x0 + 1, pos - cs, placed, s, w, ins[k][2]
        i = j


def delete_events(poa, mset, deletes, cmp_counts=True):
    """Complete deletes of one read (polisher.c:445-530): runs of deletes with
    equal y and consecutive x, flanked by matches (or the reference's ends).
    Yields (found_at, placed_at, length, weight, read_offset)."""
    ref_len = len(poa.ref[0])
    dels = sorted(deletes, key=lambda t: (t[2], t[1]))
    i = 0
    while i < len(dels):
        x0, y0 = dels[i][1], dels[i][2]
        j = i + 1
        while j < len(dels) and dels[j][2] == y0 and dels[j][1] == x0 + j - i:
            j += 1
        for k in range(i, j):
            xp = x0 + k - i - 1
            if (xp, y0) not in mset and xp > -1:
                continue
            for l in range(k, j):
                xn = x0 + l - i + 1
                if (xn, y0 + 1) not in mset and xn < ref_len:
                    continue
                length = l - k + 1
                w = UINT_MAX
                for m in range(k, l + 1):
                    w = min(w, float(dels[m][0]))
                found = x0 + k - i
                assert found >= 0
                d = substring(poa.ref, found, length)
                pos = get_shift(poa.ref, found, d, cmp_counts)
                pos -= common_suffix(poa.ref, pos, d, cmp_counts)
                yield found, pos, length, w, y0
        i = j


def augment(poa, read, strand, read_no, matches, inserts, deletes, cmp_counts=True):
    """poa_augment (polisher.c:316-534).  matches / inserts (gap-Y) / deletes
    (gap-X) are lists of (weight, x, y)."""
    nodes = poa.nodes
    # match weights and base observations, in list order (polisher.c:321-339)
    for w, x, y in matches:
        node = nodes[x + 1]
        node.base_weights[char_to_symbol(read[0][y])] += w
        node.rc_weights[min(read[1][y], poa.max_rc - 1)] += w
        node.observations.append((read_no, y, float(w)))
    mset = set((x, y) for _, x, y in matches)
    assert len(mset) == len(matches)
    for _, pos, s, _, w, off in insert_events(poa, read, mset, inserts, cmp_counts):
        _add_insert(nodes[pos], s, w, strand, (read_no, off, w))
    for _, pos, length, w, off in delete_events(poa, mset, deletes, cmp_counts):
        _add_delete(nodes[pos], length, w, strand, (read_no, off, w))


def realign_augment(ref, max_rc, reads, lists, cmp_counts=True):
    """The augment half of poa_realign (polisher.c:656-703) for one chunk:
    reads = [(rle read, strand)], lists = [(matches, inserts, deletes)] in read order."""
    poa = Poa(ref, max_rc)
    for n, ((r, strand), (m, ins, dels)) in enumerate(zip(reads, lists)):
        augment(poa, r, strand, n, m, ins, dels, cmp_counts)
    return poa


MAXIMUM_REPEAT_LENGTH = 51  # margin.h:123


def estimate_repeat_counts(poa, reads, strands, repeat, max_rl=MAXIMUM_REPEAT_LENGTH):
    """poa_estimateRepeatCountsUsingBayesianModel (polisher.c:1557-1566) with
    repeatSubMatrix_getMLRepeatCount / getLogProbForGivenRepeatCount /
    getMinAndMaxRepeatCountObservations (1796-1865) and repeatSubMatrix_setLogProb
    (1756-1770): per node after the prefix, over candidate run lengths from the
    smallest observed to the largest observed (capped at max_rl - 1), the sum
    of weight * logProb[strand row][candidate][min(observed, max_rl - 1)] / 1e7,
    first maximum; no observation -> 0 -> 1.  reads[readNo] = (bases, counts),
    strands[readNo] true for forward; repeat = [5 x max_rl x max_rl] rows by
    base (forward strand), [underlying][observed]."""
    tab = np.asarray(repeat, dtype=np.float64).reshape(5, max_rl, max_rl).tolist()
    out = []
    for node in poa.nodes[1:]:
        b = char_to_symbol(node.base)
        if b == 4:
            b = 0
        obs = node.observations
        lo, hi = max_rl, 0
        for rn, off, _ in obs:
            o = reads[rn][1][off]
            lo, hi = min(lo, o), max(hi, o)
        if hi >= max_rl:
            hi = max_rl - 1
        if lo == max_rl:
            c = 0  # no observation below the bound (lo <= hi otherwise)
        else:
            best, c = None, lo
            for u in range(lo, hi + 1):
                lp = 0.0
                for rn, off, w in obs:
                    o = min(reads[rn][1][off], max_rl - 1)
                    row = b if strands[rn] else 3 - b
                    lp += tab[row][u][o] * w
                lp /= 10000000.0
                if best is None or lp > best:
                    best, c = lp, u
        out.append(1 if c == 0 else c)
    return out


def anchor_alignments(poa, n_reads, params, poa_to_consensus=None):
    """poa_getAnchorAlignments (polisher.c:536-592): nodes in order, each
    node's observations in order; an observation above params[0] (weight /
    1e7) anchors its read at (consensus position, offset, expansion), the
    expansion stepping through the (threshold, expansion) pairs it reaches;
    kept only when strictly after the read's previous anchor."""
    out = [[] for _ in range(n_reads)]
    for i in range(1, len(poa.nodes)):
        ci = i - 1 if poa_to_consensus is None else poa_to_consensus[i - 1]
        if ci == -1:
            continue
        for rn, off, w in poa.nodes[i].observations:
            nw = w / 10000000.0
            if not nw > params[0]:
                continue
            exp = int(params[1])
            for k in range(2, len(params), 2):
                if nw >= params[k]:
                    exp = int(params[k + 1])
                else:
                    break
            lst = out[rn]
            if not lst or (lst[-1][0] < ci and lst[-1][1] < off):
                lst.append((ci, off, exp))
    return out


def _f32(v):
    return float(np.float32(v))


_CUBIC = [(_f32(1.0), [_f32(c) for c in (-0.009350833524763, 0.130659527668286, 0.498799810682272, 0.693203116424741)]),
          (_f32(2.5), [_f32(c) for c in (-0.014532321752540, 0.139942324101744, 0.495635523139337, 0.692140569840976)]),
          (_f32(4.5), [_f32(c) for c in (-0.004605031767994, 0.063427417320019, 0.695956496475118, 0.514272634594009)]),
          (float("inf"), [_f32(c) for c in (-0.000458661602210, 0.009695946122598, 0.930734667215156,
                                            0.168037164329057)])]
LOG_ZERO = float("-inf")


def log_add(x, y):
    """logAdd (pairwiseAligner.c:282-299): the 4-piece cubic of log(1 + e^d)
    with float-rounded coefficients, in double arithmetic, cut at 7.5."""
    if x < y:
        x, y = y, x
        if y == LOG_ZERO or x - y >= 7.5:
            return x
    elif y == LOG_ZERO or x - y >= 7.5:
        return x
    d = x - y
    for lim, (a, b, c, e) in _CUBIC:
        if d <= lim:
            return ((a * d + b) * d + c) * d + e + y


def max_weight(w, ref, penalty):
    """getMaxWeight (polisher.c:1025-1038)."""
    mw, mi = 0.0, -1
    for j, v in enumerate(w):
        if j != ref and v >= mw:
            mw, mi = v, j
    wr = w[ref] if 0 <= ref < len(w) else 0.0
    return ref if (wr * penalty > mw or mi < 0) else mi


def consensus(poa, rle=True, penalty=0.5):
    """poa_getConsensus (polisher.c:1040-1275): forward pass over the graph,
    greedy traceback; returns (expanded consensus, poaToConsensusMap)."""
    import math
    nodes = poa.nodes
    n = len(nodes)
    tot_out = [0.0] * n
    fwd = [LOG_ZERO] * (n + 1)
    fwd[0] = 0.0
    match_fwd = [0.0] * n
    incoming = [[] for _ in range(n + 1)]
    for i, node in enumerate(nodes):
        for e in node.deletes:
            incoming[i + e[0] + 1].append((i, e))
    for i, node in enumerate(nodes):
        indel = 0.0
        for e in node.inserts:
            indel += e[1] + e[2]
        for e in node.deletes:
            indel += e[1] + e[2]
        mt = 0.0
        if i == 0:
            if n == 1:
                mt = 1.0
            else:
                for j in range(1, n):
                    for v in nodes[j].base_weights:
                        mt += v
                mt /= n - 1
                mt -= indel
        else:
            for v in node.base_weights:
                mt += v
            mt -= indel
        mt = 0.0001 if mt <= 0.0 else mt
        tot_out[i] = mt + indel
        for e in node.inserts:
            fwd[i + 1] = log_add(fwd[i + 1], fwd[i] + math.log((e[1] + e[2]) / tot_out[i]))
        for e in node.deletes:
            t = i + e[0] + 1
            fwd[t] = log_add(fwd[t], fwd[i] + math.log((e[1] + e[2]) / tot_out[i]))
        match_fwd[i] = fwd[i] + math.log(mt / tot_out[i])
        fwd[i + 1] = log_add(fwd[i + 1], match_fwd[i])
    cmap = [-1] * (n - 1)
    pieces = []
    running = 0
    prev = "-"
    i = n
    while i > 0:
        if i < n:
            node = nodes[i]
            mb = max_weight(node.base_weights, char_to_symbol(node.base), penalty)
            base = "ACGTN"[mb]
            if rle:
                rc = max_weight(node.rc_weights, node.repeat_count, penalty)
                rc = 1 if rc == 0 else rc
                pieces.append(base * rc)
                if prev != base:
                    cmap[i - 1] = running
                    running += 1
                prev = base
            else:
                pieces.append(base)
                cmap[i - 1] = running
                running += 1
        max_ins, tot_ins, best_ins = LOG_ZERO, LOG_ZERO, None
        for e in nodes[i - 1].inserts:
            pr = math.log((e[1] + e[2]) / tot_out[i - 1]) + fwd[i - 1]
            if pr > max_ins:
                max_ins, best_ins = pr, e
            tot_ins = log_add(tot_ins, pr)
        max_del, tot_del, best_del = LOG_ZERO, LOG_ZERO, None
        for src, e in incoming[i]:
            pr = math.log((e[1] + e[2]) / tot_out[src]) + fwd[src]
            if pr > max_del:
                max_del, best_del = pr, e
            tot_del = log_add(tot_del, pr)
        if match_fwd[i - 1] >= tot_del and match_fwd[i - 1] >= tot_ins:
            i -= 1
        elif tot_ins >= tot_del:
            s = best_ins[0]
            pieces.append(expand(s))
            if rle:
                running += len(s[0]) + (0 if s[0][-1] != prev else -1)
                prev = s[0][0]
            else:
                running += len(s[0])
            i -= 1
        else:
            i -= best_del[0] + 1
    expanded = "".join(reversed(pieces))
    clen = len(rle_string(expanded)[0]) if rle else len(expanded)
    assert running == clen
    return expanded, [-1 if v == -1 else clen - 1 - v for v in cmap]


def rle_string(s):
    return rle(s)

