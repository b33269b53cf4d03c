"""Seeded synthetic workloads for the realignment path (no datasets exist here).

Each workload is a list of chunk-level read/reference pairs in run-length
(RLE) space, the shape poa_realign (polisher.c:656-703) sees per chunk:
  ref_sym   chunk reference, RLE symbols (run length << 8 | base)
  read_sym  read clipped to the chunk, RLE symbols
  anchors   strictly increasing (x, y, expansion) RLE match pairs of the
            true alignment (the CIGAR-derived anchors of
            htsIntegration.c:347-650, expansion = band half-width W)
  strand    1 forward / 0 reverse (selects the state machine)

Reads are simulated in reference orientation (as BAM stores them) with
per-base error events sub:ins:del = 4:3:3; 70% of insertions repeat the
current base (a +-1 homopolymer run-length change), deletions fall on the
base where they are drawn.

Workloads (sizes follow the benchmark configurations):
  config0  5 kb uniform-ACGT reference, 20 full-length reads, 10% error, W=32
  config1  400 kb reference (GC 45%, runs ~ Geometric(0.45)), 40 chunks of
           10 kb with 500 bp boundaries, 60x log-normal reads (mean 8 kb,
           sd 6 kb, clipped to 0.5-40 kb), 9% error, W=10
  config2  5 Mb reference (GC 50%, runs ~ Geometric(0.5), 1% tandem
           repeats), 100 kb chunks with 5 kb boundaries, 50x reads (mean
           20 kb, sd 15 kb, max 100 kb), 8% error, W=10
  config4  256 pairs of a 100 kb read against its own 100 kb reference
           (GC 50%, runs ~ Geometric(0.65): RLE ~65 kb), structural indels of
           50-500 bp covering 0.5% of the length, then 12% error; band W in
           {128, 256, 512, 1024}; for the five-state machine
"""
from __future__ import annotations

import dataclasses
from typing import List

import numpy as np

MAX_REPEAT_EXCLUSIVE = 51
_COMP = np.array([3, 2, 1, 0], dtype=np.uint8)


@dataclasses.dataclass
class Pair:
    ref_sym: np.ndarray
    read_sym: np.ndarray
    anchors: np.ndarray  # (n, 3) int64
    strand: int
    chunk: int = 0
    read: int = 0
    ref_runs: np.ndarray = None   # uncapped run lengths of ref_sym (the RleString's repeat counts)
    read_runs: np.ndarray = None  # uncapped run lengths of read_sym


@dataclasses.dataclass
class Workload:
    name: str
    pairs: List[Pair]
    W: int
    ref_len: int = 0
    chunks: int = 1


def _ref_uniform(rng, n):
    return rng.integers(0, 4, size=n, dtype=np.uint8)


def _ref_runs(rng, n, gc, p_geom, tandem_frac=0.0):
    """Reference built run by run: base ~ GC composition redrawn if equal to
    the previous run's base, run length ~ Geometric(p) >= 1."""
    pa = (1.0 - gc) / 2.0
    probs = np.array([pa, gc / 2.0, gc / 2.0, pa])
    nruns = int(n / (1.0 / p_geom) * 1.3) + 16
    bases = rng.choice(4, size=nruns, p=probs).astype(np.uint8)
    # redraw equal neighbours until none remain
    for _ in range(64):
        eq = np.flatnonzero(bases[1:] == bases[:-1]) + 1
        if eq.size == 0:
            break
        bases[eq] = rng.choice(4, size=eq.size, p=probs)
    runs = rng.geometric(p_geom, size=nruns)
    seq = np.repeat(bases, runs)[:n]
    if tandem_frac > 0:
        total = 0
        while total < tandem_frac * n:
            k = int(rng.integers(2, 7))
            copies = int(rng.integers(10, 41))
            unit = rng.integers(0, 4, size=k, dtype=np.uint8)
            rep = np.tile(unit, copies)
            pos = int(rng.integers(0, max(1, n - rep.size)))
            seq[pos:pos + rep.size] = rep[: max(0, min(rep.size, n - pos))]
            total += rep.size
    return seq


def _mutate(rng, ref, err):
    """Per-base error events. Returns (read bases, read position of each
    reference base or -1 when deleted)."""
    n = ref.size
    u = rng.random(n)
    ev = np.zeros(n, dtype=np.int8)  # 0 copy, 1 sub, 2 ins, 3 del
    ev[u < err * 0.4] = 1
    ev[(u >= err * 0.4) & (u < err * 0.7)] = 2
    ev[(u >= err * 0.7) & (u < err)] = 3
    outlen = np.where(ev == 3, 0, np.where(ev == 2, 2, 1))
    ends = np.cumsum(outlen)
    starts = ends - outlen
    read = np.empty(int(ends[-1]) if n else 0, dtype=np.uint8)
    keep = outlen > 0
    first = ref.copy()
    sub = ev == 1
    first[sub] = (ref[sub] + rng.integers(1, 4, size=int(sub.sum()), dtype=np.uint8)) % 4
    read[starts[keep]] = first[keep]
    ins = ev == 2
    homo = rng.random(n) < 0.7
    ins_base = np.where(homo, ref, rng.integers(0, 4, size=n, dtype=np.uint8))
    read[starts[ins] + 1] = ins_base[ins]
    pos = np.where(keep, starts, -1)
    return read, pos


def _rle(bases):
    if bases.size == 0:
        return bases, np.zeros(0, np.int64), np.zeros(0, np.int64)
    new = np.r_[True, bases[1:] != bases[:-1]]
    starts = np.flatnonzero(new)
    runs = np.diff(np.r_[starts, bases.size])
    run_id = np.cumsum(new) - 1
    return bases[starts], runs, run_id


def _symbols(rbases, runs):
    rl = np.minimum(runs, MAX_REPEAT_EXCLUSIVE - 1).astype(np.uint32)
    return ((rl << 8) | rbases.astype(np.uint32)).astype(np.uint16)


def _ref_rle(ref_seg):
    """Run-length form of a reference segment: (symbols, run lengths, run id
    of every base), shared by every pair of the segment."""
    rb, rr, rid = _rle(ref_seg)
    return _symbols(rb, rr), rr.astype(np.uint64), rid


def _pair(ref_seg, read_seg, ref_idx, read_idx, W, strand, chunk, read_id, ref_rle=None):
    """RLE both segments and derive strictly increasing RLE anchors from the
    aligned base pairs (ref_idx[k], read_idx[k]).  ref_rle: the segment's
    _ref_rle, when several pairs share it."""
    rsym, rruns, rid = ref_rle if ref_rle is not None else _ref_rle(ref_seg)
    qb, qr, qid = _rle(read_seg)
    X = rid[ref_idx]
    Y = qid[read_idx]
    if X.size:
        keep = np.r_[True, (X[1:] > X[:-1]) & (Y[1:] > Y[:-1])]
        X, Y = X[keep], Y[keep]
    anchors = np.stack([X, Y, np.full(X.size, W, dtype=np.int64)], axis=1).astype(np.int64)
    return Pair(rsym, _symbols(qb, qr), anchors, strand, chunk, read_id, rruns, qr.astype(np.uint64))


def config0(n_reads=20, ref_len=5000, seed=0, err=0.10, W=32, strands=(1,)):
    rng = np.random.default_rng(seed)
    ref = _ref_uniform(rng, ref_len)
    pairs = []
    for r in range(n_reads):
        read, pos = _mutate(rng, ref, err)
        al = np.flatnonzero(pos >= 0)
        strand = strands[r % len(strands)]
        pairs.append(_pair(ref, read, al, pos[al], W, strand, 0, r))
    return Workload("config0", pairs, W, ref_len, 1)


def _lognormal_lengths(rng, n, mean, sd, lo, hi):
    s2 = np.log(1.0 + (sd / mean) ** 2)
    mu = np.log(mean) - s2 / 2.0
    L = rng.lognormal(mu, np.sqrt(s2), size=n)
    return np.clip(L, lo, hi).astype(np.int64)


def chunked(name, seed, ref_len, gc, p_geom, tandem, chunk_size, boundary, depth, mean, sd, lo, hi, err, W,
            max_chunks=None):
    rng = np.random.default_rng(seed)
    ref = _ref_runs(rng, ref_len, gc, p_geom, tandem)
    nreads = int(round(depth * ref_len / mean))
    lens = np.minimum(_lognormal_lengths(rng, nreads, mean, sd, lo, hi), ref_len)
    starts = rng.integers(0, np.maximum(1, ref_len - lens + 1))
    strands = rng.integers(0, 2, size=nreads)
    nchunks = (ref_len + chunk_size - 1) // chunk_size
    if max_chunks is not None:
        nchunks = min(nchunks, max_chunks)
    chunk_reads = [[] for _ in range(nchunks)]
    for r in range(nreads):
        s, e = int(starts[r]), int(starts[r] + lens[r])
        c0 = max(0, (s - boundary) // chunk_size)
        c1 = min(nchunks - 1, (e - 1 + boundary) // chunk_size)
        if c0 > c1:
            continue
        read, pos = _mutate(rng, ref[s:e], err)
        for c in range(c0, c1 + 1):
            chunk_reads[c].append((r, s, read, pos))
    pairs = []
    for c in range(nchunks):
        cs = max(0, c * chunk_size - boundary)
        ce = min(ref_len, (c + 1) * chunk_size + boundary)
        ref_seg = ref[cs:ce]
        ref_rle = _ref_rle(ref_seg)
        for (r, s, read, pos) in chunk_reads[c]:
            # clip the read to the reference bases it aligns inside the chunk
            lo_i, hi_i = max(cs - s, 0), min(ce - s, pos.size)
            if hi_i <= lo_i:
                continue
            p = pos[lo_i:hi_i]
            al = np.flatnonzero(p >= 0)
            if al.size < 2:
                continue
            q0, q1 = int(p[al[0]]), int(p[al[-1]])
            read_seg = read[q0:q1 + 1]
            ref_idx = al + (lo_i + s - cs)
            read_idx = p[al] - q0
            pairs.append(_pair(ref_seg, read_seg, ref_idx, read_idx, W, int(strands[r]), c, r, ref_rle))
    return Workload(name, pairs, W, ref_len, nchunks)


def config1(seed=1, max_chunks=None):
    return chunked("config1", seed, 400_000, 0.45, 0.45, 0.0, 10_000, 500, 60, 8000, 6000, 500, 40_000, 0.09, 10,
                   max_chunks)


def config2(seed=2, max_chunks=None):
    return chunked("config2", seed, 5_000_000, 0.50, 0.5, 0.01, 100_000, 5_000, 50, 20_000, 15_000, 500, 100_000,
                   0.08, 10, max_chunks)


def _structural(rng, ref, frac, lo=50, hi=500):
    """Insert and delete segments of lo..hi bases (half each, non-overlapping)
    until they cover frac of the length.  Returns (sequence, source
    reference index of each base or -1 for inserted bases)."""
    n = ref.size
    events, covered, tries = [], 0, 0
    while covered < frac * n and tries < 1000:
        tries += 1
        L = int(rng.integers(lo, hi + 1))
        at = int(rng.integers(hi, max(hi + 1, n - 2 * hi)))
        if any(abs(at - a) < 2 * hi for a, _, _ in events):
            continue
        events.append((at, L, bool(rng.random() < 0.5)))
        covered += L
    events.sort()
    out, src, prev = [], [], 0
    for at, L, is_del in events:
        out.append(ref[prev:at])
        src.append(np.arange(prev, at))
        if is_del:
            prev = at + L
        else:
            out.append(rng.integers(0, 4, size=L, dtype=np.uint8))
            src.append(np.full(L, -1, dtype=np.int64))
            prev = at
    out.append(ref[prev:])
    src.append(np.arange(prev, n))
    return np.concatenate(out), np.concatenate(src)


def config4(n_pairs=256, length=100_000, seed=4, err=0.12, sv_frac=0.005, W=128, strands=(1, 0)):
    rng = np.random.default_rng(seed)
    pairs = []
    for r in range(n_pairs):
        ref = _ref_runs(rng, length, 0.5, 0.65)
        sv, src = _structural(rng, ref, sv_frac)
        read, pos = _mutate(rng, sv, err)
        keep = (src >= 0) & (pos >= 0)
        pairs.append(_pair(ref, read, src[keep], pos[keep], W, strands[r % len(strands)], r, r))
    return Workload("config4", pairs, W, length, n_pairs)


def _rle_pair(bases):
    rb, runs, _ = _rle(bases)
    return (_BASE_CHARS[rb].tobytes(), runs.astype(np.uint64))


_BASE_CHARS = np.frombuffer(b"ACGT", dtype=np.uint8)


def bubbles(n_bubbles=4000, depth=60, seed=5, err=0.09, region=(8, 40), flank=6, max_alleles=4, dup_frac=0.2):
    """Bubble-graph likelihood workload (bubbleGraph.c:895-928 shapes): per
    bubble, a reference region of `region` bases plus `flank` bases each
    side (runs ~ Geometric(0.45), GC 45%), the reference allele and 1..3
    alternatives (a substitution, a +-1 run-length change or a 1-3 base
    indel inside the region), and `depth` read substrings, each a copy of one
    allele with `err` per-base errors; a `dup_frac` share of reads repeats an
    earlier read's substring exactly (identical substrings are common in
    short windows).  Returns [(alleles, reads)] with run-length strings
    (bases, repeatCounts) and reads ((bases, counts), strand)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_bubbles):
        L = int(rng.integers(region[0], region[1] + 1)) + 2 * flank
        ref = _ref_runs(rng, L, 0.45, 0.45)
        alleles = [ref]
        for _ in range(int(rng.integers(1, max_alleles))):
            a = ref.copy()
            pos = int(rng.integers(flank, L - flank))
            kind = int(rng.integers(0, 4))
            if kind == 0:
                a[pos] = (a[pos] + 1 + int(rng.integers(0, 3))) % 4
            elif kind == 1:
                a = np.insert(a, pos, a[pos])
            elif kind == 2:
                a = np.delete(a, pos)
            else:
                k = int(rng.integers(1, 4))
                a = np.insert(a, pos, rng.integers(0, 4, size=k).astype(a.dtype))
            if not any(a.size == b.size and np.array_equal(a, b) for b in alleles):
                alleles.append(a)
        reads = []
        for r in range(depth):
            if reads and rng.random() < dup_frac:
                reads.append(reads[int(rng.integers(0, len(reads)))])
                continue
            src = alleles[int(rng.integers(0, len(alleles)))]
            mut, _ = _mutate(rng, src, err)
            if mut.size == 0:
                mut = src.copy()
            reads.append((_rle_pair(mut), int(rng.integers(0, 2))))
        out.append(([_rle_pair(a) for a in alleles], reads))
    return out


@dataclasses.dataclass
class PolishChunk:
    start: int               # draft coordinates of the chunk (with its boundaries)
    end: int
    ref: tuple               # run-length draft segment (bases, counts)
    reads: list              # [((bases, counts), strand)]
    anchors: list            # per read: (n, 3) RLE anchors (draft x, read y, expansion)


def _plain(bases):
    """Every base its own run (rleString_construct_no_rle)."""
    return bases, np.ones(bases.size, np.int64), np.arange(bases.size)


def polish_workload(seed=7, genome_len=30000, draft_err=0.03, read_err=0.08, depth=20, mean=3000, sd=1500,
                    chunk_size=10000, boundary=250, W=10, rle=True, gc=0.45, p_geom=0.45, tandem=0.0, lo=500,
                    hi=None):
    """A draft to polish: a truth genome (GC `gc`, runs ~ Geometric(p_geom),
    `tandem` of it in 2-6-mer tandem repeats), a draft = truth with
    `draft_err` errors, reads of log-normal lengths (clipped to lo..hi,
    default the genome) sampled from the truth at `read_err`; chunks of the
    draft with `boundary` bases each side, every read clipped to each chunk it
    overlaps, RLE anchors from the reads' true placement on the draft
    (rle=False: plain strings, one symbol per base).  Returns (truth bases,
    draft bases, chunks)."""
    enc = _rle if rle else _plain
    rng = np.random.default_rng(seed)
    truth = _ref_runs(rng, genome_len, gc, p_geom, tandem)
    draft, dpos = _mutate(rng, truth, draft_err)   # dpos: draft position of each truth base
    nreads = int(round(depth * genome_len / mean))
    lens = np.minimum(_lognormal_lengths(rng, nreads, mean, sd, lo, genome_len if hi is None else hi), genome_len)
    starts = rng.integers(0, np.maximum(1, genome_len - lens + 1))
    strands = rng.integers(0, 2, size=nreads)
    chars = np.frombuffer(b"ACGT", np.uint8)
    dlen = draft.size
    nchunks = (dlen + chunk_size - 1) // chunk_size
    chunks = []
    for c in range(nchunks):
        cs, ce = max(0, c * chunk_size - boundary), min(dlen, (c + 1) * chunk_size + boundary)
        rb, rr, rid = enc(draft[cs:ce])
        chunks.append(PolishChunk(cs, ce, (chars[rb].tobytes(), rr.astype(np.uint64)), [], []))
        chunks[-1]._rid = rid
    for r in range(nreads):
        s, e = int(starts[r]), int(starts[r] + lens[r])
        read, pos = _mutate(rng, truth[s:e], read_err)
        dp = dpos[s:e]
        ok = (pos >= 0) & (dp >= 0)
        D, Y = dp[ok], pos[ok]
        if D.size < 2:
            continue
        # only chunks whose span can hold two of the read's draft positions
        c0 = max(0, (int(D.min()) - boundary) // chunk_size - 1)
        c1 = min(nchunks, (int(D.max()) + boundary) // chunk_size + 2)
        for c in range(c0, c1):
            ch = chunks[c]
            sel = (D >= ch.start) & (D < ch.end)
            if sel.sum() < 2:
                continue
            d0, y0, y1 = D[sel], int(Y[sel][0]), int(Y[sel][-1])
            seg = read[y0:y1 + 1]
            qb, qr, qid = enc(seg)
            X = ch._rid[d0 - ch.start]
            Yr = qid[Y[sel] - y0]
            keep = np.r_[True, (X[1:] > X[:-1]) & (Yr[1:] > Yr[:-1])]
            X, Yr = X[keep], Yr[keep]
            ch.reads.append(((chars[qb].tobytes(), qr.astype(np.uint64)), int(strands[r])))
            ch.anchors.append(np.stack([X, Yr, np.full(X.size, W)], axis=1).astype(np.int64))
    for ch in chunks:
        del ch._rid
    return truth, draft, chunks


def polish_workload_rle(repeat_log10, seed=9, runs=12000, draft_err=0.03, read_err=0.05, depth=20, mean_runs=1500,
                        chunk_runs=4000, boundary_runs=100, W=10):
    """A run-length draft to polish, with reads drawn from the run-length
    error model the params encode: truth = `runs` runs (GC 45%, lengths ~
    Geometric(0.45)); the draft = truth with `draft_err` of its runs given a
    length off by one (so draft run i is truth run i); every read covers a
    span of truth runs, each run's observed length drawn from
    P(observed | true length) = 10^repeat_log10[base row][true][observed]
    (the reverse strand reads the complement's row) and, at `read_err`, a run
    dropped or an extra run inserted.  Chunks of the draft in run coordinates.
    Returns (truth (bases, counts), draft (bases, counts), chunks)."""
    rng = np.random.default_rng(seed)
    tab = np.asarray(repeat_log10, dtype=np.float64).reshape(-1, MAX_REPEAT_EXCLUSIVE, MAX_REPEAT_EXCLUSIVE)[:4]
    prob = 10.0 ** tab
    prob /= prob.sum(axis=2, keepdims=True)
    cdf = np.cumsum(prob, axis=2)
    pa = 0.275
    p4 = np.array([pa, 0.225, 0.225, pa])
    bases = rng.choice(4, size=runs, p=p4).astype(np.uint8)
    for _ in range(64):
        eq = np.flatnonzero(bases[1:] == bases[:-1]) + 1
        if eq.size == 0:
            break
        bases[eq] = rng.choice(4, size=eq.size, p=p4)
    lens = np.minimum(rng.geometric(0.45, size=runs), MAX_REPEAT_EXCLUSIVE - 1)
    dlens = lens.copy()
    flip = rng.random(runs) < draft_err
    dlens[flip] = np.maximum(1, dlens[flip] + np.where(rng.random(int(flip.sum())) < 0.5, -1, 1))
    chars = np.frombuffer(b"ACGT", np.uint8)
    nreads = int(round(depth * runs / mean_runs))
    chunks = []
    for c in range((runs + chunk_runs - 1) // chunk_runs):
        cs, ce = max(0, c * chunk_runs - boundary_runs), min(runs, (c + 1) * chunk_runs + boundary_runs)
        chunks.append(PolishChunk(cs, ce, (chars[bases[cs:ce]].tobytes(), dlens[cs:ce].astype(np.uint64)), [], []))
    for _ in range(nreads):
        n = int(min(runs, max(50, rng.normal(mean_runs, mean_runs / 3))))
        s0 = int(rng.integers(0, runs - n + 1))
        strand = int(rng.integers(0, 2))
        rb, rl, rx = [], [], []  # read runs and the truth run each came from (-1: inserted)
        for t in range(s0, s0 + n):
            u = rng.random()
            if u < read_err / 2:
                continue  # run dropped
            if u < read_err:
                nb = int(rng.integers(0, 4))
                if (not rb or nb != rb[-1]) and nb != bases[t]:
                    rb.append(nb)
                    rl.append(1)
                    rx.append(-1)
            b = int(bases[t])
            row = b if strand else 3 - b
            o = int(np.searchsorted(cdf[row, lens[t]], rng.random() * cdf[row, lens[t], -1]))
            o = max(1, min(o, MAX_REPEAT_EXCLUSIVE - 1))
            if rb and rb[-1] == b:  # a dropped run joined two equal runs: merge them
                rl[-1] += o
                continue
            rb.append(b)
            rl.append(o)
            rx.append(t)
        rb, rl, rx = np.array(rb, np.uint8), np.array(rl, np.uint64), np.array(rx, np.int64)
        for ch in chunks:
            sel = np.flatnonzero((rx >= ch.start) & (rx < ch.end))
            if sel.size < 2:
                continue
            a, e = sel[0], sel[-1] + 1
            X, Y = rx[sel] - ch.start, sel - a
            keep = np.r_[True, (X[1:] > X[:-1]) & (Y[1:] > Y[:-1])]
            ch.reads.append(((chars[rb[a:e]].tobytes(), rl[a:e].copy()), strand))
            ch.anchors.append(np.stack([X[keep], Y[keep], np.full(int(keep.sum()), W)], axis=1).astype(np.int64))
    return (chars[bases].tobytes(), lens.astype(np.uint64)), (chars[bases].tobytes(), dlens.astype(np.uint64)), chunks


def banded_edit_distance(a, b, band=400):
    """Levenshtein distance restricted to |i - j| <= band (exact when the
    optimal path stays inside the band, as it does for these near-equal strings)."""
    a = np.frombuffer(a.encode(), np.uint8)
    b = np.frombuffer(b.encode(), np.uint8)
    n, m = a.size, b.size
    big = 1 << 30
    prev = np.full(m + 1, big, np.int64)
    prev[: min(m, band) + 1] = np.arange(min(m, band) + 1)
    for i in range(1, n + 1):
        lo, hi = max(1, i - band), min(m, i + band)
        cur = np.full(m + 1, big, np.int64)
        if i <= band:
            cur[0] = i
        if lo <= hi:
            sub = prev[lo - 1:hi] + (a[i - 1] != b[lo - 1:hi])
            dele = prev[lo:hi + 1] + 1
            best = np.minimum(sub, dele)
            # insertions run along the row: a prefix-minimum of best - j, plus j
            j = np.arange(lo, hi + 1)
            left = cur[lo - 1] + 1
            run = np.minimum.accumulate(np.minimum(best - j, np.r_[left - lo, np.full(hi - lo, big)]))
            cur[lo:hi + 1] = np.minimum(best, run + j)
        prev = cur
    return int(prev[m])


def edit_distance(a, b):
    """Levenshtein distance by diagonal transition (Ukkonen; Landau and
    Vishkin): for e = 0, 1, ... the furthest row reached on every diagonal
    with e edits, each slide along equal characters done by numpy; O((n + m)
    d) for d edits, for near-equal strings of any length."""
    A = np.frombuffer(a.encode(), np.uint8)
    B = np.frombuffer(b.encode(), np.uint8)
    n, m = A.size, B.size

    def slide(i, j):
        step = 256
        while i < n and j < m:
            L = min(step, n - i, m - j)
            neq = np.flatnonzero(A[i:i + L] != B[j:j + L])
            if neq.size:
                return i + int(neq[0])
            i += L
            j += L
            step = min(step * 2, 1 << 16)
        return i

    goal = m - n  # the diagonal k = j - i of the end cell
    prev = {0: slide(0, 0)}
    e = 0
    while not (goal in prev and prev[goal] >= n):
        e += 1
        cur = {}
        for k in range(max(-e, -n), min(e, m) + 1):
            best = -1
            if k in prev:
                best = prev[k] + 1          # substitution
            if k - 1 in prev:
                best = max(best, prev[k - 1])  # b's character inserted (j + 1)
            if k + 1 in prev:
                best = max(best, prev[k + 1] + 1)  # a's character deleted (i + 1)
            if best < 0:
                continue
            i = min(best, n, m - k)
            if i < 0 or i + k < 0:
                continue
            cur[k] = slide(i, i + k)
        prev = cur
    return e
