"""Comparison helpers for GPU-vs-oracle parity (test infrastructure)."""
import numpy as np

PROB1 = 1e7


def as_map(lst):
    """(weight, x, y) rows -> {(x, y): weight}; asserts (x, y) unique."""
    d = {}
    for row in lst:
        w, x, y = int(row[0]), int(row[1]), int(row[2])
        assert (x, y) not in d, ("duplicate", x, y)
        d[(x, y)] = w
    return d


def compare(gpu, ora, threshold, tol=1e-4):
    """Max |p_gpu - p_oracle| over the union of (x, y) keys.  A key present
    on one side only must be within tol of the threshold on the other side
    (FP32 vs FP64 decide p >= threshold differently only there).  Returns
    (max_abs_err, n_common, n_only)."""
    g, o = as_map(gpu), as_map(ora)
    err, only = 0.0, 0
    common = set(g) & set(o)
    for k in common:
        err = max(err, abs(g[k] - o[k]) / PROB1)
    edge = (threshold + tol) * PROB1 + 1
    for k in set(g) - set(o):
        only += 1
        assert g[k] <= edge, ("gpu-only entry far above threshold", k, g[k])
    for k in set(o) - set(g):
        only += 1
        assert o[k] <= edge, ("oracle-only entry far above threshold", k, o[k])
    return err, len(common), only


def order_consistent(gpu, ora):
    """Entries present on both sides appear in the same relative order."""
    o_idx = {(int(r[1]), int(r[2])): i for i, r in enumerate(ora)}
    seq = [o_idx[(int(r[1]), int(r[2]))] for r in gpu if (int(r[1]), int(r[2])) in o_idx]
    return all(a < b for a, b in zip(seq, seq[1:]))


def as_rows(arr):
    """numpy structured (weight, x, y) -> int64 (n, 3)."""
    if arr.dtype.names:
        return np.stack([arr["weight"].astype(np.int64), arr["x"].astype(np.int64), arr["y"].astype(np.int64)],
                        axis=1) if len(arr) else np.zeros((0, 3), np.int64)
    return np.asarray(arr, np.int64).reshape(-1, 3)


def keyed(res, li):
    """Every pair's tuples of list li (0 matches, 1 gap-X, 2 gap-Y) from a
    batch's fetch(): (sorted keys pair << 42 | (x + 1) << 21 | (y + 1),
    weights in key order); x or y is -1 for a gap before the first base."""
    keys, ws = [], []
    for i, r in enumerate(res):
        a = r[li]
        keys.append((np.int64(i) << 42) | ((a["x"].astype(np.int64) + 1) << 21) | (a["y"].astype(np.int64) + 1))
        ws.append(a["weight"].astype(np.int64))
    k = np.concatenate(keys) if keys else np.zeros(0, np.int64)
    w = np.concatenate(ws) if ws else np.zeros(0, np.int64)
    o = np.argsort(k, kind="stable")
    return k[o], w[o]


def batch_list_err(ra, rb):
    """Two fetch() results of the same pairs: (max |p_a - p_b| over the
    tuples both hold, the largest probability held by one side only, the
    number of repeated (pair, x, y) keys, tuples held by both)."""
    worst, edge, dups, both = 0.0, 0.0, 0, 0
    for li in range(3):
        ka, wa = keyed(ra, li)
        kb, wb = keyed(rb, li)
        dups += int(np.sum(ka[1:] == ka[:-1]) + np.sum(kb[1:] == kb[:-1]))
        pa = np.minimum(np.searchsorted(kb, ka), max(kb.size - 1, 0))
        inA = (kb[pa] == ka) if kb.size else np.zeros(ka.size, bool)
        pb = np.minimum(np.searchsorted(ka, kb), max(ka.size - 1, 0))
        inB = (ka[pb] == kb) if ka.size else np.zeros(kb.size, bool)
        both += int(inA.sum())
        if inA.any():
            worst = max(worst, float(np.abs(wa[inA] - wb[pa[inA]]).max()) / PROB1)
        for w, only in ((wa, ~inA), (wb, ~inB)):
            if only.any():
                edge = max(edge, float(w[only].max()) / PROB1)
    return worst, edge, dups, both
