"""POA augmentation straight from a GPU batch's posterior lists
(mp_batch_poa_augment) against the per-read entry point fed with the fetched
lists and against the oracle restatement of poa_augment."""
import numpy as np
import pytest

from marginpolish_amd import pairhmm as ph
from marginpolish_amd import poa as mpoa
from marginpolish_amd import sim
from oracle import poa as P
from tests.conftest import PARAMS
from tests.poa_compare import assert_same_graph, native_as_oracle, oracle_lists

pytestmark = pytest.mark.gpu


class _AsOracle:
    def __init__(self, nodes):
        self.nodes = nodes


def _chunk_setup(w):
    chunks = sorted(set(p.chunk for p in w.pairs))
    first = {}
    for p in w.pairs:
        first.setdefault(p.chunk, p)
    refs = [mpoa.rle_from_symbols(first[c].ref_sym, first[c].ref_runs) for c in chunks]
    pair_poa = np.array([chunks.index(p.chunk) for p in w.pairs], dtype=np.int64)
    read_no = np.zeros(len(w.pairs), dtype=np.int64)
    seen = {}
    for i, p in enumerate(w.pairs):
        read_no[i] = seen.get(p.chunk, 0)
        seen[p.chunk] = read_no[i] + 1
    reads = [mpoa.rle_from_symbols(p.read_sym, p.read_runs) for p in w.pairs]
    return refs, pair_poa, read_no, reads


@pytest.fixture(scope="module")
def config1_batch():
    w = sim.config1(max_chunks=3)
    pp = ph.PolishParams(PARAMS)
    pairs = [(p.ref_sym, p.read_sym, p.anchors, p.strand) for p in w.pairs]
    print("config1 3 chunks: %d pairs" % len(pairs), flush=True)
    b = ph.Batch(pp.machines, pairs, pp.p, crop_reference=True)
    b.run()
    lists = b.fetch()
    yield w, b, lists
    b.close()


@pytest.mark.parametrize("cmp", [True, False])
def test_batch_augment_equals_per_read_augment(config1_batch, cmp):
    w, b, lists = config1_batch
    refs, pair_poa, read_no, reads = _chunk_setup(w)
    g_batch = [mpoa.Poa(r) for r in refs]
    mpoa.augment_batch(b, g_batch, pair_poa, reads, read_no, compare_repeat_counts=cmp)
    g_read = [mpoa.Poa(r) for r in refs]
    for i, p in enumerate(w.pairs):
        m, gx, gy, _ = lists[i]
        g_read[pair_poa[i]].augment(reads[i], p.strand, read_no[i], m, gy, gx, cmp)
    for a, c in zip(g_batch, g_read):
        assert_same_graph(native_as_oracle(a), _AsOracle(native_as_oracle(c)))
    s = g_batch[0].sizes()
    print("graph 0:", s, flush=True)
    assert s["inserts"] > 100 and s["deletes"] > 100 and s["nodeObservations"] > 10000
    # one insert observation per insert event, and a graph with 40000 or more
    # events runs its insert task as node-range shards: the comparison above
    # then covers the shard merge
    assert s["insertObservations"] >= 40000


def test_batch_augment_thread_counts_agree(config1_batch):
    w, b, lists = config1_batch
    refs, pair_poa, read_no, reads = _chunk_setup(w)
    out = []
    for threads in (1, 3, 0):
        g = [mpoa.Poa(r) for r in refs]
        mpoa.augment_batch(b, g, pair_poa, reads, read_no, threads=threads)
        out.append([native_as_oracle(x) for x in g])
    for other in out[1:]:
        for a, c in zip(out[0], other):
            assert_same_graph(a, _AsOracle(c))


def test_batch_augment_matches_oracle(config1_batch):
    """One chunk's graph built from the GPU lists by the library and by the
    oracle restatement (first 40 reads of the chunk)."""
    w, b, lists = config1_batch
    refs, pair_poa, read_no, reads = _chunk_setup(w)
    idx = [i for i in range(len(w.pairs)) if pair_poa[i] == 0][:40]
    ref = refs[0]
    ref_o = (ref[0].tobytes().decode(), tuple(int(c) for c in ref[1]))
    o = P.Poa(ref_o, 51)
    g = mpoa.Poa(ref)
    for i in idx:
        m, gx, gy, _ = lists[i]
        r = reads[i]
        r_o = (r[0].tobytes().decode(), tuple(int(c) for c in r[1]))
        P.augment(o, r_o, bool(w.pairs[i].strand), int(read_no[i]), oracle_lists(m), oracle_lists(gy),
                  oracle_lists(gx), True)
        g.augment(r, w.pairs[i].strand, read_no[i], m, gy, gx, True)
    assert_same_graph(native_as_oracle(g), o)
    # the maximum-likelihood run lengths of the graph (poa_estimateRepeatCountsUsingBayesianModel)
    from oracle.oracle import PolishParams as OP
    reads_by_no = {int(read_no[i]): reads[i] for i in idx}
    strands_by_no = {int(read_no[i]): w.pairs[i].strand for i in idx}
    n = max(reads_by_no) + 1
    rl = [reads_by_no.get(k, (b"", np.zeros(0, np.uint64))) for k in range(n)]
    st = [strands_by_no.get(k, 1) for k in range(n)]
    got = g.estimate_repeat_counts(ph.PolishParams(PARAMS).machines, rl, st)
    rl_o = [(bytes(r[0]).decode(), tuple(int(c) for c in r[1])) for r in rl]
    want = P.estimate_repeat_counts(o, rl_o, [bool(v) for v in st], OP(PARAMS).repeat)
    assert got.tolist() == want


def test_batch_augment_refuses_bad_input(config1_batch):
    w, b, lists = config1_batch
    refs, pair_poa, read_no, reads = _chunk_setup(w)
    g = [mpoa.Poa(r) for r in refs]
    bad = pair_poa.copy()
    bad[0] = len(g)
    with pytest.raises(Exception):
        mpoa.augment_batch(b, g, bad, reads, read_no)
    # a graph whose reference is shorter than the lists' coordinates
    short = [mpoa.Poa((r[0][:10], r[1][:10])) for r in refs]
    with pytest.raises(Exception):
        mpoa.augment_batch(b, short, pair_poa, reads, read_no)


def test_batch_augment_edge_cases(config1_batch):
    """Pairs of several graphs interleaved in the batch, a graph no pair
    uses, and read numbers given out of order: each graph still sees its
    pairs in batch order, as the per-read calls would."""
    w = sim.config1(max_chunks=2)
    pp = ph.PolishParams(PARAMS)
    idx = list(range(len(w.pairs)))
    rng = np.random.default_rng(4)
    rng.shuffle(idx)  # chunks interleaved
    idx = idx[:120]
    pairs = [(w.pairs[i].ref_sym, w.pairs[i].read_sym, w.pairs[i].anchors, w.pairs[i].strand) for i in idx]
    b = ph.Batch(pp.machines, pairs, pp.p, crop_reference=True)
    b.run()
    lists = b.fetch()
    chunks = sorted(set(p.chunk for p in w.pairs))
    first = {}
    for p in w.pairs:
        first.setdefault(p.chunk, p)
    refs = [mpoa.rle_from_symbols(first[c].ref_sym, first[c].ref_runs) for c in chunks]
    refs.append(refs[0])  # a graph no pair uses
    pair_poa = np.array([chunks.index(w.pairs[i].chunk) for i in idx], dtype=np.int64)
    read_no = rng.permutation(len(idx)).astype(np.int64)
    reads = [mpoa.rle_from_symbols(w.pairs[i].read_sym, w.pairs[i].read_runs) for i in idx]
    g_batch = [mpoa.Poa(r) for r in refs]
    mpoa.augment_batch(b, g_batch, pair_poa, reads, read_no, threads=5)
    g_read = [mpoa.Poa(r) for r in refs]
    for k, i in enumerate(idx):
        m, gx, gy, _ = lists[k]
        g_read[pair_poa[k]].augment(reads[k], w.pairs[i].strand, read_no[k], m, gy, gx, True)
    for a, c in zip(g_batch, g_read):
        assert_same_graph(native_as_oracle(a), _AsOracle(native_as_oracle(c)))
    assert g_batch[-1].sizes()["nodeObservations"] == 0
    b.close()
