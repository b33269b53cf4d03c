"""Shared helpers: the library's POA graph against the oracle's, field by field."""
import numpy as np

from oracle import poa as P


def native_as_oracle(g):
    """The library's graph (marginpolish_amd.poa.Poa) as oracle.poa nodes."""
    nd = g.nodes()
    n = len(nd["bases"])
    mrc = nd["repeatCountWeights"].shape[1]
    out = [P.Node(nd["bases"][k], int(nd["repeatCounts"][k]), mrc) for k in range(n)]
    for k in range(n):
        out[k].base_weights = [float(v) for v in nd["baseWeights"][k]]
        out[k].rc_weights = [float(v) for v in nd["repeatCountWeights"][k]]
    st, ob = g.observations(0)
    for k in range(n):
        out[k].observations = [(int(o["readNo"]), int(o["offset"]), float(o["weight"])) for o in ob[st[k]:st[k + 1]]]
    ins = g.inserts()
    ist, iob = g.observations(1)
    for q in range(len(ins["node"])):
        a, ln = ins["symbolStart"][q], ins["symbolLength"][q]
        s = (ins["bases"][a:a + ln].tobytes().decode(), tuple(int(c) for c in ins["repeatCounts"][a:a + ln]))
        obs = [(int(o["readNo"]), int(o["offset"]), float(o["weight"])) for o in iob[ist[q]:ist[q + 1]]]
        out[ins["node"][q]].inserts.append([s, float(ins["weightForwardStrand"][q]),
                                            float(ins["weightReverseStrand"][q]), obs])
    dl = g.deletes()
    dst, dob = g.observations(2)
    for q in range(len(dl["node"])):
        obs = [(int(o["readNo"]), int(o["offset"]), float(o["weight"])) for o in dob[dst[q]:dst[q + 1]]]
        out[dl["node"][q]].deletes.append([int(dl["length"][q]), float(dl["weightForwardStrand"][q]),
                                           float(dl["weightReverseStrand"][q]), obs])
    return out


def assert_same_graph(native_nodes, oracle_poa):
    """Bit-exact: bases, counts, every weight, list order, observations."""
    on = oracle_poa.nodes
    assert len(native_nodes) == len(on)
    for k, (a, b) in enumerate(zip(native_nodes, on)):
        ctx = "node %d" % k
        assert a.base == b.base, ctx
        assert a.repeat_count == b.repeat_count, ctx
        assert a.base_weights == b.base_weights, ctx
        assert a.rc_weights == b.rc_weights, ctx
        assert a.observations == b.observations, ctx
        assert [e[:3] for e in a.inserts] == [[(e[0][0], tuple(e[0][1])), e[1], e[2]] for e in b.inserts], ctx
        assert [e[3] for e in a.inserts] == [e[3] for e in b.inserts], ctx
        assert a.deletes == b.deletes, ctx


def oracle_lists(t):
    """An ALIGNED_PAIR_DTYPE array as a list of (weight, x, y) tuples."""
    t = np.asarray(t)
    return list(zip(t["weight"].tolist(), t["x"].tolist(), t["y"].tolist()))
