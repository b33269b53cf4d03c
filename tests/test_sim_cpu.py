"""Test-workload helpers: the diagonal-transition edit distance against the
banded dynamic programme (exact when the band holds the optimal path)."""
import random

from marginpolish_amd import sim


def _mutate(rng, s, rate):
    out = []
    for c in s:
        r = rng.random()
        if r < rate / 3:
            out.append(rng.choice("ACGT"))
        elif r < 2 * rate / 3:
            out.append(c + rng.choice("ACGT"))
        elif r < rate:
            continue
        else:
            out.append(c)
    return "".join(out)


def test_edit_distance_equals_banded_dp():
    rng = random.Random(3)
    for k in range(60):
        a = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 400)))
        b = _mutate(rng, a, rng.choice([0.0, 0.01, 0.05, 0.2]))
        if k % 7 == 0:
            b = b[rng.randint(0, 30):]
        band = max(len(a), len(b)) + 1
        assert sim.edit_distance(a, b) == sim.banded_edit_distance(a, b, band=band), (k, len(a), len(b))
    assert sim.edit_distance("", "") == 0
    assert sim.edit_distance("", "ACG") == 3
    assert sim.edit_distance("ACGT", "") == 4
    assert sim.edit_distance("ACGT", "TGCA") == 4
    assert sim.edit_distance("AAAA", "AAA") == 1


def test_edit_distance_long_sparse():
    """100 kb strings with a few dozen edits (the chunk-vs-chunk case)."""
    rng = random.Random(5)
    a = "".join(rng.choice("ACGT") for _ in range(100_000))
    b = list(a)
    for p in sorted(rng.sample(range(100, 99_900), 30), reverse=True):
        b[p] = "T" if b[p] != "T" else "G"
    b = "".join(b)
    b = b[:5000] + "ACGTA" + b[5000:50_000] + b[50_003:]
    assert sim.edit_distance(a, b) == 38
