"""The allocation-free forward log-likelihood entry
(mp_pairhmm_forward_log_prob_noalloc, device memory = one static per-GPU
scratch array) against the allocating entry (mp_pairhmm_forward_log_prob),
both called through the C ABI: two calls in a row, the second on a smaller
pair, so the second reuses the scratch the first left behind."""
import ctypes as C

import numpy as np
import pytest

from marginpolish_amd import _lib, crop, pairhmm, sim
from tests.conftest import PARAMS

pytestmark = pytest.mark.gpu


def _call(fn, eng, model, pair_tuple, bp):
    pa = pairhmm.PairArray([pair_tuple])
    out = C.c_double()
    st = fn(eng.handle, model.handle, pa.array.ctypes.data_as(C.POINTER(_lib.Pair)), C.byref(bp), C.byref(out))
    return st, out.value


def test_noalloc_forward_equals_allocating_entry():
    L = _lib.lib()
    pp = pairhmm.PolishParams(PARAMS)
    eng = pairhmm.default_engine()
    w = sim.config0(n_reads=2, ref_len=4000, seed=5, W=10, strands=(1, 0))
    sx, sy, an, _ = crop.crop_pair(w.pairs[0].ref_sym, w.pairs[0].read_sym, w.pairs[0].anchors)
    sx2, sy2, an2, _ = crop.crop_pair(w.pairs[1].ref_sym, w.pairs[1].read_sym, w.pairs[1].anchors)
    cut = np.asarray(an2)
    keep = cut[(cut[:, 0] < 300) & (cut[:, 1] < 300)]
    small = (sx2[: int(keep[-1, 0]) + 5], sy2[: int(keep[-1, 1]) + 5], keep, 0, True, False)
    big = (sx, sy, an, 1, False, True)
    bp = pairhmm.PairwiseAlignmentParameters(diagonalExpansion=10, dynamicAnchorExpansion=True).c()
    print("case: noalloc forward, %d then %d symbols" % (len(sx) + len(sy), len(small[0]) + len(small[1])),
          flush=True)
    for pt in (big, small):
        st_ref, ref = _call(L.mp_pairhmm_forward_log_prob, eng, pp.machines, pt, bp)
        st, got = _call(L.mp_pairhmm_forward_log_prob_noalloc, eng, pp.machines, pt, bp)
        assert st_ref == 0 and st == 0, (st_ref, st, L.mp_last_error())
        assert np.isfinite(got) and got < 0
        assert abs(got - ref) <= 1e-12 * abs(ref), (got, ref)


def test_noalloc_forward_refuses_oversize_and_wide_bands():
    L = _lib.lib()
    pp = pairhmm.PolishParams(PARAMS)
    eng = pairhmm.default_engine()
    bp = pairhmm.PairwiseAlignmentParameters(diagonalExpansion=10).c()
    n = 2_100_000  # 4.2 M anti-diagonals: more than the scratch holds
    long = np.zeros(n, np.uint16)
    print("case: noalloc refusals", flush=True)
    st, _ = _call(L.mp_pairhmm_forward_log_prob_noalloc, eng, pp.machines, (long, long, None, 1), bp)
    assert st == 6 and b"scratch" in L.mp_last_error()
    wide = pairhmm.PairwiseAlignmentParameters(diagonalExpansion=80).c()
    seq = ((np.arange(400) % 4) | 256).astype(np.uint16)  # run length 1
    st, _ = _call(L.mp_pairhmm_forward_log_prob_noalloc, eng, pp.machines, (seq, seq, None, 1), wide)
    assert st == 6 and b"64 cells" in L.mp_last_error()
    # the entry still works after the refusals
    st, got = _call(L.mp_pairhmm_forward_log_prob_noalloc, eng, pp.machines, (seq[:50], seq[:50], None, 1), bp)
    st_ref, ref = _call(L.mp_pairhmm_forward_log_prob, eng, pp.machines, (seq[:50], seq[:50], None, 1), bp)
    assert st == 0 and st_ref == 0 and abs(got - ref) <= 1e-12 * abs(ref)
