"""CPU tests of the shipped library's host side: it loads, exports every
symbol the headers in include/ declare, parses the params JSON like the
reference, and builds the same band / windows / split points as the oracle.
No kernel runs here."""
import ctypes as C
import json
import math
import os
import random
import re

import numpy as np
import pytest

from marginpolish_amd import _lib
from marginpolish_amd import pairhmm as ph
from oracle import oracle as O
from tests.conftest import PARAMS, PARAMS_NO_RLE, ROOT


def header_symbols():
    syms = set()
    for h in sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        syms |= set(re.findall(r"\b(mp_[a-z0-9_]+)\s*\(", txt))
    return sorted(syms)


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(_lib.EXPORTS)
    assert b"gfx950" in L.mp_version()


def test_model_tables_match_oracle_parse():
    pp = ph.PolishParams(PARAMS)
    op = O.PolishParams(PARAMS)
    for strand, om in ((1, op.forward), (0, op.reverse)):
        t = pp.machines.tables(strand)
        keys = ["matchContinue", "matchFromGapX", "matchFromGapY", "gapOpenX", "gapOpenY", "gapExtendX",
                "gapExtendY", "gapSwitchToX", "gapSwitchToY"]
        for i, k in enumerate(keys):
            assert t["trans"][i] == om.trans[k] or (math.isinf(t["trans"][i]) and math.isinf(om.trans[k]))
        assert t["match"] == om.match and t["gapX"] == om.gapX and t["gapY"] == om.gapY
        assert t["rle"] == om.rle
    R = 51
    rng = random.Random(1)
    for _ in range(200):
        b, s, o, u = rng.randint(0, 4), rng.randint(0, 1), rng.randint(0, 50), rng.randint(0, 50)
        bb = 0 if b == 4 else b
        row = bb if s else 3 - bb
        assert pp.machines.repeat_log10(b, s, o, u) == op.repeat[row * R * R + u * R + o]
    p = pp.p
    assert (p.threshold, p.minDiagsBetweenTraceBack, p.traceBackDiagonals, p.diagonalExpansion,
            p.splitMatrixBiggerThanThis, p.dynamicAnchorExpansion) == (0.01, 10000, 40, 10, 250000000, True)
    assert pp.useRepeatCountsInAlignment and pp.maximumRepeatLength == 51
    nr = ph.PolishParams(PARAMS_NO_RLE)
    assert not nr.useRepeatCountsInAlignment


def test_default_machine_matches_oracle():
    m = ph.StateMachines.default_nucleotide()
    om = O.default_nucleotide_model()
    t = m.tables(1)
    assert t["match"] == om.match and t["gapX"] == om.gapX and not t["rle"]


def test_params_errors_are_status_codes(tmp_path):
    js = json.load(open(PARAMS))
    js["polish"]["notAKey"] = 1
    bad = tmp_path / "bad.json"
    bad.write_text(json.dumps(js))
    with pytest.raises(_lib.MPError) as ei:
        ph.PolishParams(str(bad))
    assert ei.value.status == 2 and "notAKey" in str(ei.value)
    js = json.load(open(PARAMS))
    js["polish"]["pairwiseAlignmentParameters"]["diagonalExpansion"] = 3
    bad.write_text(json.dumps(js))
    with pytest.raises(_lib.MPError):
        ph.PolishParams(str(bad))
    js = json.load(open(PARAMS))
    del js["polish"]["hmmForwardStrandReadGivenReference"]
    bad.write_text(json.dumps(js))
    with pytest.raises(_lib.MPError):
        ph.PolishParams(str(bad))
    with pytest.raises(_lib.MPError):
        ph.PolishParams(str(tmp_path / "missing.json"))
    # polish section alone (no top-level 'polish'/'phase') is accepted, as params_readParams2
    only = tmp_path / "only.json"
    only.write_text(json.dumps(json.load(open(PARAMS))["polish"]))
    assert ph.PolishParams(str(only)).p.diagonalExpansion == 10


def _band(anchors, lX, lY, e, dyn):
    a = np.ascontiguousarray(np.asarray(anchors, np.int64).reshape(-1, 3)) if anchors else np.zeros((1, 3), np.int64)
    L = np.zeros(lX + lY + 1, np.int32)
    R = np.zeros(lX + lY + 1, np.int32)
    _lib.check(_lib.lib().mp_band_construct(a.ctypes.data_as(C.POINTER(C.c_int64)), len(anchors or []), lX, lY, e,
                                            int(dyn), L.ctypes.data_as(C.POINTER(C.c_int32)),
                                            R.ctypes.data_as(C.POINTER(C.c_int32))))
    return L, R


def test_band_known_answer_and_vs_oracle():
    L, R = _band([(1, 0, 0), (2, 1, 0), (3, 3, 0)], 6, 5, 2, False)
    want = [(0, 0), (-1, 1), (-2, 2), (-1, 3), (-2, 4), (-1, 3), (-2, 4), (-3, 3), (-2, 2), (-1, 3), (0, 2), (1, 1)]
    assert list(zip(L.tolist(), R.tolist())) == want
    rng = random.Random(9)
    for _ in range(300):
        lX, lY = rng.randint(0, 80), rng.randint(0, 80)
        anchors, x, y = [], -1, -1
        while lX and lY:
            x += rng.randint(1, 12); y += rng.randint(1, 12)
            if x >= lX or y >= lY:
                break
            anchors.append((x, y, 2 * rng.randint(0, 6)))
        e, dyn = 2 * rng.randint(0, 8), rng.random() < 0.5
        L, R = _band(anchors, lX, lY, e, dyn)
        oL, oR = O.band(anchors, lX, lY, e, dyn)
        assert (L == oL).all() and (R == oR).all()


def test_bad_anchors_rejected():
    with pytest.raises(_lib.MPError) as ei:
        _band([(3, 3, 2), (2, 5, 2)], 10, 10, 2, True)
    assert ei.value.status == 3
    with pytest.raises(_lib.MPError):
        _band([(3, 3, 3)], 10, 10, 2, True)
    with pytest.raises(_lib.MPError):
        _band([(10, 3, 2)], 10, 10, 2, True)


def _split(anchors, lX, lY, area, rl=False, rr=False):
    a = np.ascontiguousarray(np.asarray(anchors or [(0, 0, 0)], np.int64).reshape(-1, 3))
    out = np.zeros((64, 4), np.int64)
    k = _lib.lib().mp_split_points(a.ctypes.data_as(C.POINTER(C.c_int64)), len(anchors or []), lX, lY, area,
                                   int(rl), int(rr), out.ctypes.data_as(C.POINTER(C.c_int64)), 64)
    return [tuple(int(v) for v in r) for r in out[:k]]


def test_split_points_known_answer_and_vs_oracle():
    area = 2000 * 2000
    assert _split([], 20000, 25000, area, False, False) == [(0, 0, 2000, 2000), (18000, 23000, 20000, 25000)]
    assert _split([], 20000, 25000, area, True, True) == []
    anchors = [(2000, 2000, 0), (4002, 4001, 0), (5000, 5000, 0), (8000, 6000, 0), (9000, 9000, 0),
               (10000, 14000, 0), (15000, 15000, 0), (16000, 16000, 0)]
    assert _split(anchors, 20000, 25000, area) == O.split_points(anchors, 20000, 25000, area)
    rng = random.Random(4)
    for _ in range(200):
        lX, lY = rng.randint(1, 3000), rng.randint(1, 3000)
        anchors, x, y = [], -1, -1
        while True:
            x += rng.randint(1, 700); y += rng.randint(1, 700)
            if x >= lX or y >= lY:
                break
            anchors.append((x, y, 0))
        for rl in (False, True):
            for rr in (False, True):
                assert _split(anchors, lX, lY, 90000, rl, rr) == O.split_points(anchors, lX, lY, 90000, rl, rr)


def test_windows_follow_reference_rule():
    """Traceback windows: the oracle's driver fires at the same diagonals."""
    rng = random.Random(12)
    for _ in range(60):
        lX = rng.randint(50, 400)
        anchors = [(i, i, 2 * rng.randint(0, 4)) for i in range(0, lX, rng.randint(1, 9))]
        L, R = _band(anchors, lX, lX, 0, True)
        bp = ph.PairwiseAlignmentParameters(minDiagsBetweenTraceBack=rng.randint(12, 120),
                                            traceBackDiagonals=rng.randint(1, 10), diagonalExpansion=4,
                                            dynamicAnchorExpansion=True)
        out = np.zeros((256, 3), np.int32)
        bpc = bp.c()
        k = _lib.lib().mp_traceback_windows(L.ctypes.data_as(C.POINTER(C.c_int32)),
                                            R.ctypes.data_as(C.POINTER(C.c_int32)), 2 * lX, C.byref(bpc),
                                            out.ctypes.data_as(C.POINTER(C.c_int32)), 256)
        wins = out[:k]
        # restate the rule of getPosteriorProbsWithBanding directly
        want, lo = [], 0
        for d in range(1, 2 * lX + 1):
            atEnd = d == 2 * lX
            tbp = d >= lo + bp.minDiagsBetweenTraceBack and (R[d] - L[d]) // 2 + 1 <= 2 * bp.diagonalExpansion + 1
            if atEnd or tbp:
                frm = d if atEnd else d - (bp.traceBackDiagonals + 1)
                want.append((d, lo, frm))
                lo = frm
            if atEnd:
                break
        assert [tuple(w) for w in wins.tolist()] == want


def test_engine_without_device_is_an_error():
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(_lib.MPError) as ei:
        ph.Engine(0)
    assert ei.value.status == 7


def test_fit_sampler_and_gradient_plumbing():
    """The pair-HMM sampler of the fitting harness follows its transition
    matrix (path frequencies), its anchors are valid, and the autograd
    function returns the evaluator's counts as the gradient."""
    import torch
    from marginpolish_amd import fit
    T = np.array([[0.90, 0.05, 0.05], [0.70, 0.25, 0.05], [0.70, 0.05, 0.25]])
    pairs = fit.simulate_pairs(300, 400, T, np.full(16, 1 / 16), np.full(4, 0.25), np.full(4, 0.25), 10, seed=1)
    n = np.zeros((3, 3))
    for sx, sy, an, strand in pairs:
        assert strand == 1 and an.shape[1] == 3 and np.all(an[:, 2] == 10)
        assert np.all(np.diff(an[:, 0]) > 0) and np.all(np.diff(an[:, 1]) > 0)
        assert an[:, 0].max() < sx.size and an[:, 1].max() < sy.size and sx.size == 400
        # recover the state path from the anchors: between consecutive matches,
        # dx - 1 gap-X and dy - 1 gap-Y states (in some order) are hidden, so
        # count only match-to-match steps, which identify T[M, M]
        d = np.diff(an[:, :2], axis=0)
        n[0, 0] += np.sum((d[:, 0] == 1) & (d[:, 1] == 1))
        n[0, 1] += np.sum((d[:, 0] > 1) | (d[:, 1] > 1))
    assert abs(n[0, 0] / (n[0, 0] + n[0, 1]) - T[0, 0]) < 0.01
    assert fit.matrix_from_named(fit.named_from_matrix(np.log(T))).round(12).tolist() == np.log(T).round(12).tolist()

    class Quadratic:
        pairs = [0]

        def evaluate(self, logT):
            c = np.arange(9.0).reshape(3, 3)
            return float((c * logT).sum()), c

    logT = torch.zeros(3, 3, dtype=torch.float64, requires_grad=True)
    fit.log_likelihood(logT, Quadratic()).backward()
    assert np.array_equal(logT.grad.numpy(), np.arange(9.0).reshape(3, 3))


def test_device_symbol_encoding_matches_host_encoding(tmp_path):
    """The batch's symbol pool goes to the device as 16-bit raw entries and is
    encoded there: raw_sym then enc_raw equals enc_x / enc_y for every symbol,
    side and strand (compiled for the host with hipcc)."""
    import shutil
    import subprocess
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    here = os.path.dirname(os.path.abspath(__file__))
    root = os.path.dirname(here)
    exe = str(tmp_path / "symbol_encoding_check")
    subprocess.check_call([hipcc, "-O2", "-std=c++17", "-I", os.path.join(root, "include"), "-I",
                           os.path.join(root, "marginpolish_amd", "csrc"), "-o", exe,
                           os.path.join(here, "symbol_encoding_check.cpp")], stderr=subprocess.DEVNULL)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
    assert "mismatches 0" in out.stdout


def test_c_caller_compiles_against_the_headers(tmp_path):
    """A C program compiled with gcc against include/*.h and linked with the
    library: the ABI structs' sizes and offsets equal the ctypes mirrors, and
    its host calls give the Python bindings' results."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    here = os.path.dirname(os.path.abspath(__file__))
    libdir = os.path.join(ROOT, "marginpolish_amd")
    exe = str(tmp_path / "c_abi_check")
    subprocess.check_call([gcc, "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-o", exe,
                           os.path.join(here, "c_abi_check.c"), "-L", libdir, "-lmp_pairhmm",
                           "-Wl,-rpath," + libdir])
    out = subprocess.run([exe, PARAMS], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    rows = {ln.split()[0]: ln.split()[1:] for ln in out.stdout.splitlines()}

    def layout(cls):
        return [C.sizeof(cls)] + [getattr(cls, f).offset for f, _ in cls._fields_]
    assert [int(v) for v in rows["mp_pair"]] == layout(_lib.Pair)
    assert [int(v) for v in rows["mp_band_params"]] == layout(_lib.BandParams)
    assert [int(v) for v in rows["mp_aligned_pair"]] == layout(_lib.AlignedPair)
    assert [int(v) for v in rows["mp_rle_string"]] == layout(_lib.RleString)
    assert int(rows["mp_poa_sizes"][0]) == C.sizeof(_lib.PoaSizes)
    p = ph.PolishParams(PARAMS).p
    assert [float(rows["params"][0])] + [int(v) for v in rows["params"][1:]] == [
        p.threshold, p.minDiagsBetweenTraceBack, p.traceBackDiagonals, p.diagonalExpansion,
        p.splitMatrixBiggerThanThis, int(p.dynamicAnchorExpansion)]
    L, R = _band([(1, 0, 0), (2, 1, 0), (3, 3, 0)], 6, 5, 2, False)
    assert rows["band"] == ["%d:%d" % (l, r) for l, r in zip(L, R)]
    sp = O.split_points([(10, 10, 0), (11, 11, 0), (5000, 5000, 0), (5001, 5001, 0)], 6000, 6000, 1000000)
    assert int(rows["splits"][0]) == len(sp) >= 2
    assert [int(v) for v in rows["splits"][1:]] == [int(v) for t in sp for v in t]
