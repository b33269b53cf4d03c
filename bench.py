"""Benchmark: batched pair-HMM forward-backward + posteriors on one MI355X.

Default workload (the full-polish-loop configuration, one realignment
iteration): a seeded 5 Mb reference (GC 50%, runs ~ Geometric(0.5), 1% in
2-6-mer tandem repeats), 50x log-normal reads (mean 20 kb, sd 15 kb, max
100 kb) at 8% error, 50 chunks of 100 kb with 5 kb boundaries, RLE symbols,
anchors = true RLE match pairs with expansion 10, allParams.np.json 3-state
RLE model.  One step = one POA realignment iteration's device work: every
read/chunk pair (~16k) through forward, backward and posterior extraction
(the persistent kernels), then the reference's three lists per pair and the
FP64 log-likelihoods assembled on the device; inputs resident in HBM.
`--workload config1` runs the 400 kb / 40-chunk batch instead.

Prints one JSON line (rank 0, N=1)."""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pair-HMM GCUPS + reads/s: batched FP32 vs plain, posterior err <=1e-4"
WORKLOADS = {
    "config1": "config1: %d chunks x 10 kb (+-500 bp), 400 kb ref, 60x reads mean 8 kb, 9%% error, RLE, anchor "
               "expansion 10, allParams.np.json, one POA realign iteration",
    "config2": "config2: %d chunks x 100 kb (+-5 kb), 5 Mb ref, 50x reads mean 20 kb (max 100 kb), 8%% error, RLE, "
               "anchor expansion 10, allParams.np.json, one POA realign iteration",
}
FLOPS_PER_CELL = 42          # forward 18 + backward 18 + posterior 6 (FMA = 2)
FLOPS_PER_CELL_FWD = 18      # forward only
PEAK_FP32_TFLOPS = 157.3     # MI355X vector FP32
PARAMS = os.path.join(ROOT, "tests", "golden", "params", "allParams.np.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _rows(arr):
    """(weight, x, y) list of one pair (numpy structured or int rows) -> {(x, y): weight}."""
    if getattr(arr, "dtype", None) is not None and arr.dtype.names:
        return {(int(x), int(y)): int(w) for w, x, y in zip(arr["weight"], arr["x"], arr["y"])}
    return {(int(r[1]), int(r[2])): int(r[0]) for r in arr}


def _posterior_diff(a, b, threshold, tol):
    """Max |p_a - p_b| over the (x, y) both lists hold, or 1.0 when an entry
    held by one side only lies more than tol above the threshold (the two
    sides may decide p >= threshold differently only near it)."""
    da, db = _rows(a), _rows(b)
    worst = 0.0
    for k in set(da) & set(db):
        worst = max(worst, abs(da[k] - db[k]) / 1e7)
    edge = (threshold + tol) * 1e7 + 1
    for d, other in ((da, db), (db, da)):
        for k in set(d) - set(other):
            if d[k] > edge:
                return 1.0
    return worst


def _host_threads():
    """Host threads for the CPU baseline: the cores this process may use
    (affinity), capped by OMP_NUM_THREADS when the machine sets it (the GPU
    box allots 16 per GPU)."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    cap = int(omp) if omp.isdigit() and int(omp) > 0 else aff
    return max(1, min(aff, cap)), aff, os.cpu_count() or 1


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--plain-sample", type=int, default=64)
    ap.add_argument("--cpu-pairs", type=int, default=2000)
    ap.add_argument("--err-pairs", type=int, default=48)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-poa", action="store_true", help="skip the POA-augmentation timing")
    ap.add_argument("--exact-chain", action="store_true",
                    help="run each sub-problem's traceback windows on one forward chain (no window jobs)")
    ap.add_argument("--workload", choices=("config1", "config2", "config4", "bubbles", "polish"), default="config2",
                    help="config2 (the headline: 5 Mb reference, 50x 20 kb reads, 100 kb chunks), config1 (400 kb, "
                         "40 chunks of 10 kb), config4 (five-state long-band stress: FP32 vs FP64 GPU vs CPU FP64 at "
                         "one --band) or bubbles (bubble-graph per-read x per-allele forward likelihoods) or polish "
                         "(a simulated draft polished end to end: realign, consensus rounds, run lengths, merge)")
    ap.add_argument("--band", type=int, default=128, help="config4: anchor expansion W (128, 256, 512, 1024)")
    ap.add_argument("--pairs", type=int, default=256, help="config4: number of 100 kb pairs")
    ap.add_argument("--max-chunks", type=int, default=None, help="limit the number of chunks (config2)")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a known-size kernel that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE counters to bytes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path here (chunks would shard across "
              "GPUs with no exchange between them). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from marginpolish_amd import crop, pairhmm, sim

    if args.workload == "config4":
        return run_config4(args)
    if args.workload == "bubbles":
        return run_bubbles(args)
    if args.workload == "polish":
        return run_polish(args)

    if args.copy_calib_mib > 0:
        import torch
        n = args.copy_calib_mib << 18  # float32 elements
        a = torch.ones(n, device="cuda")
        c = torch.empty_like(a)
        c.copy_(a)
        torch.cuda.synchronize()
        log("copy calibration: %d MiB read + %d MiB written" % (args.copy_calib_mib, args.copy_calib_mib))
        del a, c

    t0 = time.time()
    w = sim.config1() if args.workload == "config1" else sim.config2(max_chunks=args.max_chunks)
    pairs = [(p.ref_sym, p.read_sym, p.anchors, p.strand) for p in w.pairs]
    pair_array = pairhmm.PairArray(pairs)  # the mp_pair array, laid out once as a C caller holds it
    log("workload: %d pairs built in %.1f s" % (len(pairs), time.time() - t0))
    pp = pairhmm.PolishParams(PARAMS)
    eng = pairhmm.default_engine()
    exact = args.exact_chain

    # reads/s: one POA iteration end to end (host prep + upload + device work + the three lists and their
    # per-pair offsets in host memory, as the C ABI hands them over), timed on the second iteration of the
    # engine (the first also pays one-time allocations: pinned staging, code load)
    iters = []
    for it in range(2):
        t1 = time.time()
        b = pairhmm.Batch(pp.machines, pair_array, pp.p, crop_reference=True, engine=eng, exact_chain=exact)
        t_prep = time.time() - t1
        b.run()
        b.fetch_arrays()
        iters.append((t_prep, time.time() - t1))
        if it == 0:
            b.close()
    res = b.fetch()  # per-pair views, for the posterior error sample
    t_prep, t_iter = iters[1]
    reads_per_s = len(pairs) / t_iter
    cells = b.cells
    log("iteration: first %.3f s; second: prep %.3f s, total %.3f s, %d cells" % (iters[0][1], t_prep, t_iter, cells))

    poa = {}
    if not args.no_poa:
        poa = _poa_iteration(w, pairs, pair_array, pp, eng, exact)

    # hardware tiling picked by the autotuner on this batch (launch geometry only; results unchanged)
    tuned = b.autotune(3)
    log("autotune: %d-thread narrow workgroups, best %.3f ms" % (tuned[0], tuned[2]))
    for _ in range(args.warmup):
        b.run()
    kms, lms = [], []
    t2 = time.perf_counter()
    for _ in range(args.steps):
        b.run()
        kms.append(b.kernel_ms)
        lms.append(b.list_ms)
    wall = time.perf_counter() - t2
    ms_per_step = wall * 1e3 / args.steps
    kernel_ms = float(np.mean(kms))
    list_ms = float(np.mean(lms))
    gcups = cells / (ms_per_step * 1e-3) / 1e9
    kernel_gcups = cells / (kernel_ms * 1e-3) / 1e9
    total_ll = b.total_log_prob()
    njobs, rec_cap, recs = b.stats()
    log("batched: %.3f ms/step wall, %.3f ms kernel, %.3f ms list assembly, %.1f GCUPS"
        % (ms_per_step, kernel_ms, list_ms, gcups))

    # plain per-read kernel (one launch per pair) on a sample of pairs
    order = np.argsort([-len(p[0]) for p in pairs])
    samp = [pairs[i] for i in order[:: max(1, len(pairs) // args.plain_sample)][: args.plain_sample]]
    bp = pairhmm.Batch(pp.machines, samp, pp.p, crop_reference=True, engine=eng, exact_chain=exact)
    bp.run_plain()
    plain_gcups = bp.cells / (bp.kernel_ms * 1e-3) / 1e9
    log("plain: %d pairs, %.3f ms, %.2f GCUPS" % (len(samp), bp.kernel_ms, plain_gcups))
    bp.close()

    out = {
        "metric": METRIC, "value": round(gcups, 3), "unit": "GCUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": WORKLOADS[args.workload] % w.chunks,
                   "pairs": len(pairs), "cells": int(cells), "band_half_width": 10,
                   "traceback_windows": "exact chain" if exact else "window jobs (512 anti-diagonal warm-up)"},
        "kernel_ms": round(kernel_ms, 4), "kernel_gcups": round(kernel_gcups, 3),
        "list_assembly_ms": round(list_ms, 4), "device_jobs": njobs, "posterior_records": recs,
        "record_capacity": rec_cap,
        "reads_per_s": round(reads_per_s, 1), "iteration_s": round(t_iter, 4), "iteration_prep_s": round(t_prep, 4),
        "first_iteration_s": round(iters[0][1], 4), "total_log_likelihood_fp64": total_ll,
        "plain_gcups": round(plain_gcups, 4), "batched_over_plain": round(kernel_gcups / plain_gcups, 2),
        "plain_sample_pairs": len(samp), "autotuned_narrow_threads": tuned[0],
    }
    out.update(poa)
    achieved = FLOPS_PER_CELL * cells / (kernel_ms * 1e-3) / 1e12
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_pairhmm_split_kernel_%s.json" % args.workload)
    if os.path.exists(tf) and not exact:  # measured on this workload's launch (window jobs)
        traffic = json.load(open(tf)).get("hbm_bytes_per_launch")
    out["roofline"] = {"bound": "valu", "achieved": round(achieved, 4), "peak": PEAK_FP32_TFLOPS,
                       "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 5), "traffic": traffic}

    if not args.no_cpu:
        out.update(_cpu_leg(args, pairs, res, pp))
    print(json.dumps(out), flush=True)


def _poa_iteration(w, pairs, pair_array, pp, eng, exact):
    """The same iteration through the POA augmentation: every chunk's graph
    built from the batch's lists inside the library (no list handoff to
    Python); timed on the second run, plus the augment alone on one thread
    and the run-length estimation of every graph."""
    import numpy as np
    from marginpolish_amd import pairhmm
    from marginpolish_amd import poa as mpoa
    chunk_ids = sorted(set(p.chunk for p in w.pairs))
    first_pair = {}
    for p in w.pairs:
        first_pair.setdefault(p.chunk, p)
    refs = [mpoa.rle_from_symbols(first_pair[c].ref_sym, first_pair[c].ref_runs) for c in chunk_ids]
    cidx = {c: k for k, c in enumerate(chunk_ids)}
    pair_poa = np.array([cidx[p.chunk] for p in w.pairs], dtype=np.int64)
    read_no = np.zeros(len(pairs), dtype=np.int64)
    seen = {}
    for i, p in enumerate(w.pairs):
        read_no[i] = seen.get(p.chunk, 0)
        seen[p.chunk] = read_no[i] + 1
    rle_reads = [mpoa.rle_from_symbols(p.read_sym, p.read_runs) for p in w.pairs]
    packed_reads = mpoa.PackedReads(rle_reads)  # the reads' RleStrings, laid out once
    poa_iters = []
    for it in range(2):
        t1 = time.time()
        bb = pairhmm.Batch(pp.machines, pair_array, pp.p, crop_reference=True, engine=eng, exact_chain=exact)
        tc = time.time()
        bb.run()
        ta = time.time()
        log("POA iteration %d: batch create %.3f s, run %.3f s" % (it, tc - t1, ta - tc))
        graphs = [mpoa.Poa(r) for r in refs]
        mpoa.augment_batch(bb, graphs, pair_poa, packed_reads, read_no)
        poa_iters.append((time.time() - ta, time.time() - t1))
        if it == 1:
            tb = time.time()
            g1 = [mpoa.Poa(r) for r in refs]
            mpoa.augment_batch(bb, g1, pair_poa, packed_reads, read_no, threads=1)
            aug1 = time.time() - tb
            gsz = [g.sizes() for g in graphs]
            del g1
            # the polished run lengths of every graph (poa_estimateRepeatCountsUsingBayesianModel)
            by_chunk = [[] for _ in refs]
            for i in range(len(pairs)):
                by_chunk[pair_poa[i]].append(i)
            chunk_reads = [mpoa.PackedReads([rle_reads[i] for i in ids]) for ids in by_chunk]
            chunk_strands = [[w.pairs[i].strand for i in ids] for ids in by_chunk]
            tr = time.time()
            mpoa.estimate_repeat_counts_many(graphs, pp.machines, chunk_reads, chunk_strands)
            t_rc = time.time() - tr
        bb.close()
        del graphs
    t_aug, t_poa_iter = poa_iters[1]
    log("POA iteration: %.3f s (augment %.3f s on all cores, %.3f s on one thread; %d graphs, %d inserts, "
        "%d deletes); repeat-count estimation %.3f s" % (t_poa_iter, t_aug, aug1, len(refs),
                                                         sum(s["inserts"] for s in gsz),
                                                         sum(s["deletes"] for s in gsz), t_rc))
    return {"poa_iteration_s": round(t_poa_iter, 4), "poa_reads_per_s": round(len(pairs) / t_poa_iter, 1),
            "augment_s": round(t_aug, 4), "augment_1thread_s": round(aug1, 4),
            "repeat_count_estimation_s": round(t_rc, 4)}


def _cpu_leg(args, pairs, res, pp):
    """CPU baseline (the oracle's C restatement of the reference algorithm,
    FP64 log space with the polynomial logAdd, built -march=native on this
    host, OpenMP over pairs) on a bounded sample, and the GPU's posterior
    error and log-likelihood drift against the exact FP64 oracle."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np
    from marginpolish_amd import crop
    from oracle import oracle as O
    native = O.use_native()
    op = O.PolishParams(PARAMS)
    nthr, aff, nproc = _host_threads()
    rng = np.random.default_rng(0)
    pick = rng.choice(len(pairs), size=min(args.cpu_pairs, len(pairs)), replace=False)
    cpu_jobs = []
    for i in pick:
        sx, sy, an, _ = crop.crop_pair(*pairs[i][:3])
        cpu_jobs.append((sx, sy, an, pairs[i][3]))
    t3 = time.perf_counter()
    ccells, _ = O.batch_cells((op.reverse, op.forward), cpu_jobs, op.p, O.LOGADD_POLY, nthr)
    tc = time.perf_counter() - t3
    one = cpu_jobs[: max(1, len(cpu_jobs) // 16)]
    t4 = time.perf_counter()
    c1, _ = O.batch_cells((op.reverse, op.forward), one, op.p, O.LOGADD_POLY, 1)
    t1c = time.perf_counter() - t4
    # the exact-arithmetic oracle (O_exact) on the single-thread sample, for reference
    t5 = time.perf_counter()
    c2, _ = O.batch_cells((op.reverse, op.forward), one, op.p, O.LOGADD_EXACT, 1)
    t2c = time.perf_counter() - t5
    log("cpu: %d pairs on %d threads %.2f s; %d pairs on 1 thread %.2f s" % (len(cpu_jobs), nthr, tc, len(one), t1c))
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            cpu_model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), "")
    except OSError:
        pass
    out = {"cpu_baseline": {"value": round(ccells / tc / 1e9, 5), "unit": "GCUPS", "cores": nthr, "kind": "port",
                            "sample": "%d of %d %s pairs (%d cells), reference algorithm restated in C (FP64 log "
                                      "space, polynomial logAdd) built -march=native on this host, OpenMP over pairs"
                                      % (len(cpu_jobs), len(pairs), args.workload, ccells),
                            "reads_per_s": round(len(cpu_jobs) / tc, 2),
                            "single_core_gcups": round(c1 / t1c / 1e9, 5),
                            "single_core_reads_per_s": round(len(one) / t1c, 3),
                            "o_exact_single_core_gcups": round(c2 / t2c / 1e9, 5),
                            "cpu_model": cpu_model, "nproc": nproc, "affinity_cores": aff,
                            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
                            "library": os.path.relpath(native, ROOT)}}

    # max-abs posterior error and log-likelihood drift vs the exact FP64 oracle, and the reference
    # arithmetic (polynomial logAdd) against exact log-sum-exp, on a sample (pairs in parallel)
    def one_pair(i):
        sx, sy, an, off = crop.crop_pair(*pairs[i][:3])
        m = op.forward if pairs[i][3] else op.reverse
        ex = O.posteriors(m, sx, sy, an, op.p, mode=O.LOGADD_EXACT)
        fa = O.posteriors(m, sx, sy, an, op.p, mode=O.LOGADD_POLY)
        gm, gx, gy, glp = res[i]
        err = 0.0
        for g, o in ((gm, ex[0]), (gx, ex[1]), (gy, ex[2])):
            o = o.copy()
            o[:, 1] += off
            err = max(err, _posterior_diff(g, o, op.p.threshold, 1e-4))
        fe = max(_posterior_diff(a, b_, op.p.threshold, 1.0) for a, b_ in zip(fa[:3], ex[:3]))
        return err, abs(glp - ex[3]), fe

    t6 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=nthr) as pool:
        got = list(pool.map(one_pair, pick[: args.err_pairs]))
    log("posterior error sample: %d pairs in %.1f s" % (len(got), time.perf_counter() - t6))
    out["max_abs_posterior_err"] = max(g[0] for g in got)
    out["max_log_likelihood_drift"] = max(g[1] for g in got)
    out["posterior_err_sample_pairs"] = len(got)
    out["o_faithful_vs_o_exact_max_abs"] = max(g[2] for g in got)
    return out


FLOPS_PER_CELL_5 = 120  # five states: forward 55 (5 emission muls, 25 FMA sends), backward 55, posterior 10


def _keyed(res, li):
    """All pairs' tuples of list li as (sorted keys pair<<42 | (x+1)<<21 | (y+1), weights); x or y is -1 for
    gap tuples before the first base of the other sequence."""
    import numpy as np
    keys, ws = [], []
    for i, r in enumerate(res):
        a = r[li]
        keys.append((np.int64(i) << 42) | ((a["x"].astype(np.int64) + 1) << 21) | (a["y"].astype(np.int64) + 1))
        ws.append(a["weight"].astype(np.int64))
    k = np.concatenate(keys) if keys else np.zeros(0, np.int64)
    w = np.concatenate(ws) if ws else np.zeros(0, np.int64)
    o = np.argsort(k, kind="stable")
    return k[o], w[o]


def _list_err(ra, rb, thr, tol):
    """Max |p_a - p_b| over tuples both hold, the largest weight held by one
    side only (it must stay within tol of the threshold), and the number of
    repeated (pair, x, y) keys (0 in a correct list)."""
    import numpy as np
    worst, edge, dups = 0.0, 0.0, 0
    for li in range(3):
        ka, wa = _keyed(ra, li)
        kb, wb = _keyed(rb, li)
        dups += int(np.sum(ka[1:] == ka[:-1]) + np.sum(kb[1:] == kb[:-1]))
        pa = np.minimum(np.searchsorted(kb, ka), max(kb.size - 1, 0))
        inA = (kb.size > 0) & (kb[pa] == ka) if kb.size else np.zeros(ka.size, bool)
        pb = np.minimum(np.searchsorted(ka, kb), max(ka.size - 1, 0))
        inB = (ka.size > 0) & (ka[pb] == kb) if ka.size else np.zeros(kb.size, bool)
        if inA.any():
            worst = max(worst, float(np.abs(wa[inA] - wb[pa[inA]]).max()) / 1e7)
        for w, only in ((wa, ~inA), (wb, ~inB)):
            if only.any():
                edge = max(edge, float(w[only].max()) / 1e7)
    return worst, edge, dups


def run_config4(args):
    """Config 4: the five-state machine on 100 kb pairs at one band W, FP32
    (per-cell exponents) vs the FP64 GPU path, and both vs the CPU FP64
    oracle on a bounded sample (the first ~6k RLE positions of one pair)."""
    import numpy as np
    from marginpolish_amd import pairhmm, sim
    t0 = time.time()
    w = sim.config4(n_pairs=args.pairs, W=args.band)
    pairs = [(p.ref_sym, p.read_sym, p.anchors, p.strand) for p in w.pairs]
    log("config4: %d pairs, W=%d, built in %.1f s" % (len(pairs), args.band, time.time() - t0))
    pp = pairhmm.PolishParams(PARAMS)
    mach = pp.machines
    mach.add_long_gaps()
    bp = pairhmm.PairwiseAlignmentParameters(threshold=0.01, minDiagsBetweenTraceBack=10000, traceBackDiagonals=40,
                                             diagonalExpansion=args.band, splitMatrixBiggerThanThis=250_000_000,
                                             dynamicAnchorExpansion=True)
    eng = pairhmm.default_engine()
    legs = {}
    for tag, fp64 in (("fp32", False), ("fp64", True)):
        t1 = time.time()
        b = pairhmm.Batch(mach, pairs, bp, engine=eng, fp64=fp64)
        b.run()
        res = b.fetch()
        t_first = time.time() - t1
        kms = []
        for _ in range(args.steps):
            b.run()
            kms.append(b.kernel_ms)
        kernel_ms = float(np.mean(kms))
        legs[tag] = {"kernel_ms": kernel_ms, "gcups": b.cells / (kernel_ms * 1e-3) / 1e9, "cells": int(b.cells),
                     "first_iteration_s": t_first, "res": res}
        log("%s: %.1f ms kernel, %.2f GCUPS, first iteration %.1f s" % (tag, kernel_ms, legs[tag]["gcups"], t_first))
        b.close()
        del b
    r32, r64 = legs["fp32"]["res"], legs["fp64"]["res"]
    err, edge, dups = _list_err(r32, r64, 0.01, 1e-4)
    lp32 = np.array([r[3] for r in r32])
    lp64 = np.array([r[3] for r in r64])
    drift = np.abs(lp32 - lp64)
    cells = legs["fp32"]["cells"]
    kernel_ms = legs["fp32"]["kernel_ms"]
    achieved = FLOPS_PER_CELL_5 * cells / (kernel_ms * 1e-3) / 1e12
    out = {
        "metric": "five-state pair-HMM GCUPS (long-band stress), FP32 per-cell exponents vs FP64 GPU vs CPU FP64",
        "value": round(legs["fp32"]["gcups"], 3), "unit": "GCUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": 1, "ms_per_step": round(kernel_ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "config4: %d pairs of a 100 kb read vs its reference (RLE ~65 kb), 12%% error, "
                               "structural indels 50-500 bp over 0.5%%, five-state machine (long gaps open 0.0018, "
                               "extend 0.9965), anchor expansion W=%d, windows every 10000 anti-diagonals"
                               % (len(pairs), args.band), "pairs": len(pairs), "cells": cells,
                   "band_half_width": args.band},
        "fp64_gcups": round(legs["fp64"]["gcups"], 3), "fp64_kernel_ms": round(legs["fp64"]["kernel_ms"], 3),
        "fp32_vs_fp64_gpu": {"max_abs_posterior_err": err, "max_one_sided_posterior": edge, "repeated_keys": dups,
                             "max_abs_ll_drift": float(drift.max()),
                             "max_rel_ll_drift": float((drift / np.abs(lp64)).max())},
        "roofline": {"bound": "valu", "achieved": round(achieved, 4), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 5), "traffic": None,
                     "flops_per_cell": FLOPS_PER_CELL_5},
    }
    if not args.no_cpu:
        from oracle import oracle as O
        op = O.PolishParams(PARAMS)
        T5 = mach.transitions5()
        sx, sy, an, strand = pairs[0]
        n = 6000
        an = an[(an[:, 0] < n) & (an[:, 1] < n)]
        cut = int(an[-1, 0]) + 1, int(an[-1, 1]) + 1
        sx, sy, an = sx[: cut[0]], sy[: cut[1]], an[:-1]
        base = op.forward if strand else op.reverse
        om5 = O.Model(base.trans, base.match, base.gapX, base.gapY, base.rle, base.strand, base.repeat, trans5=T5)
        opar = O.Params(threshold=0.01, minDiagsBetweenTraceBack=10000, traceBackDiagonals=40,
                        diagonalExpansion=args.band, splitMatrixBiggerThanThis=250_000_000,
                        dynamicAnchorExpansion=True)
        t3 = time.perf_counter()
        o = O.posteriors(om5, sx, sy, an, opar, mode=O.LOGADD_EXACT)
        tc = time.perf_counter() - t3
        cpu = {"value": round(o[4] / tc / 1e9, 6), "unit": "GCUPS", "cores": 1, "kind": "port",
               "sample": "first %d x %d RLE positions of pair 0 (%d cells), five-state O_exact (FP64 log space, "
                         "exact log-sum-exp), one thread" % (cut[0], cut[1], o[4])}
        for tag, fp64 in (("fp32", False), ("fp64", True)):
            b = pairhmm.Batch(mach, [(sx, sy, an, strand)], bp, engine=eng, fp64=fp64)
            b.run()
            g = b.fetch()
            orows = [{"weight": l[:, 0], "x": l[:, 1], "y": l[:, 2]} for l in o[:3]]
            e, ed, du = _list_err(g, [tuple(orows) + (o[3],)], 0.01, 1e-4)
            cpu[tag + "_max_abs_posterior_err"] = e
            cpu[tag + "_max_one_sided_posterior"] = ed
            cpu[tag + "_repeated_keys"] = du
            cpu[tag + "_ll_drift"] = abs(g[0][3] - o[3])
        out["cpu_baseline"] = cpu
        log("cpu sample: %.1f s, %s" % (tc, json.dumps(cpu)))
    print(json.dumps(out), flush=True)


def run_bubbles(args):
    """Bubble-graph likelihoods (bubbleGraph.c:895-928): 4000 bubbles x 60
    read substrings x 2-4 alleles, every (allele, distinct read) pair's
    forward log-likelihood in one forward-only launch.  value = allele-read
    likelihoods per second of the whole call (host string handling + upload
    + kernel + download); the kernel line is the same pairs as a resident
    batch."""
    import numpy as np
    from marginpolish_amd import pairhmm, sim
    t0 = time.time()
    bub = sim.bubbles(n_bubbles=4000, depth=60, seed=5)
    n_sup = sum(len(a) * len(r) for a, r in bub)
    log("bubbles: %d bubbles, %d supports, built in %.1f s" % (len(bub), n_sup, time.time() - t0))
    pp = pairhmm.PolishParams(PARAMS)
    eng = pairhmm.default_engine()
    pk = pairhmm.pack_bubbles(bub)  # the strings as a C caller holds them (timed region starts after)
    calls = []
    for _ in range(1 + max(1, args.warmup)):
        t1 = time.perf_counter()
        flat, npairs = pairhmm.bubble_allele_supports_packed(pp.machines, pp.p, pk, engine=eng)
        calls.append(time.perf_counter() - t1)
    t_call = min(calls[1:])
    got, at = [], 0
    for alleles, reads in bub:
        got.append(flat[at:at + len(alleles) * len(reads)].reshape(len(alleles), len(reads)))
        at += len(alleles) * len(reads)
    # the same (allele, distinct read) pairs as a resident batch, for the kernel's own rate
    pairs = []
    for alleles, reads in bub:
        seen = set()
        asym = [pairhmm.rleString_constructSymbolString(a) for a in alleles]
        for r, strand in reads:
            key = (bytes(r[0]), tuple(int(x) for x in r[1]))
            if key in seen:
                continue
            seen.add(key)
            rs = pairhmm.rleString_constructSymbolString(r)
            pairs.extend((a, rs, None, strand) for a in asym)
    assert len(pairs) == npairs
    t4 = time.perf_counter()
    b = pairhmm.Batch(pp.machines, pairs, pp.p, mode=2, engine=eng)
    t_create = time.perf_counter() - t4
    b.run()
    kms = []
    t2 = time.perf_counter()
    for _ in range(args.steps):
        b.run()
        kms.append(b.kernel_ms)
    ms_per_step = (time.perf_counter() - t2) * 1e3 / args.steps
    kernel_ms = float(np.mean(kms))
    cells = b.cells
    log("bubbles: call %.3f s (%d forward pairs), batch create from Python %.3f s, kernel %.3f ms, %.1f GCUPS"
        % (t_call, npairs, t_create, kernel_ms, cells / kernel_ms / 1e6))
    achieved = FLOPS_PER_CELL_FWD * cells / (kernel_ms * 1e-3) / 1e12
    out = {
        "metric": "bubble-graph allele-read forward log-likelihoods per second", "value": round(n_sup / t_call, 1),
        "unit": "likelihoods/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_call * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "bubbles: 4000 bubbles of 8-40 bp (+6 bp flanks), 2-4 alleles, 60 read substrings at "
                               "9% error (20% repeated), allParams.np.json, forward only, no anchors",
                   "bubbles": len(bub), "supports": n_sup, "forward_pairs": npairs, "cells": int(cells)},
        "kernel_ms": round(kernel_ms, 4), "kernel_gcups": round(cells / (kernel_ms * 1e-3) / 1e9, 3),
        "batch_ms_per_step": round(ms_per_step, 4), "batch_create_s": round(t_create, 4),
        "roofline": {"bound": "valu", "achieved": round(achieved, 4), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 5), "traffic": None,
                     "flops_per_cell": FLOPS_PER_CELL_FWD},
    }
    # parity on a sample against the FP64 oracle (computeForwardProbability restated)
    from oracle import oracle as O
    op = O.PolishParams(PARAMS)
    worst = 0.0
    for (alleles, reads), g in list(zip(bub, got))[:40]:
        first = {}
        for k, (r, strand) in enumerate(reads):
            key = (bytes(r[0]), tuple(int(x) for x in r[1]))
            if key in first:
                continue
            first[key] = k
            m = op.forward if strand else op.reverse
            rs = pairhmm.rleString_constructSymbolString(r)
            for j, a in enumerate(alleles):
                want = O.forward_prob(m, pairhmm.rleString_constructSymbolString(a), rs, [], op.p)
                worst = max(worst, abs(float(g[j, k]) - want) / max(1.0, abs(want)))
    out["max_rel_ll_err_vs_fp64_oracle"] = worst
    if not args.no_cpu:
        samp = pairs[:: max(1, len(pairs) // 20000)][:20000]
        mods = (op.reverse, op.forward)
        t3 = time.perf_counter()
        c_cells = 0
        for a, rs, _, strand in samp:
            O.forward_prob(mods[strand], a, rs, [], op.p, mode=O.LOGADD_POLY)
        tc = time.perf_counter() - t3
        bs = pairhmm.Batch(pp.machines, samp, pp.p, mode=2, engine=eng)
        c_cells = bs.cells
        bs.close()
        out["cpu_baseline"] = {"value": round(len(samp) * (n_sup / npairs) / tc, 1), "unit": "likelihoods/s",
                               "cores": 1, "kind": "port",
                               "sample": "%d of %d forward pairs (%d cells), computeForwardProbability restated in C "
                                         "(FP64 log space, polynomial logAdd), one thread, called per pair"
                                         % (len(samp), npairs, c_cells),
                               "gcups": round(c_cells / tc / 1e9, 5)}
    print(json.dumps(out), flush=True)


def run_polish(args):
    """A 400 kb draft (3% error) polished end to end from 60x reads at 8%
    error (plain-string params, 3 consensus rounds): realign batch + POA
    graphs, consensus rounds, merge.  value = reads/s of the whole run."""
    import numpy as np
    from marginpolish_amd import pairhmm, polish, sim
    from marginpolish_amd.sim import banded_edit_distance
    params_path = os.path.join(ROOT, "tests", "golden", "params", "allParams.np.no_rle.json")
    t0 = time.time()
    truth, draft, chunks = sim.polish_workload(seed=11, genome_len=400_000, depth=60, mean=8000, sd=6000,
                                               chunk_size=10_000, boundary=250, rle=False)
    nreads = sum(len(c.reads) for c in chunks)
    log("polish: %d chunks, %d chunk reads, built in %.1f s" % (len(chunks), nreads, time.time() - t0))
    params = pairhmm.PolishParams(params_path)
    cfg = polish.PolishConfig(params_path)
    cfg.maxRealignmentPolishIterations = 0
    cfg.maxPoaConsensusIterations = 3
    cfg.chunkBoundary = 250
    refs, reads, anchors = [c.ref for c in chunks], [c.reads for c in chunks], [c.anchors for c in chunks]
    eng = pairhmm.default_engine()
    runs = []
    for _ in range(2):
        t1 = time.perf_counter()
        merged = polish.polish_contig(params, cfg, refs, reads, anchors, engine=eng)
        runs.append(time.perf_counter() - t1)
    chars = np.frombuffer(b"ACGT", np.uint8)
    t_str, d_str = chars[truth].tobytes().decode(), chars[draft].tobytes().decode()
    d0 = banded_edit_distance(d_str, t_str, 1500)
    d1 = banded_edit_distance(merged, t_str, 1500)
    log("polish: %.2f s (first %.2f s); edit distance to truth: draft %d, polished %d" % (runs[1], runs[0], d0, d1))
    out = {"metric": "end-to-end polishing reads/s (realign + consensus rounds + merge)",
           "value": round(nreads / runs[1], 1), "unit": "reads/s", "n_gpus": 1, "steps": 1, "warmup": 1,
           "ms_per_step": round(runs[1] * 1e3, 1), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": "polish: 400 kb draft at 3% error, 60x reads (mean 8 kb) at 8%, 40 chunks of 10 kb "
                                  "(+-250 bp), plain-string params (allParams.np.no_rle.json), 3 consensus rounds",
                      "chunk_reads": nreads},
           "draft_edit_distance": d0, "polished_edit_distance": d1}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
