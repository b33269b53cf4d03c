// mp_pairhmm.hip -- the C-ABI batch driver around the gfx950 pair-HMM
// kernels (mp_split.hpp: posterior modes, mp_seq.hpp: forward-only and
// expectation modes, mp_wide.hpp: bands wider than 64 cells).
//
// A batch is every read/reference pair of one realignment iteration
// (poa_realign, polisher.c:656-703).  mp_batch_create crops, splits
// (getSplitPoints, pairwiseAligner.c:899-926), builds bands and traceback
// windows on the host once, uploads them with the encoded symbols, and sizes
// the device rings; mp_batch_run is one persistent launch per kernel; the
// fetch functions lay the posterior records out as the reference's three
// lists per pair.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <chrono>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <numeric>
#include <string>
#include <thread>
#include <map>
#include <memory>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "mp_internal.hpp"
#include "mp_lists.hpp"
#include "mp_ref.hpp"
#include "mp_seq.hpp"
#include "mp_split.hpp"
#include "mp_wide.hpp"
#include "mp_wide5.hpp"

using namespace mpk;

namespace {

__global__ void fill_i32(int32_t *p, int32_t v, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

#define HIP_TRY(call)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) {                                                               \
            mp::set_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #call); \
            return MP_EHIP;                                                                   \
        }                                                                                     \
    } while (0)

struct HostJob {
    int64_t pair;
    int64_t xOff, yOff;  // split offsets inside the (cropped) pair
    int64_t symX, symY;
    int32_t lX, lY;
    int32_t flags;
    int64_t bandOff;
    int32_t winOff, nWin;
    int32_t wmax;
    int32_t cls;  // 0..2 narrow rows of 16/32/64 lanes, 3 wide
    int64_t cells;
    int32_t cap[2];
    int64_t outOff;
    int32_t ringNeed;
    int32_t fstart = 0;  // forward start anti-diagonal (window jobs)
};

// A sub-problem's anchors, in place in the caller's array: n triples from a,
// coordinates read as (x - xo, y - yo).
struct JobAnc {
    const int64_t *a = nullptr;
    int64_t n = 0, xo = 0, yo = 0;
};

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    mp_status alloc(size_t count) {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = count;
        if (count == 0) return MP_OK;
        hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            mp::set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
            return MP_ENOMEM;
        }
        return MP_OK;
    }
    mp_status upload(const T *v, size_t count, hipStream_t s) {
        mp_status st = alloc(count);
        if (st != MP_OK || count == 0) return st;
        HIP_TRY(hipMemcpyAsync(p, v, count * sizeof(T), hipMemcpyHostToDevice, s));
        return MP_OK;
    }
    mp_status upload(const std::vector<T> &v, hipStream_t s) { return upload(v.data(), v.size(), s); }
};

}  // namespace

// Device scratch of the allocation-free entry (mp_pairhmm_forward_log_prob_noalloc):
// one copy per GPU, held until the process ends.  Host code reaches it only
// through the address hipGetSymbolAddress returns for the current device.
constexpr size_t kNoAllocScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

struct mp_engine {
    int device = 0;
    unsigned char *scratch = nullptr;  // this device's copy of `scratch`
    int cus = 0;
    int splitBlocksPerCU = 1;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    // pinned double buffer for large device-to-host result copies (allocated at first use)
    static constexpr size_t kStageBytes = (size_t)128 << 20;
    void *stage[2] = {nullptr, nullptr};
    hipEvent_t stageEv[2] = {nullptr, nullptr};
    // pinned host copies of a batch's three lists for the POA augmentation (grown on demand, reused)
    void *listHost[3] = {nullptr, nullptr, nullptr};
    size_t listHostBytes[3] = {0, 0, 0};
    // pinned staging of a batch's symbol pool [0] and band entries [1]: the planner writes them
    // there and they go to the device by DMA (grown on demand with headroom, reused); batch
    // creation holds createMu while it uses them
    void *upHost[2] = {nullptr, nullptr};
    size_t upHostBytes[2] = {0, 0};
    std::mutex createMu;
    // autotuner results: batch shape key -> (narrow workgroup threads, five-state waves per workgroup)
    std::map<uint64_t, std::pair<int32_t, int32_t>> tuned;
};

struct mp_batch {
    mp_engine *eng = nullptr;
    int32_t mode = 0;
    int32_t flags = 0;
    int64_t npairs = 0;
    std::vector<HostJob> jobs;          // in device order
    std::vector<int64_t> pairCrop;      // reference crop offset per pair
    std::vector<uint8_t> pairStrand;    // 1 = forward-strand read
    std::vector<int64_t> pairFirstJob;  // index into `order` of a pair's first slot
    std::vector<int32_t> order;         // slots: pair-major device job indices (split order, a sub-problem's
                                        // window jobs last window first; -1 = empty sub-problem)
    std::vector<int32_t> slotK;         // per slot: window jobs of its sub-problem at its first slot, else 0
    std::vector<int32_t> pairNJobs;     // slots per pair
    std::vector<DevWin> winHost;
    int64_t cells = 0;
    int32_t ringLen = kK;
    int32_t classBegin[kNumClasses] = {0, 0, 0}, classEnd[kNumClasses] = {0, 0, 0};
    KParams kp;
    DevBuf<DevJob> dJobs;
    DevBuf<uint32_t> dSym;
    DevBuf<int32_t> dBand;
    DevBuf<DevWin> dWin;
    DevBuf<int4> dOut;
    DevBuf<int32_t> dJobMap;            // a partial re-run's job map
    DevBuf<int32_t> dCounts;
    DevBuf<int32_t> dWinEnd;
    DevBuf<int2> dLogTotal;
    DevBuf<int2> dChain;
    DevBuf<float> dRing;
    DevBuf<int32_t> dRingExp;
    DevBuf<float> dTables;
    DevBuf<int32_t> dQueue;
    DevBuf<double> dExpect;
    DevBuf<double> dEmit;  // expected emission counts by emission-table entry (expectation mode)
    DevBuf<int32_t> dError;
    DevBuf<float> dRingWide;
    int32_t wideBegin = 0, wideEnd = 0, wideC = 0, ringWideLen = 0, wideWaves = 0;
    int64_t ringWideCells = 0;          // 3-state wide kernel: ring cells per wave
    // the last run's exact-size re-runs, jobs over their record capacity and their missing records
    int32_t lastReruns = 0;
    int64_t lastOvfJobs = 0, lastOvfMissing = 0;
    // host time of mp_batch_create's phases (ms): checks and split points, symbols and jobs,
    // bands and windows, window jobs, device order and record regions, slots, upload and
    // device buffers, list buffers
    double prepMs[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    DevBuf<int32_t> dRingWideIdx;       // 3-state wide kernel: row offsets, ringWideLen per wave
    int32_t modelStates = 3;
    hipGraphExec_t graph = nullptr;  // the run's resets and launches, captured once per parameter block
    hipGraphExec_t graphLists = nullptr;  // the run's list assembly and log-likelihoods
    int32_t narrowBlocks = 0;
    int32_t narrowThreads = 512;  // workgroup size of the narrow kernels (the autotuner may pick 256)
    int32_t narrowWaves = 0;      // resident narrow wavefronts the rings are sized for
    int32_t wide5Waves = 0;       // waves per five-state workgroup (0: min(columns, 8))
    uint64_t tuneKey = 0;
    int64_t outTotal = 0;
    double kernelMs = 0.0;
    double listMs = 0.0;
    bool ran = false;
    std::vector<int32_t> hCounts;  // 2 per job, fetched after each run
    // lists and log-likelihoods, assembled on the device by every run (mp_lists.hpp)
    bool listsReady = false;            // listStart / pairLogProb copied to the host since the last run
    std::vector<mpk::JobLists> jlHost;  // per device job
    DevBuf<mpk::JobLists> dJL;
    DevBuf<int32_t> dSlotJob, dSlotK, dPairSlot, dOverflow;
    DevBuf<int64_t> dSlotBase, dStarts;
    DevBuf<int4> dSegCnt;     // per window and side: records of each list
    DevBuf<int64_t> dSegBase;  // per window and side: first position in each list
    // tiles of mpk::kBigTile records over the one-window job sides (by record capacity)
    DevBuf<int32_t> dSegTile;
    DevBuf<int4> dTileMeta, dTileCnt, dTilePre;
    int64_t nListTiles = 0;
    bool multiWin = false;  // some job has several windows (exact-chain batches)
    DevBuf<int4> dListTotals;
    int64_t listCap[3] = {0, 0, 0};     // tuples the device lists hold (grown when a run needs more)
    // bits of x in a device list entry (mpk::put_tuple): 64-bit entries when the batch's
    // longest reference and read fit beside the 24-bit weight, else 0 (16-byte entries)
    int32_t packBits = 0;
    int64_t records = 0;                // posterior records of the last run
    DevBuf<double> dPairLL, dBatchLL;
    DevBuf<mpk::ListTuple> dList[3];    // the lists on the device (copied out by fetch)
    std::vector<int64_t> listStart[3];  // per pair, npairs + 1
    std::vector<double> pairLogProb;
    double batchLogProb = 0.0;
    // reference-arithmetic batches (MP_BATCH_REFERENCE_ARITHMETIC, mp_ref.hpp)
    bool ref = false;
    std::vector<mpk::RefJob> refJobs;
    mpk::RefParams rp{};
    int32_t refBlocks = 0;
    int32_t refThreads = 256;  // workgroup size of the reference-arithmetic kernel (256 or 512)
    DevBuf<mpk::RefJob> dRefJobs;
    DevBuf<uint16_t> dRawSym;
    DevBuf<int2> dLR;
    DevBuf<double> dRepeat, dRefF, dRefB, dRefDots, dRefLogTotal;
    DevBuf<int64_t> dRefOffs, dSlotBase3;
    DevBuf<mpk::ListTuple> dRaw;
};

namespace {

bool use_split(int32_t mode) { return mode == MP_MODE_POSTERIORS || mode == MP_MODE_MATCH_ONLY; }

mp_status launch_narrow(mp_batch *b, int grid, int block, const KParams &kp) {
    if (use_split(b->mode)) {
        hipLaunchKernelGGL(pairhmm_split_kernel, dim3(grid), dim3(block), kSplitLdsBytes, b->eng->stream, kp);
    } else {
        hipLaunchKernelGGL(pairhmm_seq_kernel, dim3(grid), dim3(block),
                           b->mode == MP_MODE_EXPECTATIONS ? kSeqLdsBytesExp : kSeqLdsBytes, b->eng->stream, kp);
    }
    HIP_TRY(hipGetLastError());
    return MP_OK;
}

mp_status launch_wide(mp_batch *b, int grid, const KParams &kp) {
    const int w5 = b->wide5Waves > 0 ? std::min(b->wide5Waves, wide5_waves(kp.wideC)) : wide5_waves(kp.wideC);
    const dim3 block5(64 * w5);
    if (kp.nStates == 5 && (b->flags & MP_BATCH_FP64))
        hipLaunchKernelGGL(pairhmm_wide5_kernel<double>, dim3(grid), block5, wide5_lds_bytes<double>(kp.wideC),
                           b->eng->stream, kp);
    else if (kp.nStates == 5)
        hipLaunchKernelGGL(pairhmm_wide5_kernel<float>, dim3(grid), block5, wide5_lds_bytes<float>(kp.wideC),
                           b->eng->stream, kp);
    else {
        // grid = job waves: workgroups of up to wide_group_waves(C) waves
        const int n = std::max(1, std::min(grid, wide_group_waves(kp.wideC)));
        hipLaunchKernelGGL(pairhmm_wide_kernel, dim3((grid + n - 1) / n), dim3(64 * n), wide_lds_bytes(kp.wideC, n),
                           b->eng->stream, kp);
    }
    HIP_TRY(hipGetLastError());
    return MP_OK;
}

mp_status reset_i32(mp_batch *b, int32_t *q, int n) {
    hipLaunchKernelGGL(fill_i32, dim3((n + 255) / 256), dim3(256), 0, b->eng->stream, q, 0, n);
    HIP_TRY(hipGetLastError());
    return MP_OK;
}

bool lists_mode(const mp_batch *b) { return b->mode == MP_MODE_POSTERIORS || b->mode == MP_MODE_MATCH_ONLY; }

// Device list l with room for `tuples` entries (8 or 16 bytes each).
mp_status alloc_list(mp_batch *b, int l, int64_t tuples) {
    const int64_t units = b->packBits ? (tuples + 1) / 2 : tuples;  // in 16-byte units
    mp_status st = b->dList[l].alloc((size_t)std::max<int64_t>(units, 1));
    if (!st) b->listCap[l] = tuples;
    return st;
}

// Per-job list metadata (record regions, capacities, slots) and the device
// buffers of the list assembly, sized by the record capacity b->outTotal.
mp_status alloc_lists(mp_batch *b) {
    if (!lists_mode(b)) return MP_OK;  // (no records: forward-only and expectation batches)
    const int64_t nj = (int64_t)b->jobs.size(), ns = (int64_t)b->order.size();
    std::vector<int32_t> jobSlot((size_t)std::max<int64_t>(nj, 1), -1);
    for (int64_t sl = 0; sl < ns; sl++)
        if (b->order[sl] >= 0) jobSlot[b->order[sl]] = (int32_t)sl;
    b->jlHost.resize((size_t)nj);
    for (int64_t q = 0; q < nj; q++) {
        const HostJob &hj = b->jobs[q];
        mpk::JobLists &J = b->jlHost[q];
        J.outF = hj.outOff;
        J.outB = hj.outOff + hj.cap[0];
        J.bandOff = hj.bandOff;
        J.capF = hj.cap[0];
        J.capB = hj.cap[1];
        J.winOff = hj.winOff;
        J.nWin = hj.nWin;
        J.xAdd = (int32_t)(hj.xOff + b->pairCrop[hj.pair] - 1);
        J.yAdd = (int32_t)(hj.yOff - 1);
        J.slot = jobSlot[q];
        J.pairSlot = (int32_t)b->pairFirstJob[hj.pair];
    }
    hipStream_t s = b->eng->stream;
    mp_status st;
    if ((st = b->dJL.upload(b->jlHost, s))) return st;
    const int64_t cap = std::max<int64_t>(b->outTotal, 1);
    const size_t nseg = 2 * std::max<size_t>(b->winHost.size(), 1);
    if ((st = b->dSegCnt.alloc(nseg)) || (st = b->dSegBase.alloc(3 * nseg)) ||
        (st = b->dSlotBase3.alloc(3 * (b->order.size() + 1))) || (st = b->dListTotals.alloc(1)))
        return st;
    // (host arrays kept to the stream synchronisation below: the uploads are asynchronous)
    std::vector<int32_t> segTile((size_t)std::max<int64_t>(2 * nj, 1), 0);
    std::vector<int4> tileMeta;
    {
        int64_t nt = 0;
        b->multiWin = false;
        for (int64_t q = 0; q < nj; q++) {
            const mpk::JobLists &J = b->jlHost[q];
            b->multiWin |= J.nWin > 1;
            for (int side = 0; side < 2; side++) {
                segTile[2 * q + side] = (int32_t)nt;
                if (J.nWin == 1) nt += ((int64_t)(side ? J.capB : J.capF) + mpk::kBigTile - 1) / mpk::kBigTile;
                if (nt >= INT32_MAX) {
                    mp::set_error("batch too large: more than 2^31 list tiles (split it into several batches)");
                    return MP_EUNSUPPORTED;
                }
            }
        }
        tileMeta.assign((size_t)std::max<int64_t>(nt, 1), make_int4(0, 0, 0, 0));
        for (int64_t q = 0; q < nj; q++)
            for (int side = 0; side < 2; side++) {
                const mpk::JobLists &J = b->jlHost[q];
                const int64_t k0 = segTile[2 * q + side], k1 = 2 * q + side + 1 < 2 * nj ? segTile[2 * q + side + 1] : nt;
                for (int64_t k = k0; k < k1; k++)
                    tileMeta[k] = make_int4((int)(2 * q + side), (int)((k - k0) * mpk::kBigTile), 2 * J.winOff + side, 0);
            }
        b->nListTiles = nt;
        if ((st = b->dTileMeta.upload(tileMeta, s)) || (st = b->dSegTile.upload(segTile, s)) ||
            (st = b->dTileCnt.alloc((size_t)std::max<int64_t>(nt, 1))) ||
            (st = b->dTilePre.alloc((size_t)std::max<int64_t>(nt, 1))))
            return st;
    }
    // first guess of the list lengths (a record holds at most one tuple of each list; the
    // capacities count two records per posterior anti-diagonal and side): half the records
    for (int l = 0; l < 3; l++) {
        const int64_t want = std::max<int64_t>(cap / 2, 1024);
        if (b->listCap[l] < want && (st = alloc_list(b, l, want))) return st;
    }
    HIP_TRY(hipStreamSynchronize(s));
    return MP_OK;
}

// Reference-arithmetic batch (mp_ref.hpp): raw symbols, the full band of every
// sub-problem as (xmyL, xmyR) per anti-diagonal, per-workgroup working memory
// sized by the largest job, tuple regions per job and list.
mp_status create_reference_batch(mp_batch *b, const mp_model *m, const mp_pair *pairs, const mp_band_params *p,
                                 const std::vector<int64_t> &symOff, const std::vector<int64_t> &pairFirstSub,
                                 std::vector<HostJob> &jobs, const std::vector<JobAnc> &jobAnchors) {
    mp_engine *e = b->eng;
    const int64_t np = b->npairs, nj = (int64_t)jobs.size();
    b->ref = true;
    // raw symbols at the same pool positions as the encoded ones (N before each sequence)
    std::vector<uint16_t> raw((size_t)(symOff[np] + kSymPad), (uint16_t)4);
    for (int64_t i = 0; i < np; i++) {
        const mp_pair &pr = pairs[i];
        const int64_t baseX = symOff[i] + 1, lXc = symOff[i + 1] - symOff[i] - 2 - pr.lY, baseY = baseX + lXc + 1;
        for (int64_t x = 0; x < lXc; x++) raw[baseX + x] = pr.sX[b->pairCrop[i] + x];
        for (int64_t y = 0; y < pr.lY; y++) raw[baseY + y] = pr.sY[y];
    }
    // bands
    std::vector<int64_t> lrOff(nj + 1, 0);
    for (int64_t j = 0; j < nj; j++) lrOff[j + 1] = lrOff[j] + jobs[j].lX + jobs[j].lY + 1;
    std::vector<int2> lr((size_t)std::max<int64_t>(lrOff[nj], 1));
    std::vector<int64_t> cells(nj, 0);
    std::vector<int32_t> wmax(nj, 0);
    std::atomic<int> bad{0};
    {
        const int nth = (int)std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
        std::atomic<int64_t> next{0};
        mp::parallel_run(nth, [&](int) {
                std::vector<int32_t> L, R;
                for (int64_t j; (j = next.fetch_add(1)) < nj;) {
                    const HostJob &hj = jobs[j];
                    const int64_t D = hj.lX + hj.lY;
                    L.resize(D + 1);
                    R.resize(D + 1);
                    const JobAnc &a = jobAnchors[j];
                    if (!mp::build_band(a.a, a.n, a.xo, a.yo, hj.lX, hj.lY, p->diagonalExpansion,
                                        p->dynamicAnchorExpansion != 0, L.data(), R.data())) {
                        bad = 1;
                        continue;
                    }
                    int64_t c = 0;
                    int32_t wm = 0;
                    for (int64_t d = 0; d <= D; d++) {
                        lr[lrOff[j] + d] = make_int2(L[d], R[d]);
                        const int32_t wd = (R[d] - L[d]) / 2 + 1;
                        c += wd;
                        wm = std::max(wm, wd);
                    }
                    cells[j] = D > 0 ? c : 0;
                    wmax[j] = wm;
                }
            });
    }
    if (bad) {
        mp::set_error("invalid band (anchors inconsistent with the sequences)");
        return MP_EANCHORS;
    }
    // device order: largest jobs first (the queue is taken in order)
    std::vector<int32_t> ord(nj);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int32_t a, int32_t c) { return cells[a] > cells[c]; });
    std::vector<int32_t> devIndex(nj);
    b->refJobs.resize(nj);
    b->jobs.resize(nj);
    int64_t maxCells = 1, maxD = 1, outOff = 0, sumD = 0;
    int32_t maxW = 1;
    b->cells = 0;
    for (int64_t q = 0; q < nj; q++) {
        const int32_t j = ord[q];
        devIndex[j] = (int32_t)q;
        const HostJob &hj = jobs[j];
        const int64_t D = hj.lX + hj.lY;
        mpk::RefJob &R = b->refJobs[q];
        R.symX = hj.symX;
        R.symY = hj.symY;
        R.lrOff = lrOff[j];
        R.lX = hj.lX;
        R.lY = hj.lY;
        R.flags = hj.flags;
        R.xAdd = (int32_t)(hj.xOff + b->pairCrop[hj.pair] - 1);
        R.yAdd = (int32_t)(hj.yOff - 1);
        for (int l = 0; l < 3; l++) {
            // (at most one tuple of each list per cell; a job that writes more than its region
            // holds makes the run size every region exactly and run again)
            R.cap[l] = (l == 0 || b->mode == MP_MODE_POSTERIORS) ? (int32_t)std::min<int64_t>(D + 64, INT32_MAX / 4) : 0;
            R.outOff[l] = outOff;
            outOff += R.cap[l];
        }
        b->jobs[q] = hj;
        maxCells = std::max(maxCells, cells[j]);
        maxD = std::max(maxD, D);
        maxW = std::max(maxW, wmax[j]);
        sumD += D;
        b->cells += cells[j];
    }
    b->outTotal = outOff;
    b->refThreads = maxW > 512 ? 512 : 256;
    {
        std::vector<int32_t> slots, slotK;
        for (int64_t i = 0; i < np; i++) {
            b->pairFirstJob[i] = (int64_t)slots.size();
            const int32_t nsub = (i + 1 < np ? (int32_t)(pairFirstSub[i + 1] - pairFirstSub[i])
                                             : (int32_t)(nj - pairFirstSub[i]));
            for (int32_t sp = 0; sp < nsub; sp++) {
                const int64_t j = pairFirstSub[i] + sp;
                slots.push_back(jobs[j].lX + jobs[j].lY > 0 ? devIndex[j] : -1);
                slotK.push_back(1);
            }
            b->pairNJobs[i] = nsub;
        }
        b->order.swap(slots);
        b->slotK.swap(slotK);
    }
    // per-workgroup working memory, as many workgroups as half the free memory holds (at most 2 per CU)
    const size_t perWG = sizeof(double) * (size_t)(3 * maxCells + 9 * (int64_t)maxW + 2 * (int64_t)maxW) +
                         sizeof(int64_t) * (size_t)(maxD + 2);
    size_t freeB = 0, totalB = 0;
    HIP_TRY(hipMemGetInfo(&freeB, &totalB));
    const int64_t fit = (int64_t)(freeB / 2 / perWG);
    b->refBlocks = (int32_t)std::min<int64_t>({std::max<int64_t>(nj, 1), (int64_t)e->cus * 2, fit});
    if (b->refBlocks < 1) {
        mp::set_error("reference arithmetic: a sub-problem's working memory (" + std::to_string(perWG) +
                      " bytes) exceeds the free device memory");
        return MP_ENOMEM;
    }
    hipStream_t s = e->stream;
    mp_status st;
    std::vector<int32_t> pairSlot((size_t)np + 1);
    for (int64_t i = 0; i < np; i++) pairSlot[i] = (int32_t)b->pairFirstJob[i];
    pairSlot[np] = (int32_t)b->order.size();
    const size_t G = (size_t)b->refBlocks;
    if ((st = b->dRefJobs.upload(b->refJobs, s)) || (st = b->dRawSym.upload(raw, s)) || (st = b->dLR.upload(lr, s)) ||
        (st = b->dRefF.alloc(G * 3 * (size_t)maxCells)) || (st = b->dRefB.alloc(G * 9 * (size_t)maxW)) ||
        (st = b->dRefDots.alloc(G * 2 * (size_t)maxW)) || (st = b->dRefOffs.alloc(G * (size_t)(maxD + 2))) ||
        (st = b->dRaw.alloc((size_t)std::max<int64_t>(outOff, 1))) ||
        (st = b->dCounts.alloc(3 * (size_t)std::max<int64_t>(nj, 1))) ||
        (st = b->dRefLogTotal.alloc((size_t)std::max<int64_t>(nj, 1))) || (st = b->dQueue.alloc(kNumClasses + 1)) ||
        (st = b->dError.alloc(1)) || (st = b->dOverflow.alloc(1)) || (st = b->dListTotals.alloc(1)) ||
        (st = b->dSlotJob.upload(b->order, s)) || (st = b->dSlotK.upload(b->slotK, s)) ||
        (st = b->dPairSlot.upload(pairSlot, s)) || (st = b->dSlotBase3.alloc(3 * (b->order.size() + 1))) ||
        (st = b->dStarts.alloc(3 * ((size_t)np + 1))) || (st = b->dPairLL.alloc(std::max<size_t>((size_t)np, 1))) ||
        (st = b->dBatchLL.alloc(1)))
        return st;
    const bool rle = m->mach[1].rle != 0;
    if (rle && (st = b->dRepeat.upload(m->repeatLog10.data(), (size_t)4 * 51 * 51, s))) return st;
    for (int l = 0; l < 3; l++) {
        if ((st = alloc_list(b, l, std::max<int64_t>(sumD / 2, 1024)))) return st;
    }
    mpk::RefParams &rp = b->rp;
    std::memset(&rp, 0, sizeof(rp));
    const mp::Machine &f = m->mach[1];
    const int ti[9] = {mp::kMatchContinue, mp::kMatchFromGapX, mp::kMatchFromGapY, mp::kGapOpenX, mp::kGapOpenY,
                       mp::kGapExtendX, mp::kGapExtendY, mp::kGapSwitchToX, mp::kGapSwitchToY};
    for (int k = 0; k < 9; k++) rp.model.trans[k] = f.trans[ti[k]];
    for (int sd = 0; sd < 2; sd++) {
        const mp::Machine &mc = m->mach[sd];
        for (int k = 0; k < 16; k++) rp.model.match[sd][k] = mc.match[k];
        for (int k = 0; k < 4; k++) {
            rp.model.gapX[sd][k] = mc.gapX[k];
            rp.model.gapY[sd][k] = mc.gapY[k];
        }
    }
    rp.model.rle = rle ? 1 : 0;
    rp.jobs = b->dRefJobs.p;
    rp.njobs = (int32_t)nj;
    rp.sym = b->dRawSym.p;
    rp.lr = b->dLR.p;
    rp.repeat = rle ? b->dRepeat.p : nullptr;
    rp.threshold = p->threshold;
    rp.withIndels = b->mode == MP_MODE_POSTERIORS;
    rp.poly = 1;
    rp.minDiags = p->minDiagsBetweenTraceBack;
    rp.tbd = p->traceBackDiagonals;
    rp.expansion = p->diagonalExpansion;
    rp.F = b->dRefF.p;
    rp.fStride = 3 * maxCells;
    rp.B = b->dRefB.p;
    rp.bStride = 3 * (int64_t)maxW;
    rp.offs = b->dRefOffs.p;
    rp.oStride = maxD + 2;
    rp.dots = b->dRefDots.p;
    rp.wStride = maxW;
    rp.out = b->dRaw.p;
    rp.counts = b->dCounts.p;
    rp.logTotal = b->dRefLogTotal.p;
    rp.queue = b->dQueue.p;
    rp.error = b->dError.p;
    HIP_TRY(hipStreamSynchronize(s));
    return MP_OK;
}

// fp64: every job runs on the five-state kernel in FP64; a 3-state machine is
// embedded there with its long gaps unreachable.
void build_tables(const mp_model *m, bool fp64, std::vector<float> &tab, KParams &kp) {
    tab.assign(kTabFloats, 0.f);
    for (int s = 0; s < 2; s++) {
        const mp::Machine &mc = m->mach[s];
        for (int i = 0; i < 5; i++) {
            tab[kTabGX + s * 5 + i] = (float)std::exp(i >= 4 ? -1.386294361 : mc.gapX[i]);
            tab[kTabGY + s * 5 + i] = (float)std::exp(i >= 4 ? -1.386294361 : mc.gapY[i]);
            for (int j = 0; j < 5; j++) {
                const double lp = (i >= 4 || j >= 4) ? -2.772588722 : mc.match[i * 4 + j];
                tab[kTabE + s * 25 + i * 5 + j] = (float)std::exp(lp);
            }
        }
    }
    const bool rle = m->mach[1].rle != 0;
    for (int row = 0; row < 4; row++)
        for (int u = 0; u < kRL; u++)
            for (int o = 0; o < kRL; o++) {
                const double l10 =
                    m->repeatLog10[(size_t)row * mp::kMaxRepeat * mp::kMaxRepeat + (size_t)u * mp::kMaxRepeat + o];
                tab[kTabRL + (row * kRL + u) * kRL + o] = rle ? (float)std::exp(2.3025 * l10) : 1.f;
            }
    const mp::Machine &f = m->mach[1];
    auto P = [&](int i) { return (float)std::exp(f.trans[i]); };
    // T[from*3 + to]
    kp.T[0] = P(mp::kMatchContinue); kp.T[1] = P(mp::kGapOpenX);     kp.T[2] = P(mp::kGapOpenY);
    kp.T[3] = P(mp::kMatchFromGapX); kp.T[4] = P(mp::kGapExtendX);   kp.T[5] = P(mp::kGapSwitchToY);
    kp.T[6] = P(mp::kMatchFromGapY); kp.T[7] = P(mp::kGapSwitchToX); kp.T[8] = P(mp::kGapExtendY);
    kp.startV[0][0] = 1.f; kp.startV[0][1] = 0.f; kp.startV[0][2] = 0.f;
    kp.startV[1][0] = 0.f; kp.startV[1][1] = 1.f; kp.startV[1][2] = 1.f;
    const double *et = m->endPinned ? m->endTrans : f.trans;
    auto E = [&](int i) { return (float)std::exp(et[i]); };
    kp.endV[0][0] = E(mp::kMatchContinue);
    kp.endV[0][1] = E(mp::kMatchFromGapX);
    kp.endV[0][2] = E(mp::kMatchFromGapY);
    kp.endV[1][0] = (float)std::exp((et[mp::kGapOpenX] + et[mp::kGapOpenY]) / 2.0);
    kp.endV[1][1] = E(mp::kGapExtendX);
    kp.endV[1][2] = E(mp::kGapExtendY);
    kp.nStates = f.nStates == 5 || fp64 ? 5 : 3;
    if (f.nStates == 5) {
        // Boundary vectors of the five-state machine, the three-state ones
        // (stateMachine.c:518-557) extended to the long gaps: start in M;
        // ragged start in any gap state; end through each state's transition
        // to M; ragged end: M at the mean of the short-gap opens (log space),
        // each gap state at its own extension.
        const double *l5 = f.trans5;
        for (int i = 0; i < 25; i++) kp.T5[i] = std::exp(l5[i]);
        for (int st = 0; st < 5; st++) {
            kp.startV5[0][st] = st == mp::kS5Match ? 1.0 : 0.0;
            kp.startV5[1][st] = st == mp::kS5Match ? 0.0 : 1.0;
            kp.endV5[0][st] = std::exp(l5[st * 5 + mp::kS5Match]);
            kp.endV5[1][st] = st == mp::kS5Match ? std::exp((l5[mp::kS5ShortX] + l5[mp::kS5ShortY]) / 2.0)
                                                 : std::exp(l5[st * 5 + st]);
        }
    } else if (fp64) {
        // the 3-state machine on the five-state kernel: long gaps unreachable
        const double *t = f.trans;
        const int M = mp::kS5Match, X = mp::kS5ShortX, Y = mp::kS5ShortY;
        for (int i = 0; i < 25; i++) kp.T5[i] = 0.0;
        kp.T5[M * 5 + M] = std::exp(t[mp::kMatchContinue]);
        kp.T5[M * 5 + X] = std::exp(t[mp::kGapOpenX]);
        kp.T5[M * 5 + Y] = std::exp(t[mp::kGapOpenY]);
        kp.T5[X * 5 + M] = std::exp(t[mp::kMatchFromGapX]);
        kp.T5[X * 5 + X] = std::exp(t[mp::kGapExtendX]);
        kp.T5[X * 5 + Y] = std::exp(t[mp::kGapSwitchToY]);
        kp.T5[Y * 5 + M] = std::exp(t[mp::kMatchFromGapY]);
        kp.T5[Y * 5 + X] = std::exp(t[mp::kGapSwitchToX]);
        kp.T5[Y * 5 + Y] = std::exp(t[mp::kGapExtendY]);
        for (int st = 0; st < 5; st++) kp.startV5[0][st] = kp.startV5[1][st] = kp.endV5[0][st] = kp.endV5[1][st] = 0.0;
        kp.startV5[0][M] = 1.0;
        kp.startV5[1][X] = kp.startV5[1][Y] = 1.0;
        kp.endV5[0][M] = std::exp(et[mp::kMatchContinue]);
        kp.endV5[0][X] = std::exp(et[mp::kMatchFromGapX]);
        kp.endV5[0][Y] = std::exp(et[mp::kMatchFromGapY]);
        kp.endV5[1][M] = std::exp((et[mp::kGapOpenX] + et[mp::kGapOpenY]) / 2.0);
        kp.endV5[1][X] = std::exp(et[mp::kGapExtendX]);
        kp.endV5[1][Y] = std::exp(et[mp::kGapExtendY]);
    }
}

// Crossing diagonal of window k for the concurrent sweeps: the forward row
// reaches cx after cx + 1 anti-diagonals of the job, the backward row reaches
// cx + 1 after the earlier windows' backward sweeps plus tb - cx; balance the
// two, prefer cx = 7 (mod 8) (both sweeps then cross at cell boundaries), and
// keep cx in [previous tb, tb - 1] ([0, tb - 1] for the first window).
void choose_crossings(std::vector<DevWin> &w) {
    int64_t bwdDone = 0;
    for (size_t k = 0; k < w.size(); k++) {
        const int64_t low = k == 0 ? 0 : w[k - 1].tb, high = (int64_t)w[k].tb - 1;
        int64_t c = (bwdDone + w[k].tb - 1) / 2;
        c = ((c + 4) & ~(int64_t)(kK - 1)) - 1;
        c = std::min(std::max(c, low), high);
        w[k].cx = (int32_t)c;
        bwdDone += w[k].tb - w[k].lo;
    }
}

// Narrow band entries: flag the 8-diagonal cells whose neighbour exchange is
// not a one-lane shift (going forward, going backward).
void mark_slow(const int32_t *XL, int64_t D, int32_t *bj) {
    auto in = [](int64_t v, int64_t a, int64_t b) { return v >= a && v <= b; };
    for (int64_t d = 0; d <= D; d++) {
        const bool fastF = d == 0 || (in(XL[d] - XL[d - 1], 0, 1) && (d < 2 || in(XL[d] - XL[d - 2], 0, 2)));
        const bool fastB = d == D || (in(XL[d] - XL[d + 1], -1, 0) && (d + 2 > D || in(XL[d] - XL[d + 2], -2, 0)));
        const int64_t c0 = d & ~(int64_t)(kK - 1);
        if (!fastF)
            for (int64_t q = c0; q < c0 + kK; q++) bj[q] |= kSlowF;
        if (!fastB)
            for (int64_t q = c0; q < c0 + kK; q++) bj[q] |= kSlowB;
    }
}

// First anti-diagonal of the forward sweep of window job k (window start lo):
// kWarmup before lo, on a cell boundary; the first window starts at 0.
inline int64_t window_fstart(size_t k, int64_t lo) {
    return k == 0 ? 0 : (std::max<int64_t>(0, lo - kWarmup) & ~(int64_t)(kK - 1));
}

// One job's band entries (bj: padEnd entries, the band padded to whole cells),
// band class, cell count and traceback windows with their crossings and ring
// need: the planner's per-job step (pairwiseAligner.c:120-226, 715-783).
// Returns false when the anchors do not fit the sequences.
struct PlanScratch {
    std::vector<int32_t> L, R, XL, WD;  // band extent, first x and width per anti-diagonal
    std::vector<mp::Window> w;
    // per window: the band's width beyond the anchors' own 2 W + 1 cells (at most
    // 16 per anti-diagonal) over its posterior anti-diagonals (lo, from], and the
    // band cells of [lo, tb]
    std::vector<int64_t> wex, wcells;
};
bool plan_job(HostJob &hj, const JobAnc &anc, const mp_band_params &bp, int32_t mode, bool five, int32_t *bj,
              int64_t padEnd, PlanScratch &sc, std::vector<DevWin> &wj) {
    const mp_band_params *p = &bp;
    const bool split = use_split(mode);
    std::vector<int32_t> &L = sc.L, &R = sc.R, &XL = sc.XL;
    std::vector<mp::Window> &w = sc.w;
    const int64_t D = hj.lX + hj.lY;
    L.resize(D + 1);
    R.resize(D + 1);
    const bool dyn = mode != MP_MODE_FORWARD_ONLY && p->dynamicAnchorExpansion != 0;
    if (!mp::build_band(anc.a, anc.n, anc.xo, anc.yo, hj.lX, hj.lY, p->diagonalExpansion, dyn, L.data(), R.data()))
        return false;
    // traceback windows (they read only the band's extent)
    w.clear();
    if (D > 0 && mode == MP_MODE_FORWARD_ONLY) w.push_back(mp::Window{(int32_t)D, 0, (int32_t)D});
    else if (D > 0) mp::build_windows(L.data(), R.data(), D, *p, w);
    const size_t nw = w.size();
    sc.wex.assign(nw, 0);
    sc.wcells.assign(nw, 0);
    // Passes over the anti-diagonals (each a simple loop the compiler vectorises):
    // width, cells, xL = x of the band's first cell ((d + xmyL) / 2) and the band
    // entry (width - 1 in the low bits; the narrow kernels' flags are added below);
    // the x-alignment of rows of 16 lanes (the cells of three consecutive
    // anti-diagonals, with the neighbour each side reads, must span at most 16 x);
    // per window the excess width over its posterior anti-diagonals (lo, from]
    // (consecutive ranges) and the cells of [lo, tb] (a wide window's ring).
    XL.resize(D + 1);
    sc.WD.resize(D + 1);
    int32_t *xlA = XL.data(), *wdA = sc.WD.data();
    const int32_t *Lp = L.data(), *Rp = R.data();
    int32_t wmax = 0;
    int64_t cells = 0;
    const int32_t nominal = 2 * p->diagonalExpansion + 1;
    for (int64_t d = 0; d <= D; d++) {
        const int32_t wd = (Rp[d] - Lp[d]) / 2 + 1;
        const int32_t xl = (int32_t)((d + Lp[d]) / 2);
        wmax = std::max(wmax, wd);
        cells += wd;
        xlA[d] = xl;
        wdA[d] = wd;
        bj[d] = (int32_t)((uint32_t)xl << kBandShift) | (std::min(wd, 2048) - 1);
    }
    // forward reads of d: d - 1, d - 2 and the cell left of xL; backward reads of d - 2:
    // d - 1, d and the cell right of its last x
    int bad = 0;
    for (int64_t d = 0; d <= D; d++) {
        const int32_t xl = xlA[d], xr = xl + wdA[d] - 1;
        int32_t lo = xl - 1, hi = xr;
        if (d >= 1) {
            lo = std::min(lo, xlA[d - 1]);
            hi = std::max(hi, xlA[d - 1] + wdA[d - 1] - 1);
        }
        if (d >= 2) {
            lo = std::min(lo, xlA[d - 2]);
            hi = std::max(hi, xlA[d - 2] + wdA[d - 2] - 1);
            const int32_t lo2 = std::min(xlA[d - 2], std::min(xlA[d - 1], xl));
            const int32_t hi2 = std::max(xlA[d - 2] + wdA[d - 2], std::max(xlA[d - 1] + wdA[d - 1] - 1, xr));
            bad |= hi2 - lo2 > 15;
        }
        bad |= hi - lo > 15;
    }
    bool fits = !bad;
    if (fits) {  // the backward reads of the last two anti-diagonals (truncated at D)
        for (int64_t d = std::max<int64_t>(0, D - 1); d <= D; d++) {
            int32_t lo2 = xlA[d], hi2 = xlA[d] + wdA[d];
            for (int64_t e = d + 1; e <= D; e++) {
                lo2 = std::min(lo2, xlA[e]);
                hi2 = std::max(hi2, xlA[e] + wdA[e] - 1);
            }
            fits = fits && hi2 - lo2 <= 15;
        }
    }
    // anti-diagonal d joins the excess of the first window whose from >= d when d > its
    // lo, and the cells of every window whose [lo, tb] holds it: with lo, tb and from
    // non-decreasing over the windows, sums over ranges
    bool mono = true;
    for (size_t k = 1; k < nw; k++)
        mono &= w[k].lo >= w[k - 1].lo && w[k].tb >= w[k - 1].tb && w[k].from >= w[k - 1].from;
    if (mono) {
        for (size_t k = 0; k < nw; k++) {
            const int64_t a = std::max<int64_t>(w[k].lo, k ? w[k - 1].from : -1), e = std::min<int64_t>(w[k].from, D);
            int64_t ex = 0;
            for (int64_t d = std::max<int64_t>(a + 1, 0); d <= e; d++) ex += std::min(16, std::max(0, wdA[d] - nominal));
            sc.wex[k] = ex;
            int64_t c = 0;
            for (int64_t d = std::max<int64_t>(0, w[k].lo), f = std::min<int64_t>(w[k].tb, D); d <= f; d++) c += wdA[d];
            sc.wcells[k] = c;
        }
    } else {
        size_t kp = 0, kr = 0;
        for (int64_t d = 0; d <= D; d++) {
            while (kp < nw && d > w[kp].from) kp++;
            if (kp < nw && d > w[kp].lo) sc.wex[kp] += std::min(16, std::max(0, wdA[d] - nominal));
            while (kr < nw && d > w[kr].tb) kr++;
            for (size_t k = kr; k < nw && w[k].lo <= d; k++) sc.wcells[k] += wdA[d];
        }
    }
    hj.wmax = wmax;
    hj.cells = D > 0 ? cells : 0;
    hj.cls = five ? 3 : (wmax <= 16 ? (fits ? 0 : 1) : (wmax <= 32 ? 1 : (wmax <= 64 ? 2 : 3)));
    const bool narrow = hj.cls < 3;
    for (int64_t d = D + 1; d < padEnd; d++) bj[d] = bj[D];
    if (narrow && hj.cls > 0) mark_slow(XL.data(), D, bj);
    wj.clear();
    for (auto &x : w) wj.push_back(DevWin{x.tb, x.lo, x.from, x.tb - 1});
    if (split && narrow) choose_crossings(wj);
    // live ring entries.  Narrow kernels (concurrent sweeps): forward and
    // backward values of (lo_k, tb_{k+1}].  Wide kernels (the backward of
    // window k runs after its forward): forward values of [lo_k, tb_k].
    int64_t need = 1;
    for (size_t k = 0; k < wj.size(); k++) {
        const int64_t hi = narrow ? wj[std::min(k + 1, wj.size() - 1)].tb : wj[k].tb;
        need = std::max<int64_t>(need, hi - wj[k].lo + (narrow ? 3 * kK : 2));
    }
    hj.ringNeed = (int32_t)((need + kK - 1) / kK * kK);
    return true;
}

// A wide sub-problem that runs as window jobs: each window's own class from
// its band over [fstart - 2, tb + 2] (rows of 32 or 64 lanes, or wide; the
// narrow kernels read the band two anti-diagonals either side of their sweeps)
// and width maximum; when some window fits a narrow row, the band again in the
// narrow encoding with the general-exchange flags (widths above 64 clamped: no
// narrow window's cells reach them), appended to `alt`.  Uses plan_job's
// L, R, XL.  Returns the copy's offset in `alt`, or -1 without one.
int64_t window_classes(const HostJob &hj, const PlanScratch &sc, const std::vector<DevWin> &wj, int64_t padEnd,
                       std::vector<int8_t> &wcls, std::vector<int32_t> &wwm, std::vector<int32_t> &alt) {
    const int64_t D = hj.lX + hj.lY;
    bool mixed = false;
    for (size_t k = 0; k < wj.size(); k++) {
        const int64_t a = std::max<int64_t>(0, window_fstart(k, wj[k].lo) - 2);
        const int64_t e = std::min<int64_t>(D, (int64_t)wj[k].tb + 2);
        int32_t m = 1;
        for (int64_t d = a; d <= e; d++) m = std::max(m, (sc.R[d] - sc.L[d]) / 2 + 1);
        const int8_t c = m <= 32 ? 1 : (m <= 64 ? 2 : 3);
        wcls.push_back(c);
        wwm.push_back(m);
        mixed |= c < 3;
    }
    if (!mixed) return -1;
    const int64_t off = (int64_t)alt.size();
    alt.resize(alt.size() + (size_t)padEnd);
    int32_t *bj = alt.data() + off;
    for (int64_t d = 0; d <= D; d++) {
        const int32_t wd = std::min((sc.R[d] - sc.L[d]) / 2 + 1, 64);
        bj[d] = (int32_t)((uint32_t)sc.XL[d] << kBandShift) | (wd - 1);
    }
    for (int64_t d = D + 1; d < padEnd; d++) bj[d] = bj[D];
    mark_slow(sc.XL.data(), D, bj);
    return off;
}


}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" mp_status mp_engine_create(int32_t device, mp_engine **out) {
    if (!out) {
        mp::set_error("mp_engine_create: null argument");
        return MP_EINVAL;
    }
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) {
        mp::set_error("no HIP device " + std::to_string(device) + " visible");
        return MP_ENODEVICE;
    }
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos) {
        mp::set_error(std::string("device is ") + prop.gcnArchName + ", this library is built for gfx950");
        return MP_ENODEVICE;
    }
    HIP_TRY(hipFuncSetAttribute((const void *)pairhmm_split_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                kSplitLdsBytes));
    HIP_TRY(hipFuncSetAttribute((const void *)pairhmm_seq_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                kSeqLdsBytesExp));
    HIP_TRY(hipFuncSetAttribute((const void *)pairhmm_wide_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                kWideLdsLimit));
    HIP_TRY(hipFuncSetAttribute((const void *)pairhmm_wide5_kernel<float>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, wide5_lds_bytes<float>(kWide5MaxC32)));
    HIP_TRY(hipFuncSetAttribute((const void *)pairhmm_wide5_kernel<double>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, wide5_lds_bytes<double>(kWide5MaxC64)));
    int bpc = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, (const void *)pairhmm_split_kernel, 512,
                                                         kSplitLdsBytes));
    void *scratchPtr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&scratchPtr, HIP_SYMBOL(scratch)));
    mp_engine *e = new mp_engine();
    e->device = device;
    e->scratch = static_cast<unsigned char *>(scratchPtr);
    e->cus = prop.multiProcessorCount;
    e->splitBlocksPerCU = std::max(1, bpc);
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess ||
        hipEventCreate(&e->ev2) != hipSuccess) {
        delete e;
        mp::set_error("cannot create HIP stream/events");
        return MP_EHIP;
    }
    *out = e;
    return MP_OK;
}

extern "C" void mp_engine_free(mp_engine *e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    for (int k = 0; k < 2; k++) {
        if (e->stage[k]) (void)hipHostFree(e->stage[k]);
        if (e->stageEv[k]) (void)hipEventDestroy(e->stageEv[k]);
    }
    for (int l = 0; l < 3; l++)
        if (e->listHost[l]) (void)hipHostFree(e->listHost[l]);
    for (int k = 0; k < 2; k++)
        if (e->upHost[k]) (void)hipHostFree(e->upHost[k]);
    if (e->ev0) (void)hipEventDestroy(e->ev0);
    if (e->ev1) (void)hipEventDestroy(e->ev1);
    if (e->ev2) (void)hipEventDestroy(e->ev2);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

extern "C" int32_t mp_engine_cu_count(const mp_engine *e) { return e ? e->cus : 0; }

// A batch's symbol pool from its raw entries (raw_sym) to the emission-table
// offsets the kernels read (enc_x / enc_y): rleString_constructSymbolString's
// symbols (stateMachine.c:124-146) as table addresses, on the device.
__global__ __launch_bounds__(256) void encode_symbols(const uint16_t *raw, uint32_t *sym, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        sym[i] = enc_raw(raw[i]);
}

// The engine's pinned upload staging k, at least `bytes` long (grown with a
// quarter of headroom); the caller holds e->createMu.
static mp_status upload_stage(mp_engine *e, int k, size_t bytes, void **out) {
    if (e->upHostBytes[k] < bytes) {
        if (e->upHost[k]) (void)hipHostFree(e->upHost[k]);
        e->upHost[k] = nullptr;
        e->upHostBytes[k] = 0;
        const size_t want = bytes + bytes / 4 + ((size_t)1 << 20);
        if (hipHostMalloc(&e->upHost[k], want, hipHostMallocDefault) != hipSuccess) {
            e->upHost[k] = nullptr;
            mp::set_error("mp_batch_create: pinned host allocation of the upload staging failed");
            return MP_ENOMEM;
        }
        e->upHostBytes[k] = want;
    }
    *out = e->upHost[k];
    return MP_OK;
}

static void discard_batch(mp_batch *b, mp_batch **inFlight) {
    *inFlight = nullptr;
    // (the symbol upload may still read the engine's staging and write the batch's buffers)
    if (b->eng && b->eng->stream) (void)hipStreamSynchronize(b->eng->stream);
    delete b;
}

static mp_status batch_create(mp_engine *e, const mp_model *m, const mp_pair *pairs, int64_t npairs,
                               const mp_band_params *p, int32_t mode, int32_t flags, mp_batch **out,
                               mp_batch **inFlight) {
    if (!e || !m || !p || !out || (npairs > 0 && !pairs) || npairs < 0) {
        mp::set_error("mp_batch_create: null or negative argument");
        return MP_EINVAL;
    }
    if (mode < MP_MODE_POSTERIORS || mode > MP_MODE_EXPECTATIONS) {
        mp::set_error("mp_batch_create: unknown mode");
        return MP_EINVAL;
    }
    if (p->traceBackDiagonals < 1 || p->diagonalExpansion < 0 || (p->diagonalExpansion & 1) ||
        p->minDiagsBetweenTraceBack < 2 || p->traceBackDiagonals + 1 >= p->minDiagsBetweenTraceBack) {
        mp::set_error("mp_batch_create: band parameters violate the reference prerequisites "
                      "(pairwiseAligner.c:692-696)");
        return MP_EINVAL;
    }
    const bool fp64 = (flags & MP_BATCH_FP64) != 0;
    const bool refArith = (flags & MP_BATCH_REFERENCE_ARITHMETIC) != 0;
    if (refArith && (fp64 || m->mach[1].nStates != 3 || !use_split(mode))) {
        mp::set_error("mp_batch_create: reference arithmetic runs the three-state machine in the posterior and "
                      "match-only modes");
        return MP_EUNSUPPORTED;
    }
    const bool five = m->mach[1].nStates == 5 || fp64;  // every job on the five-state kernel
    if (five && mode == MP_MODE_EXPECTATIONS) {
        mp::set_error("mp_batch_create: expected transition counts are computed for three-state machines in FP32 "
                      "only");
        return MP_EUNSUPPORTED;
    }
    *out = nullptr;
    HIP_TRY(hipSetDevice(e->device));
    std::lock_guard<std::mutex> createLock(e->createMu);
    mp_batch *b = new mp_batch();
    *inFlight = b;
    b->eng = e;
    b->mode = mode;
    b->flags = flags;
    b->npairs = npairs;
    {
        // 64-bit list entries when x + 1 < 2^bx and y + 1 < 2^by for every pair (list
        // coordinates run from -1 to the length - 1) and 24 + bx + by <= 64
        int64_t mx = 1, my = 1;
        for (int64_t i = 0; i < npairs; i++) {
            mx = std::max(mx, pairs[i].lX + 1);
            my = std::max(my, pairs[i].lY + 1);
        }
        auto bits = [](int64_t v) {  // bits of the values 0 .. v - 1
            int n = 1;
            while (n < 62 && ((int64_t)1 << n) < v) n++;
            return n;
        };
        const int bx = bits(mx), by = bits(my);
        b->packBits = 24 + bx + by <= 64 ? bx : 0;
    }
    b->pairCrop.assign(npairs, 0);
    b->pairStrand.resize(npairs);
    for (int64_t i = 0; i < npairs; i++) b->pairStrand[i] = pairs[i].strand ? 1 : 0;
    b->pairNJobs.assign(npairs, 0);
    b->pairFirstJob.assign(npairs, 0);
    std::vector<int64_t> pairFirstSub(npairs, 0);  // a pair's first sub-problem in `jobs`
    const bool split = use_split(mode);
    auto tPhase = std::chrono::steady_clock::now();
    auto mark = [&](int ph) {
        const auto t = std::chrono::steady_clock::now();
        b->prepMs[ph] += std::chrono::duration<double, std::milli>(t - tPhase).count();
        tPhase = t;
    };

    // ---- 1. encoded symbols, cropping, split points -> jobs (pair order) ----
    // Two parallel passes over the pairs: (A) checks, reference crop, split
    // points and sizes; (B) after prefix sums, symbol encoding into the pool and
    // the jobs with their anchors.  The first failing pair (in order) is reported.
    const int nthreads = (int)std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    // items claimed in chunks of n / (64 threads) (16 to 4096): few contended claims on
    // hundreds of thousands of small items, balance on thousands of large ones; no more
    // threads than chunks
    auto parallel_for = [&](int64_t n, const std::function<void(int64_t)> &fn) {
        const int64_t chunk = std::min<int64_t>(4096, std::max<int64_t>(16, n / (64 * (int64_t)nthreads)));
        const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(nthreads, (n + chunk - 1) / chunk));
        std::atomic<int64_t> next{0};
        mp::parallel_run(nt, [&](int) {
            for (;;) {
                const int64_t i0 = next.fetch_add(chunk);
                if (i0 >= n) break;
                for (int64_t i = i0; i < std::min(n, i0 + chunk); i++) fn(i);
            }
        });
    };
    struct PairPrep {
        int64_t first = 0, end = 0;
        int64_t one[4];             // the split tuple when there is exactly one (no allocation)
        std::vector<int64_t> more;  // split points (x1, y1, x2, y2) when there are several
        int32_t nsp = 0;            // number of sub-problems
        int err = 0;                // 1 bad arguments, 2 bad anchors, 3 sub-problem too long
        const int64_t *sp() const { return more.empty() ? one : more.data(); }
    };
    std::vector<PairPrep> prep(npairs);
    parallel_for(npairs, [&](int64_t i) {
        const mp_pair &pr = pairs[i];
        PairPrep &pp = prep[i];
        if (pr.lX < 0 || pr.lY < 0 || (pr.lX > 0 && !pr.sX) || (pr.lY > 0 && !pr.sY) ||
            (pr.nAnchors > 0 && !pr.anchors) || pr.nAnchors < 0 || pr.lX > INT32_MAX / 2 || pr.lY > INT32_MAX / 2) {
            pp.err = 1;
            return;
        }
        if (!mp::anchors_valid(pr.anchors, pr.nAnchors, pr.lX, pr.lY)) {
            pp.err = 2;
            return;
        }
        pp.first = 0;
        pp.end = pr.lX;
        if ((flags & MP_BATCH_CROP_REFERENCE) && pr.nAnchors > 0) {
            const int64_t *a0 = pr.anchors, *al = pr.anchors + 3 * (pr.nAnchors - 1);
            pp.first = std::max<int64_t>(a0[0] - a0[1], 0);
            pp.end = std::min<int64_t>(1 + al[0] + (pr.lY - al[1]), pr.lX);
        }
        const int64_t lX = pp.end - pp.first, lY = pr.lY;
        if (mode == MP_MODE_FORWARD_ONLY) {  // computeForwardProbability does not split
            pp.one[0] = pp.one[1] = 0;
            pp.one[2] = lX;
            pp.one[3] = lY;
            pp.nsp = 1;
        } else {
            mp::split_points(pr.anchors, pr.nAnchors, pp.first, lX, lY, p->splitMatrixBiggerThanThis,
                             pr.raggedLeft != 0,
                             pr.raggedRight != 0, pp.more);
            pp.nsp = (int32_t)(pp.more.size() / 4);
            if (pp.nsp == 1) {
                std::copy(pp.more.begin(), pp.more.end(), pp.one);
                std::vector<int64_t>().swap(pp.more);
            }
        }
        const int64_t *spp = pp.sp();
        for (int32_t s = 0; s < pp.nsp; s++)
            if ((spp[4 * s + 2] - spp[4 * s]) + (spp[4 * s + 3] - spp[4 * s + 1]) >= kMaxDiagonals) pp.err = 3;
    });
    for (int64_t i = 0; i < npairs; i++) {
        if (!prep[i].err) continue;
        discard_batch(b, inFlight);
        if (prep[i].err == 1) {
            mp::set_error("pair " + std::to_string(i) + ": bad sequence or anchor arguments");
            return MP_EINVAL;
        }
        if (prep[i].err == 2) {
            mp::set_error("pair " + std::to_string(i) + ": anchors must be strictly increasing, inside the "
                          "sequences, with even non-negative expansion");
            return MP_EANCHORS;
        }
        mp::set_error("pair " + std::to_string(i) + ": a sub-problem has more than " +
                      std::to_string(kMaxDiagonals - 1) + " anti-diagonals");
        return MP_EUNSUPPORTED;
    }
    mark(0);
    // Symbol pool with padding at both ends (kernels read up to one row width
    // beyond a sequence without clamping).  Each sequence is preceded by an N
    // entry of its own kind: the cells x = 0 (y = 0) of the band read the symbol
    // before the sequence (stateMachine emissions there are never used,
    // pairwiseAligner.c:528-549, but must be finite).
    std::vector<int64_t> symOff(npairs + 1, kSymPad);
    int64_t njobs = 0;
    for (int64_t i = 0; i < npairs; i++) {
        symOff[i + 1] = symOff[i] + 2 + (prep[i].end - prep[i].first) + pairs[i].lY;
        pairFirstSub[i] = njobs;
        b->pairCrop[i] = prep[i].first;
        njobs += prep[i].nsp;
    }
    if (symOff[npairs] + kSymPad >= ((int64_t)1 << 30)) {
        discard_batch(b, inFlight);
        mp::set_error("batch too large: more than 2^30 symbols (split it into several batches)");
        return MP_EUNSUPPORTED;
    }
    // every entry is written below (pairs in parallel, padding here): no value-initialisation pass
    const size_t nsym = (size_t)(symOff[npairs] + kSymPad);
    uint16_t *sym = nullptr;  // raw entries, encoded on the device (encode_symbols)
    {
        void *st = nullptr;
        const mp_status ss = upload_stage(e, 0, nsym * sizeof(uint16_t), &st);
        if (ss) {
            discard_batch(b, inFlight);
            return ss;
        }
        sym = static_cast<uint16_t *>(st);
    }
    for (int64_t i = 0; i < kSymPad; i++) sym[i] = raw_sym(4, 0, 0);
    for (size_t i = (size_t)symOff[npairs]; i < nsym; i++) sym[i] = raw_sym(4, 0, 0);
    sym[0] = kRawZero;  // read by cells outside the band
    std::vector<HostJob> jobs((size_t)njobs);
    std::vector<JobAnc> jobAnchors((size_t)njobs);
    parallel_for(npairs, [&](int64_t i) {
        const mp_pair &pr = pairs[i];
        const PairPrep &pp = prep[i];
        const int strand = pr.strand ? 1 : 0;
        uint16_t *o = sym + symOff[i];
        *o++ = raw_sym(4, 0, strand);
        const int64_t baseX = symOff[i] + 1;
        for (int64_t x = pp.first; x < pp.end; x++) *o++ = raw_sym(pr.sX[x], 0, strand);
        *o++ = raw_sym(4, 1, strand);
        const int64_t baseY = baseX + (pp.end - pp.first) + 1;
        for (int64_t y = 0; y < pr.lY; y++) *o++ = raw_sym(pr.sY[y], 1, strand);
        const int64_t ns = pp.nsp;
        const int64_t *spp = pp.sp();
        int64_t j = 0;
        for (int64_t s = 0; s < ns; s++) {
            const int64_t x1 = spp[4 * s], y1 = spp[4 * s + 1], x2 = spp[4 * s + 2], y2 = spp[4 * s + 3];
            const int64_t j0 = j;
            while (j < pr.nAnchors && (pr.anchors[3 * j] - pp.first) + pr.anchors[3 * j + 1] < x2 + y2) j++;
            jobAnchors[pairFirstSub[i] + s] = JobAnc{pr.anchors + 3 * j0, j - j0, pp.first + x1, y1};
            HostJob &hj = jobs[pairFirstSub[i] + s];
            hj.pair = i;
            hj.xOff = x1;
            hj.yOff = y1;
            hj.symX = baseX + x1;
            hj.symY = baseY + y1;
            hj.lX = (int32_t)(x2 - x1);
            hj.lY = (int32_t)(y2 - y1);
            const bool rl = pr.raggedLeft || s > 0, rr = pr.raggedRight || s < ns - 1;
            hj.flags = strand | (rl ? 2 : 0) | (rr ? 4 : 0);
        }
    });

    mark(1);
    if (refArith) {
        mp_status st = create_reference_batch(b, m, pairs, p, symOff, pairFirstSub, jobs, jobAnchors);
        if (st) {
            discard_batch(b, inFlight);
            return st;
        }
        *out = b;
        return MP_OK;
    }
    // the symbol pool is final: its DMA and the device encoding run while the bands are planned
    hipStream_t s = e->stream;
    DevBuf<uint16_t> rawUp;  // the raw pool on the device until encode_symbols has run (freed on return)
    {
        mp_status st;
        if ((st = rawUp.upload(sym, nsym, s)) || (st = b->dSym.alloc(nsym))) {
            discard_batch(b, inFlight);
            return st;
        }
        const unsigned g = (unsigned)std::min<size_t>((nsym + 255) / 256, (size_t)e->cus * 16);
        hipLaunchKernelGGL(encode_symbols, dim3(std::max(g, 1u)), dim3(256), 0, s, rawUp.p, b->dSym.p, (int64_t)nsym);
        if (hipGetLastError() != hipSuccess) {
            discard_batch(b, inFlight);
            mp::set_error("mp_batch_create: symbol encoding launch failed");
            return MP_EHIP;
        }
    }

    // ---- 2. bands, windows, classes (parallel over jobs) ----
    const int64_t nj = (int64_t)jobs.size();
    std::vector<int64_t> bandOff(nj + 1, 0);
    // Each job's band padded to whole cells of kK anti-diagonals plus one extra cell.
    for (int64_t j = 0; j < nj; j++) bandOff[j + 1] = bandOff[j] + ((jobs[j].lX + jobs[j].lY) / kK + 2) * kK;
    if (bandOff[nj] >= ((int64_t)1 << 30)) {
        discard_batch(b, inFlight);
        mp::set_error("batch too large: more than 2^30 anti-diagonals (split it into several batches)");
        return MP_EUNSUPPORTED;
    }
    // every entry of a job's region is written by its band pass below (anti-diagonals 0..D, then padding)
    int32_t *band = nullptr;
    {
        void *st = nullptr;
        const mp_status ss = upload_stage(e, 1, (size_t)std::max<int64_t>(bandOff[nj], 1) * sizeof(int32_t), &st);
        if (ss) {
            discard_batch(b, inFlight);
            return ss;
        }
        band = static_cast<int32_t *>(st);
    }
    // each job's windows live in its thread's pool: (thread, offset, count) per job,
    // with each window's class and width maximum in parallel pools; a wide job whose
    // windows run as window jobs may keep a narrow copy of its band in the thread's
    // band pool (alt: offset there, or -1)
    struct WinRef {
        int32_t t, n;
        int64_t off;
        int64_t alt = -1;
    };
    std::vector<WinRef> winRef(nj);
    const int nthW = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    std::vector<std::vector<DevWin>> winPool(nthW);
    std::vector<std::vector<int8_t>> clsPool(nthW);
    std::vector<std::vector<int32_t>> wmPool(nthW), altPool(nthW), exPool(nthW);
    std::vector<std::vector<int64_t>> rcPool(nthW);  // wide kernel ring cells of each window
    const bool winJobs = split && (flags & MP_BATCH_EXACT_CHAIN) == 0 && !five;
    std::atomic<int> err{0};
    {
        std::atomic<int64_t> next{0};
        // jobs claimed in runs of nj / (256 threads), 1 to 32 (one shared counter): long
        // sub-problems balance, hundreds of thousands of short ones claim rarely
        const int64_t runLen = std::min<int64_t>(32, std::max<int64_t>(1, nj / (256 * (int64_t)nthW)));
        mp::parallel_run(nthW, [&](int t) {
                PlanScratch sc;
                std::vector<DevWin> wj;
                std::vector<DevWin> &pool = winPool[t];
                for (int64_t j = 0, jEnd = 0;; j++) {
                    if (j >= jEnd) {
                        j = next.fetch_add(runLen);
                        jEnd = std::min<int64_t>(j + runLen, nj);
                    }
                    if (j >= nj) break;
                    HostJob &hj = jobs[j];
                    hj.bandOff = bandOff[j];
                    if (!plan_job(hj, jobAnchors[j], *p, mode, five, band + bandOff[j], bandOff[j + 1] - bandOff[j],
                                  sc, wj)) {
                        err = 1;
                        continue;
                    }
                    winRef[j] = WinRef{t, (int32_t)wj.size(), (int64_t)pool.size()};
                    pool.insert(pool.end(), wj.begin(), wj.end());
                    // the wide kernel's ring holds a window's forward values of [lo, tb], one row
                    // per anti-diagonal, plus one row lost where the ring wraps
                    for (size_t k = 0; k < wj.size(); k++) rcPool[t].push_back(sc.wcells[k] + hj.wmax);
                    // excess band width over each window's posterior anti-diagonals (lo, from]
                    for (size_t k = 0; k < wj.size(); k++)
                        exPool[t].push_back(five ? 0 : (int32_t)std::min<int64_t>(sc.wex[k], INT32_MAX / 8));
                    if (winJobs && hj.cls == 3 && wj.size() > 1) {
                        winRef[j].alt = window_classes(hj, sc, wj, bandOff[j + 1] - bandOff[j], clsPool[t], wmPool[t],
                                                       altPool[t]);
                    } else {
                        clsPool[t].insert(clsPool[t].end(), wj.size(), (int8_t)hj.cls);
                        wmPool[t].insert(wmPool[t].end(), wj.size(), hj.wmax);
                    }
                }
            });
    }
    // the narrow band copies follow the jobs' bands in the device band array
    std::vector<int64_t> altBase(nthW + 1, bandOff[nj]);
    for (int t = 0; t < nthW; t++) altBase[t + 1] = altBase[t] + (int64_t)altPool[t].size();
    const int64_t bandTotal = altBase[nthW];
    if (bandTotal >= ((int64_t)1 << 30)) {
        discard_batch(b, inFlight);
        mp::set_error("batch too large: more than 2^30 anti-diagonals (split it into several batches)");
        return MP_EUNSUPPORTED;
    }
    if (err) {
        discard_batch(b, inFlight);
        mp::set_error("invalid band (anchors inconsistent with the sequences)");
        return MP_EANCHORS;
    }
    const int32_t wideLimit = 64 * (fp64 ? kWide5MaxC64 : (five ? kWide5MaxC32 : kWideMaxC));
    for (int64_t j = 0; j < nj; j++) {
        if (jobs[j].wmax > wideLimit) {
            discard_batch(b, inFlight);
            mp::set_error("pair " + std::to_string(jobs[j].pair) + ": band width " + std::to_string(jobs[j].wmax) +
                          " exceeds the " + std::to_string(wideLimit) + "-cell kernel limit" +
                          (fp64 ? " of FP64 batches" : (five ? " of the five-state machine" : "")));
            return MP_EUNSUPPORTED;
        }
    }
    {
        // the bands are final: the jobs' bands, then the narrow band copies of the threads'
        // pools, go to the device while the rest is planned
        hipError_t he = hipSuccess;
        if (b->dBand.alloc((size_t)bandTotal)) {
            discard_batch(b, inFlight);
            return MP_ENOMEM;
        }
        if (bandOff[nj] > 0)
            he = hipMemcpyAsync(b->dBand.p, band, sizeof(int32_t) * (size_t)bandOff[nj], hipMemcpyHostToDevice, s);
        for (int t = 0; t < nthW && he == hipSuccess; t++)
            if (!altPool[t].empty())
                he = hipMemcpyAsync(b->dBand.p + altBase[t], altPool[t].data(), sizeof(int32_t) * altPool[t].size(),
                                    hipMemcpyHostToDevice, s);
        if (he != hipSuccess) {
            discard_batch(b, inFlight);
            mp::set_error(std::string("band upload: ") + hipGetErrorString(he));
            return MP_EHIP;
        }
    }

    mark(2);
    // ---- 2b. window jobs (kWarmup): in the posterior modes a three-state sub-problem
    // with several traceback windows runs each window as a job of its own (narrow
    // ones on the split kernel, wide ones on the wide-band kernel) ----
    int64_t subCells = 0;
    for (int64_t j = 0; j < nj; j++) subCells += jobs[j].cells;
    std::vector<int64_t> subFirstX(nj + 1, 0);
    {
        int64_t nx = 0;
        for (int64_t j = 0; j < nj; j++) {
            subFirstX[j] = nx;
            nx += (winJobs && winRef[j].n > 1) ? winRef[j].n : 1;
        }
        subFirstX[nj] = nx;
        if (nx >= ((int64_t)1 << 30)) {
            discard_batch(b, inFlight);
            mp::set_error("batch too large: more than 2^30 jobs (split it into several batches)");
            return MP_EUNSUPPORTED;
        }
        if (nx != nj) {
            std::vector<HostJob> xj((size_t)nx);
            std::vector<WinRef> xw((size_t)nx);
            parallel_for(nj, [&](int64_t j) {
                const int64_t x0 = subFirstX[j], n = subFirstX[j + 1] - x0;
                if (n == 1) {
                    xj[x0] = jobs[j];
                    xw[x0] = winRef[j];
                    return;
                }
                const WinRef &wr = winRef[j];
                DevWin *wv = winPool[wr.t].data() + wr.off;
                for (int64_t k = 0; k < n; k++) {
                    DevWin &w = wv[k];
                    HostJob h = jobs[j];
                    // the window's own class: a narrow window of a wide job reads the narrow copy
                    h.cls = clsPool[wr.t][wr.off + k];
                    h.wmax = wmPool[wr.t][wr.off + k];
                    const bool wide = h.cls == 3;
                    if (!wide && jobs[j].cls == 3) h.bandOff = altBase[wr.t] + wr.alt;
                    const int32_t *bj = h.bandOff < bandOff[nj] ? band + h.bandOff
                                                                 : altPool[wr.t].data() + (h.bandOff - altBase[wr.t]);
                    const int64_t fs = window_fstart((size_t)k, w.lo);
                    h.fstart = (int32_t)fs;
                    if (fs > 0) h.flags |= kJobRestart;
                    if (!wide) {
                        // crossing: the forward row (from fs) and the backward row (from tb) meet
                        // halfway, at 7 (mod 8), inside [lo, tb - 1]
                        int64_t c = (fs + w.tb - 1) / 2;
                        c = ((c + 4) & ~(int64_t)(kK - 1)) - 1;
                        w.cx = (int32_t)std::min<int64_t>(std::max<int64_t>(c, w.lo), w.tb - 1);
                    }
                    int64_t cells = 0;  // the job's work, for the device order
                    for (int64_t d = fs; d <= w.tb; d++)
                        cells += (wide ? band_wm1_wide(bj[d]) : band_wm1_narrow(bj[d])) + 1;
                    h.cells = cells;
                    // live ring entries: narrow, forward and backward values of (lo, tb]; wide,
                    // forward values of [lo, tb] (the backward runs after the forward)
                    h.ringNeed = (int32_t)((w.tb - w.lo + (wide ? 2 + kK - 1 : 4 * kK - 1)) / kK * kK);
                    xj[x0 + k] = h;
                    xw[x0 + k] = WinRef{wr.t, 1, wr.off + k};
                }
            });
            jobs.swap(xj);
            winRef.swap(xw);
        }
    }
    const int64_t nxj = (int64_t)jobs.size();

    mark(3);
    // ---- 3. device order: by class, then band area descending ----
    // (class ascending, cells descending, job index ascending): a stable LSD
    // radix sort of the jobs on (class, maxCells - cells)
    std::vector<int32_t> ord;
    {
        int64_t maxCells = 0;
        for (int64_t j = 0; j < nxj; j++)
            if (jobs[j].lX + jobs[j].lY > 0) {
                ord.push_back((int32_t)j);
                maxCells = std::max(maxCells, jobs[j].cells);
            }
        int bitsC = 1;
        while (bitsC < 60 && ((int64_t)1 << bitsC) <= maxCells) bitsC++;
        std::vector<uint64_t> key((size_t)nxj);
        for (int32_t j : ord) key[j] = ((uint64_t)jobs[j].cls << bitsC) | (uint64_t)(maxCells - jobs[j].cells);
        std::vector<int32_t> tmp(ord.size());
        std::vector<size_t> cnt(2049);
        for (int sh = 0; sh < bitsC + 2; sh += 11) {
            std::fill(cnt.begin(), cnt.end(), 0);
            for (int32_t j : ord) cnt[((key[j] >> sh) & 2047) + 1]++;
            for (int k = 0; k < 2048; k++) cnt[k + 1] += cnt[k];
            for (int32_t j : ord) tmp[cnt[(key[j] >> sh) & 2047]++] = j;
            ord.swap(tmp);
        }
    }
    const int64_t nq = (int64_t)ord.size();
    std::vector<int32_t> devIndex(nxj, -1);
    std::vector<DevJob> dj((size_t)nq);
    std::vector<DevWin> &dw = b->winHost;
    const bool post = mode <= MP_MODE_MATCH_ONLY;
    // record capacities (1.5 per posterior anti-diagonal of each side; exact re-run on overflow),
    // gathered per device slot so the offset pass below reads memory in order
    struct Slot {
        int64_t cells, out, ringCells;
        int32_t nWin, cls, ringNeed, wmax;
    };
    std::vector<Slot> slot((size_t)nq);
    parallel_for(nq, [&](int64_t q) {
        HostJob &hj = jobs[ord[q]];
        const WinRef &wr = winRef[ord[q]];
        const DevWin *wv = winPool[wr.t].data() + wr.off;
        int64_t nF = 0, nB = 0;
        // where a three-state band is wider than the anchors' own 2 W + 1 cells the anchors are
        // sparse and the posteriors spread over several alignments: a window's excess width (at
        // most 16 per anti-diagonal) is added to each side that holds its records
        for (int32_t k = 0; k < wr.n; k++) {
            const DevWin &w = wv[k];
            const int64_t ex = exPool[wr.t][wr.off + k];
            if (hj.cls < 3) {
                nF += std::max(0, w.from - w.cx) + ex;
                nB += std::max(0, std::min(w.cx, w.from) - w.lo) + ex;
            } else {
                nB += w.from - w.lo + ex;
            }
        }
        // (measured on 100 kb chunks: 0.9 records per posterior anti-diagonal; a job that writes
        // more than its region holds makes the run size every region exactly and run again)
        hj.cap[0] = post && hj.cls < 3 ? (int32_t)std::min<int64_t>(nF + nF / 2 + 64, INT32_MAX / 4) : 0;
        hj.cap[1] = post ? (int32_t)std::min<int64_t>(nB + nB / 2 + 64, INT32_MAX / 4) : 0;
        hj.nWin = wr.n;
        int64_t rc = 0;
        if (hj.cls == 3)
            for (int32_t k = 0; k < wr.n; k++) rc = std::max(rc, rcPool[wr.t][wr.off + k]);
        slot[q] = Slot{hj.cells, (int64_t)hj.cap[0] + hj.cap[1], rc, wr.n, hj.cls, hj.ringNeed, hj.wmax};
    });
    // offsets in device order, batch-wide maxima, class ranges
    std::vector<int64_t> slotOut((size_t)nq), slotWin((size_t)nq);
    int64_t outOff = 0, winTotal = 0;
    int32_t ringLen = 4 * kK;
    for (int c = 0; c < kNumClasses; c++) b->classBegin[c] = b->classEnd[c] = (int32_t)nq;
    b->wideBegin = b->wideEnd = (int32_t)nq;
    for (int64_t q = 0; q < nq; q++) {
        const Slot &sl = slot[q];
        slotOut[q] = outOff;
        outOff += sl.out;
        slotWin[q] = winTotal;
        winTotal += sl.nWin;
        if (mode != MP_MODE_FORWARD_ONLY && sl.cls < 3) ringLen = std::max(ringLen, sl.ringNeed);
        if (sl.cls == 3) {
            b->wideC = std::max(b->wideC, (sl.wmax + 63) / 64);
            if (mode != MP_MODE_FORWARD_ONLY) {
                b->ringWideLen = std::max(b->ringWideLen, sl.ringNeed);
                b->ringWideCells = std::max(b->ringWideCells, sl.ringCells);
            }
            b->wideBegin = std::min<int32_t>(b->wideBegin, (int32_t)q);
            b->wideEnd = (int32_t)q + 1;
        } else {
            b->classBegin[sl.cls] = std::min<int32_t>(b->classBegin[sl.cls], (int32_t)q);
            b->classEnd[sl.cls] = (int32_t)q + 1;
        }
    }
    b->cells = subCells;
    for (int c = 0; c < kNumClasses; c++)
        if (b->classBegin[c] > b->classEnd[c]) b->classBegin[c] = b->classEnd[c];
    dw.resize((size_t)winTotal);
    b->jobs.resize((size_t)nq);
    parallel_for(nq, [&](int64_t q) {
        HostJob &hj = jobs[ord[q]];
        hj.outOff = slotOut[q];
        hj.winOff = (int32_t)slotWin[q];
        devIndex[ord[q]] = (int32_t)q;
        const WinRef &wr = winRef[ord[q]];
        std::copy(winPool[wr.t].data() + wr.off, winPool[wr.t].data() + wr.off + wr.n, dw.data() + hj.winOff);
        DevJob &d = dj[q];
        d.symX = hj.symX; d.symY = hj.symY; d.bandOff = hj.bandOff; d.outOff = hj.outOff;
        d.lX = hj.lX; d.lY = hj.lY; d.winOff = hj.winOff; d.nWin = hj.nWin;
        d.cap[0] = hj.cap[0]; d.cap[1] = hj.cap[1];
        d.flags = hj.flags;
        d.fstart = hj.fstart;
        b->jobs[q] = hj;
    });
    mark(4);
    // ---- 3b. slots: per pair its sub-problems in split order, a sub-problem's window
    // jobs last window first (the order of the reference's lists) ----
    {
        std::vector<int32_t> slots, slotK;
        slots.reserve((size_t)nxj);
        slotK.reserve((size_t)nxj);
        for (int64_t i = 0; i < npairs; i++) {
            const int64_t f = (int64_t)slots.size();
            for (int32_t sp = 0; sp < prep[i].nsp; sp++) {
                const int64_t j = pairFirstSub[i] + sp;
                const int64_t x0 = subFirstX[j], n = subFirstX[j + 1] - x0;
                for (int64_t k = n - 1; k >= 0; k--) {
                    slots.push_back(devIndex[x0 + k]);
                    slotK.push_back(k == n - 1 ? (int32_t)n : 0);
                }
            }
            b->pairFirstJob[i] = f;
            b->pairNJobs[i] = (int32_t)(slots.size() - f);
        }
        b->order.swap(slots);
        b->slotK.swap(slotK);
    }
    b->ringLen = ringLen;
    b->outTotal = outOff;

    mark(5);
    // ---- 4. upload ----
    std::vector<float> tab;
    KParams &kp = b->kp;
    std::memset(&kp, 0, sizeof(kp));
    build_tables(m, fp64, tab, kp);
    b->modelStates = m->mach[1].nStates;
    const size_t nwin = std::max<size_t>(dw.size(), 1);
    mp_status st;
    if ((st = b->dJobs.upload(dj, s)) ||
        (st = b->dWin.upload(dw, s)) || (st = b->dTables.upload(tab, s)) ||
        (st = b->dOut.alloc(std::max<int64_t>(outOff, 1))) || (st = b->dCounts.alloc(std::max<size_t>(2 * dj.size(), 1))) ||
        (st = b->dWinEnd.alloc(2 * nwin)) || (st = b->dLogTotal.alloc(std::max<size_t>(dj.size(), 1))) ||
        (st = b->dChain.alloc(std::max<size_t>(dj.size(), 1))) || (st = b->dSlotJob.upload(b->order, s)) ||
        (st = b->dSlotK.upload(b->slotK, s)) || (st = b->dSlotBase.alloc(b->order.size() + 1)) ||
        (st = b->dOverflow.alloc(1)) || (st = b->dStarts.alloc(3 * ((size_t)npairs + 1))) ||
        (st = b->dPairLL.alloc(std::max<size_t>((size_t)npairs, 1))) || (st = b->dBatchLL.alloc(1)) ||
        (st = b->dQueue.alloc(kNumClasses + 1)) || (st = b->dError.alloc(1)) ||
        (st = b->dExpect.alloc(mode == MP_MODE_EXPECTATIONS ? std::max<size_t>(10 * dj.size(), 1) : 1)) ||
        (st = b->dEmit.alloc(mode == MP_MODE_EXPECTATIONS ? kTabFloats : 1))) {
        discard_batch(b, inFlight);
        return st;
    }
    // Rings: one slice per resident wavefront.
    const bool anyNarrow = b->classEnd[0] > b->classBegin[0] || b->classEnd[1] > b->classBegin[1] ||
                           b->classEnd[2] > b->classBegin[2];
    b->narrowWaves = e->cus * (split ? e->splitBlocksPerCU : 1) * 8;
    b->narrowBlocks = b->narrowWaves * 64 / b->narrowThreads;
    {
        // batch shape key for the autotuner's cache: mode, kernel family, dominant narrow class, widest
        // band in columns, log2 of the job count
        int domCls = 0, domN = -1;
        for (int c = 0; c < kNumClasses; c++)
            if (b->classEnd[c] - b->classBegin[c] > domN) {
                domN = b->classEnd[c] - b->classBegin[c];
                domCls = c;
            }
        int lg = 0;
        while ((1ll << lg) < (int64_t)ord.size() && lg < 40) lg++;
        b->tuneKey = (uint64_t)mode | ((uint64_t)(five ? 1 : 0) << 2) | ((uint64_t)(fp64 ? 1 : 0) << 3) |
                     ((uint64_t)domCls << 4) | ((uint64_t)std::min(b->wideC, 63) << 6) | ((uint64_t)lg << 12);
        auto it = e->tuned.find(b->tuneKey);
        if (it != e->tuned.end()) {
            b->narrowThreads = it->second.first;
            b->narrowBlocks = b->narrowWaves * 64 / b->narrowThreads;
            b->wide5Waves = it->second.second;
        }
    }
    if (anyNarrow) {
        const size_t waves = (size_t)b->narrowWaves;
        if (split) {
            if ((st = b->dRing.alloc(waves * (size_t)(ringLen / kK) * (kRingCellBytes / 4))) ||
                (st = b->dRingExp.alloc(waves * (size_t)(ringLen / kK) * 4))) {
                discard_batch(b, inFlight);
                return st;
            }
        } else {
            if ((st = b->dRing.alloc(waves * (size_t)ringLen * 64 * 4))) {
                discard_batch(b, inFlight);
                return st;
            }
        }
    }
    if (b->wideEnd > b->wideBegin) {
        // wide jobs: one wavefront each; ring of one float4 per cell (3-state: values and the exponent)
        // or six values (five-state kernel: values and the cell's exponent, FP32 or FP64); the resident
        // waves' rings take at most ~70% of the free device memory
        const size_t perCell = five ? 6 * (fp64 ? sizeof(double) : sizeof(float)) : sizeof(float4);
        const size_t perWave = five ? (size_t)std::max(b->ringWideLen, 1) * 64 * b->wideC * perCell
                                    : (size_t)std::max<int64_t>(b->ringWideCells, 1) * perCell +
                                          (size_t)std::max(b->ringWideLen, 1) * sizeof(int32_t);
        size_t freeB = 0, totalB = 0;
        HIP_TRY(hipMemGetInfo(&freeB, &totalB));
        const size_t cap = std::max<size_t>(1, freeB / 10 * 7 / perWave);
        // resident waves per CU: five-state, one workgroup of up to 8 waves per job (4 jobs); 3-state, one
        // job per wave in workgroups of wide_group_waves(C), as many workgroups as the LDS holds
        size_t perCU = 4;
        if (!five) {
            const int n = wide_group_waves(std::max(b->wideC, 1));
            perCU = (size_t)n * (size_t)std::max(1, kWideLdsLimit / wide_lds_bytes(std::max(b->wideC, 1), n));
        }
        b->wideWaves = (int32_t)std::min<size_t>({(size_t)(b->wideEnd - b->wideBegin), (size_t)e->cus * perCU, cap});
        const size_t ringBytes = five ? perWave : (size_t)std::max<int64_t>(b->ringWideCells, 1) * perCell;
        if ((st = b->dRingWide.alloc((size_t)b->wideWaves * ringBytes / sizeof(float))) ||
            (!five && (st = b->dRingWideIdx.alloc((size_t)b->wideWaves * std::max(b->ringWideLen, 1))))) {
            discard_batch(b, inFlight);
            return st;
        }
    }
    kp.jobs = b->dJobs.p;
    kp.sym = b->dSym.p;
    kp.band = b->dBand.p;
    kp.win = b->dWin.p;
    kp.out = b->dOut.p;
    kp.counts = b->dCounts.p;
    kp.winEnd = b->dWinEnd.p;
    kp.logTotal = b->dLogTotal.p;
    kp.chain = b->dChain.p;
    kp.tbd = (int32_t)p->traceBackDiagonals;
    kp.ring = b->dRing.p;
    kp.ringExp = b->dRingExp.p;
    kp.ringWide = b->dRingWide.p;
    kp.ringWideLen = std::max(b->ringWideLen, 1);
    kp.ringWideCells = std::max<int64_t>(b->ringWideCells, 1);
    kp.ringWideIdx = b->dRingWideIdx.p;
    kp.wideC = std::max(b->wideC, 1);
    kp.wideWaves = b->wideWaves;
    kp.wideBegin = b->wideBegin;
    kp.wideEnd = b->wideEnd;
    kp.tables = b->dTables.p;
    kp.queue = b->dQueue.p;
    kp.ringLen = ringLen;
    kp.withIndels = mode == MP_MODE_POSTERIORS;
    kp.mode = mode;
    kp.expect = b->dExpect.p;
    kp.emitCount = b->dEmit.p;
    kp.error = b->dError.p;
    kp.threshold = (float)p->threshold;
    for (int c = 0; c < kNumClasses; c++) {
        kp.jobBegin[c] = b->classBegin[c];
        kp.jobEnd[c] = b->classEnd[c];
    }
    {
        std::vector<int32_t> pairSlot((size_t)npairs + 1);
        for (int64_t i = 0; i < npairs; i++) pairSlot[i] = (int32_t)b->pairFirstJob[i];
        pairSlot[npairs] = (int32_t)b->order.size();
        HIP_TRY(hipStreamSynchronize(s));
        mark(6);
        if ((st = b->dPairSlot.upload(pairSlot, s)) || (st = alloc_lists(b))) {
            discard_batch(b, inFlight);
            return st;
        }
    }
    HIP_TRY(hipStreamSynchronize(s));
    mark(7);
    *out = b;
    return MP_OK;
}

// A batch that batch_create leaves behind on an early HIP error or an
// exception (host memory exhausted) is freed here; the error paths that
// free the batch themselves clear `inFlight` first.
extern "C" mp_status mp_batch_create(mp_engine *e, const mp_model *m, const mp_pair *pairs, int64_t npairs,
                                     const mp_band_params *p, int32_t mode, int32_t flags, mp_batch **out) {
    mp_batch *inFlight = nullptr;
    mp_status st;
    try {
        st = batch_create(e, m, pairs, npairs, p, mode, flags, out, &inFlight);
    } catch (const std::bad_alloc &) {
        mp::set_error("mp_batch_create: out of host memory");
        st = MP_ENOMEM;
    }
    if (st != MP_OK && inFlight) {
        delete inFlight;
        if (out) *out = nullptr;
    }
    return st;
}

static void drop_graph(mp_batch *b) {
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    if (b->graphLists) (void)hipGraphExecDestroy(b->graphLists);
    b->graph = nullptr;
    b->graphLists = nullptr;
}

// The device work of one batched run: clear the counters, then the narrow
// and wide launches.  Enqueued on the batch's stream, or captured into a graph.
static mp_status enqueue_run(mp_batch *b) {
    const int32_t nj = (int32_t)b->jobs.size();
    mp_status st;
    if (b->ref) {
        if ((st = reset_i32(b, b->dError.p, 1)) || (st = reset_i32(b, b->dQueue.p, 1))) return st;
        const dim3 grid(b->refBlocks), block(b->refThreads);
        if (b->rp.poly && b->refThreads == 512)
            hipLaunchKernelGGL((mpk::pairhmm_ref_kernel<true, 512>), grid, block, 0, b->eng->stream, b->rp);
        else if (b->rp.poly)
            hipLaunchKernelGGL((mpk::pairhmm_ref_kernel<true, 256>), grid, block, 0, b->eng->stream, b->rp);
        else if (b->refThreads == 512)
            hipLaunchKernelGGL((mpk::pairhmm_ref_kernel<false, 512>), grid, block, 0, b->eng->stream, b->rp);
        else
            hipLaunchKernelGGL((mpk::pairhmm_ref_kernel<false, 256>), grid, block, 0, b->eng->stream, b->rp);
        HIP_TRY(hipGetLastError());
        return MP_OK;
    }
    if (b->mode == MP_MODE_EXPECTATIONS && (st = reset_i32(b, reinterpret_cast<int32_t *>(b->dExpect.p), 20 * nj)))
        return st;
    if (b->mode == MP_MODE_EXPECTATIONS &&
        (st = reset_i32(b, reinterpret_cast<int32_t *>(b->dEmit.p), 2 * kTabFloats)))
        return st;
    if ((st = reset_i32(b, b->dError.p, 1))) return st;
    if ((st = reset_i32(b, b->dQueue.p, kNumClasses + 1))) return st;
    const bool anyNarrow = b->classEnd[0] > b->classBegin[0] || b->classEnd[1] > b->classBegin[1] ||
                           b->classEnd[2] > b->classBegin[2];
    if (anyNarrow && (st = launch_narrow(b, b->narrowBlocks, b->narrowThreads, b->kp))) return st;
    if (b->wideEnd > b->wideBegin && (st = launch_wide(b, b->wideWaves, b->kp))) return st;
    return MP_OK;
}

// The run's list assembly and log-likelihoods (mp_lists.hpp): every size is
// read from device memory, so the same launches serve every run.
static mp_status enqueue_lists(mp_batch *b) {
    hipStream_t s = b->eng->stream;
    const int64_t np = b->npairs;
    const int32_t ns = (int32_t)b->order.size(), nj = (int32_t)b->jobs.size();
    mp_status st;
    if (b->ref) {
        const unsigned gp = (unsigned)std::min<int64_t>((np + 256) / 256, 65535);
        if ((st = reset_i32(b, b->dOverflow.p, 1))) return st;
        hipLaunchKernelGGL(mpk::ref_slot_scan, dim3(1), dim3(1024), 0, s, b->dSlotJob.p, ns, b->dCounts.p,
                           b->dRefJobs.p, b->dError.p, b->dSlotBase3.p, b->dOverflow.p);
        const int4 cap = make_int4((int)std::min<int64_t>(b->listCap[0], INT32_MAX),
                                   (int)std::min<int64_t>(b->listCap[1], INT32_MAX),
                                   (int)std::min<int64_t>(b->listCap[2], INT32_MAX), 0);
        hipLaunchKernelGGL(mpk::ref_lists_copy, dim3((unsigned)std::max(1, std::min(ns, 65535))), dim3(256), 0, s,
                           b->dSlotJob.p, ns, b->dCounts.p, b->dRefJobs.p, b->dRaw.p, b->dSlotBase3.p, b->dOverflow.p,
                           cap, b->packBits, b->dList[0].p, b->dList[1].p, b->dList[2].p);
        hipLaunchKernelGGL(mpk::ref_pair_starts, dim3(gp), dim3(256), 0, s, b->dPairSlot.p, np, b->dSlotBase3.p,
                           b->dOverflow.p, b->dStarts.p, b->dListTotals.p);
        if (np > 0) {
            hipLaunchKernelGGL(mpk::ref_pair_log_likelihood, dim3(gp), dim3(256), 0, s, b->dPairSlot.p, np,
                               b->dSlotJob.p, b->dRefLogTotal.p, b->dError.p, b->dPairLL.p);
            hipLaunchKernelGGL(mpk::batch_log_likelihood, dim3(1), dim3(1024), 0, s, b->dPairLL.p, np, b->dError.p,
                               b->dBatchLL.p);
        }
        HIP_TRY(hipGetLastError());
        return MP_OK;
    }
    if (lists_mode(b)) {
        if ((st = reset_i32(b, b->dOverflow.p, 1))) return st;
        hipLaunchKernelGGL(mpk::slot_record_scan, dim3(1), dim3(1024), 0, s, b->dSlotJob.p, ns, b->dCounts.p,
                           b->dJL.p, b->dError.p, b->dSlotBase.p, b->dOverflow.p);
        const dim3 gseg((unsigned)std::max(1, std::min(2 * nj, 65535)));
        const int64_t nt = b->nListTiles;
        const dim3 gtile((unsigned)std::max<int64_t>(1, std::min<int64_t>(nt, (int64_t)1 << 22)));
        if (nt > 0)
            hipLaunchKernelGGL(mpk::tile_counts, gtile, dim3(mpk::kTile), 0, s, b->dOut.p, b->dJL.p, b->dCounts.p,
                               b->dWinEnd.p, b->dBand.p, b->dTileMeta.p, nt, b->dOverflow.p, b->dError.p,
                               b->kp.threshold, b->kp.withIndels, b->dTileCnt.p);
        hipLaunchKernelGGL(mpk::side_tile_scan, dim3((unsigned)std::max(1, std::min((2 * nj + 255) / 256, 65535))),
                           dim3(256), 0, s, b->dJL.p, b->dCounts.p, b->dWinEnd.p, b->dSegTile.p, 2 * nj, b->dTileCnt.p,
                           b->dOverflow.p, b->dError.p, b->dTilePre.p, b->dSegCnt.p);
        if (b->multiWin)
            hipLaunchKernelGGL(mpk::window_counts, gseg, dim3(mpk::kTile), 0, s, b->dOut.p, b->dJL.p, b->dCounts.p,
                               b->dWinEnd.p, b->dBand.p, b->dOverflow.p, b->dError.p, b->kp.threshold,
                               b->kp.withIndels, 2 * nj, b->dSegCnt.p);
        hipLaunchKernelGGL(mpk::slot_list_scan, dim3(1), dim3(1024), 0, s, b->dSlotJob.p, ns, b->dJL.p, b->dSegCnt.p,
                           b->dOverflow.p, b->dError.p, b->dSlotBase3.p, b->dSegBase.p);
        const int4 cap = make_int4((int)std::min<int64_t>(b->listCap[0], INT32_MAX),
                                   (int)std::min<int64_t>(b->listCap[1], INT32_MAX),
                                   (int)std::min<int64_t>(b->listCap[2], INT32_MAX), 0);
        if (nt > 0)
            hipLaunchKernelGGL(mpk::emit_tiles, gtile, dim3(mpk::kTile), 0, s, b->dOut.p, b->dJL.p, b->dCounts.p,
                               b->dWinEnd.p, b->dBand.p, b->dTileMeta.p, nt, b->dTilePre.p,
                               b->dSegCnt.p, b->dSegBase.p, b->dOverflow.p, b->dError.p, b->kp.threshold,
                               b->kp.withIndels, cap, b->packBits, b->dList[0].p, b->dList[1].p, b->dList[2].p);
        if (b->multiWin)
            hipLaunchKernelGGL(mpk::emit_windows, gseg, dim3(mpk::kTile), 0, s, b->dOut.p, b->dJL.p, b->dCounts.p,
                               b->dWinEnd.p, b->dBand.p, b->dSegCnt.p, b->dSegBase.p, b->dOverflow.p, b->dError.p,
                               b->kp.threshold, b->kp.withIndels, 2 * nj, cap, b->packBits, b->dList[0].p,
                               b->dList[1].p, b->dList[2].p);
        const unsigned gp = (unsigned)std::min<int64_t>((np + 256) / 256, 65535);
        hipLaunchKernelGGL(mpk::ref_pair_starts, dim3(gp), dim3(256), 0, s, b->dPairSlot.p, np, b->dSlotBase3.p,
                           b->dOverflow.p, b->dStarts.p, b->dListTotals.p);
        HIP_TRY(hipGetLastError());
    }
    if (np > 0) {
        const unsigned gp = (unsigned)std::min<int64_t>((np + 255) / 256, 65535);
        hipLaunchKernelGGL(mpk::pair_log_likelihood, dim3(gp), dim3(256), 0, s, b->dPairSlot.p, np, b->dSlotJob.p,
                           b->dSlotK.p, b->dJobs.p, b->dLogTotal.p, b->dChain.p, b->dError.p, b->dPairLL.p);
        hipLaunchKernelGGL(mpk::batch_log_likelihood, dim3(1), dim3(1024), 0, s, b->dPairLL.p, np, b->dError.p,
                           b->dBatchLL.p);
        HIP_TRY(hipGetLastError());
    }
    return MP_OK;
}

// hipGraph of enqueue_run or enqueue_lists (the per-iteration graphs: captured
// once, replayed every run until the parameter block changes).  The graphs hold
// kernel nodes only; counters are cleared by fill kernels, not memset nodes.
static mp_status capture(mp_batch *b, mp_status (*enqueue)(mp_batch *), hipGraphExec_t *exec) {
    hipStream_t s = b->eng->stream;
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const mp_status st = enqueue(b);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(s, &g);
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (ec != hipSuccess) {
        mp::set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
        return MP_EHIP;
    }
    const hipError_t ei = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        *exec = nullptr;
        mp::set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        return MP_EHIP;
    }
    return MP_OK;
}

static mp_status batch_run_impl(mp_batch *b, bool plain) {
    mp_engine *e = b->eng;
    HIP_TRY(hipSetDevice(e->device));
    hipStream_t s = e->stream;
    const int32_t nj = (int32_t)b->jobs.size();
    b->listsReady = false;
    if (nj == 0) {
        b->kernelMs = b->listMs = 0.0;
        b->ran = true;
        b->hCounts.clear();
        for (int l = 0; l < 3; l++) b->listStart[l].assign(b->npairs + 1, 0);
        b->pairLogProb.assign(b->npairs, 0.0);
        b->batchLogProb = 0.0;
        b->listsReady = true;
        return MP_OK;
    }
    mp_status st;
    float ms = 0.f, msLists = 0.f;
    if (b->ref) plain = false;  // reference arithmetic: one workgroup per job either way
    const bool useGraph = !plain && (b->flags & MP_BATCH_NO_GRAPH) == 0;
    if (useGraph && !b->graph && (st = capture(b, enqueue_run, &b->graph))) return st;
    if (useGraph && !b->graphLists && (st = capture(b, enqueue_lists, &b->graphLists))) return st;
    if (!plain) {
        HIP_TRY(hipEventRecord(e->ev0, s));
        if (useGraph) {
            HIP_TRY(hipGraphLaunch(b->graph, s));
        } else if ((st = enqueue_run(b))) {
            return st;
        }
        HIP_TRY(hipEventRecord(e->ev1, s));
    } else {
        if (b->mode == MP_MODE_EXPECTATIONS &&
            (st = reset_i32(b, reinterpret_cast<int32_t *>(b->dExpect.p), 20 * nj)))
            return st;
        if (b->mode == MP_MODE_EXPECTATIONS &&
            (st = reset_i32(b, reinterpret_cast<int32_t *>(b->dEmit.p), 2 * kTabFloats)))
            return st;
        if ((st = reset_i32(b, b->dError.p, 1))) return st;
        // one launch per job (one row, or one forward/backward pair of rows), in device order
        DevBuf<int32_t> q;
        const int nq = kNumClasses + 1;
        if ((st = q.alloc((size_t)nj * nq))) return st;
        if ((st = reset_i32(b, q.p, nj * nq))) return st;
        HIP_TRY(hipEventRecord(e->ev0, s));
        for (int32_t j = 0; j < nj; j++) {
            KParams kp = b->kp;
            for (int c = 0; c < kNumClasses; c++) kp.jobBegin[c] = kp.jobEnd[c] = j;
            kp.wideBegin = kp.wideEnd = j;
            const int c = b->jobs[j].cls;
            kp.queue = q.p + (size_t)j * nq;
            if (c < 3) {
                kp.jobEnd[c] = j + 1;
                if ((st = launch_narrow(b, 1, use_split(b->mode) ? 128 : 64, kp))) return st;
            } else {
                kp.wideEnd = j + 1;
                if ((st = launch_wide(b, 1, kp))) return st;
            }
        }
        HIP_TRY(hipEventRecord(e->ev1, s));
    }
    // list assembly and log-likelihoods on the device
    if (useGraph) {
        HIP_TRY(hipGraphLaunch(b->graphLists, s));
    } else if ((st = enqueue_lists(b))) {
        return st;
    }
    HIP_TRY(hipEventRecord(e->ev2, s));
    HIP_TRY(hipEventSynchronize(e->ev2));
    HIP_TRY(hipEventElapsedTime(&ms, e->ev0, e->ev1));
    HIP_TRY(hipEventElapsedTime(&msLists, e->ev1, e->ev2));
    b->kernelMs = ms;
    b->listMs = msLists;
    int32_t kerr = 0, ovf = 0;
    HIP_TRY(hipMemcpy(&kerr, b->dError.p, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (kerr) {
        mp::set_error("pair-HMM kernel: forward/backward sweeps lost step (code " + std::to_string(kerr) + ")");
        return MP_EHIP;
    }
    if (lists_mode(b)) HIP_TRY(hipMemcpy(&ovf, b->dOverflow.p, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (ovf && b->ref) {
        // tuple regions too small: size them exactly and run again
        std::vector<int32_t> c((size_t)nj * 3);
        HIP_TRY(hipMemcpy(c.data(), b->dCounts.p, sizeof(int32_t) * 3 * nj, hipMemcpyDeviceToHost));
        int64_t off = 0;
        b->lastReruns++;
        for (int32_t j = 0; j < nj; j++)
            for (int l = 0; l < 3; l++) {
                mpk::RefJob &R = b->refJobs[j];
                R.cap[l] = std::max(R.cap[l], c[3 * j + l]);
                R.outOff[l] = off;
                off += R.cap[l];
            }
        if ((st = b->dRefJobs.upload(b->refJobs, s)) || (st = b->dRaw.alloc((size_t)std::max<int64_t>(off, 1))))
            return st;
        b->rp.jobs = b->dRefJobs.p;
        b->rp.out = b->dRaw.p;
        b->outTotal = off;
        drop_graph(b);
        return batch_run_impl(b, plain);
    }
    // Jobs that wrote more records than their regions hold run again alone, into
    // regions of their exact size appended to the record buffer; every other
    // job's records stay where they are.
    if (ovf && b->lastReruns == 0) {
        b->hCounts.resize((size_t)nj * 2);
        HIP_TRY(hipMemcpy(b->hCounts.data(), b->dCounts.p, sizeof(int32_t) * 2 * nj, hipMemcpyDeviceToHost));
        std::vector<DevJob> dj(nj);
        HIP_TRY(hipMemcpy(dj.data(), b->dJobs.p, sizeof(DevJob) * nj, hipMemcpyDeviceToHost));
        std::vector<uint8_t> over((size_t)nj, 0);
        int64_t add = 0;
        b->lastReruns++;
        for (int32_t j = 0; j < nj; j++) {
            for (int l = 0; l < 2; l++)
                if (b->hCounts[2 * j + l] > b->jobs[j].cap[l]) {
                    over[j] = 1;
                    b->lastOvfMissing += b->hCounts[2 * j + l] - b->jobs[j].cap[l];
                }
            if (!over[j]) continue;
            b->lastOvfJobs++;
            for (int l = 0; l < 2; l++) b->jobs[j].cap[l] = std::max(b->jobs[j].cap[l], b->hCounts[2 * j + l]);
            add += (int64_t)b->jobs[j].cap[0] + b->jobs[j].cap[1];
        }
        const int64_t oldTotal = b->outTotal;
        {
            DevBuf<int4> grown;
            if ((st = grown.alloc((size_t)(oldTotal + add)))) return st;
            HIP_TRY(hipMemcpyAsync(grown.p, b->dOut.p, sizeof(int4) * (size_t)oldTotal, hipMemcpyDeviceToDevice, s));
            HIP_TRY(hipStreamSynchronize(s));
            std::swap(grown.p, b->dOut.p);
            std::swap(grown.n, b->dOut.n);
        }
        int64_t off = oldTotal;
        for (int32_t j = 0; j < nj; j++) {
            if (!over[j]) continue;
            b->jobs[j].outOff = off;
            dj[j].outOff = off;
            dj[j].cap[0] = b->jobs[j].cap[0];
            dj[j].cap[1] = b->jobs[j].cap[1];
            off += (int64_t)dj[j].cap[0] + dj[j].cap[1];
        }
        HIP_TRY(hipMemcpy(b->dJobs.p, dj.data(), sizeof(DevJob) * nj, hipMemcpyHostToDevice));
        b->kp.out = b->dOut.p;
        b->outTotal = oldTotal + add;
        if ((st = alloc_lists(b))) return st;
        drop_graph(b);
        // the over-capacity jobs of each kernel, in device order, as the launches' job map
        std::vector<int32_t> map;
        KParams kr = b->kp;
        for (int c = 0; c < kNumClasses; c++) {
            kr.jobBegin[c] = (int32_t)map.size();
            for (int32_t q = b->classBegin[c]; q < b->classEnd[c]; q++)
                if (over[q]) map.push_back(q);
            kr.jobEnd[c] = (int32_t)map.size();
        }
        kr.wideBegin = (int32_t)map.size();
        for (int32_t q = b->wideBegin; q < b->wideEnd; q++)
            if (over[q]) map.push_back(q);
        kr.wideEnd = (int32_t)map.size();
        if ((st = b->dJobMap.upload(map, s))) return st;
        kr.jobMap = b->dJobMap.p;
        float msRe = 0.f, msReLists = 0.f;
        HIP_TRY(hipEventRecord(e->ev0, s));
        if ((st = reset_i32(b, b->dError.p, 1)) || (st = reset_i32(b, b->dQueue.p, kNumClasses + 1))) return st;
        const bool anyNarrow = kr.jobEnd[0] > kr.jobBegin[0] || kr.jobEnd[1] > kr.jobBegin[1] ||
                               kr.jobEnd[2] > kr.jobBegin[2];
        if (anyNarrow && (st = launch_narrow(b, b->narrowBlocks, b->narrowThreads, kr))) return st;
        if (kr.wideEnd > kr.wideBegin && (st = launch_wide(b, b->wideWaves, kr))) return st;
        HIP_TRY(hipEventRecord(e->ev1, s));
        if ((st = enqueue_lists(b))) return st;
        HIP_TRY(hipEventRecord(e->ev2, s));
        HIP_TRY(hipEventSynchronize(e->ev2));
        HIP_TRY(hipEventElapsedTime(&msRe, e->ev0, e->ev1));
        HIP_TRY(hipEventElapsedTime(&msReLists, e->ev1, e->ev2));
        b->kernelMs += msRe;
        b->listMs += msReLists;
        HIP_TRY(hipMemcpy(&kerr, b->dError.p, sizeof(int32_t), hipMemcpyDeviceToHost));
        if (kerr) {
            mp::set_error("pair-HMM kernel: forward/backward sweeps lost step (code " + std::to_string(kerr) + ")");
            return MP_EHIP;
        }
        HIP_TRY(hipMemcpy(&ovf, b->dOverflow.p, sizeof(int32_t), hipMemcpyDeviceToHost));
    }
    // Still overflowing (not expected after the partial run): size every region exactly
    // and run the whole batch again.
    if (ovf) {
        b->hCounts.resize((size_t)nj * 2);
        HIP_TRY(hipMemcpy(b->hCounts.data(), b->dCounts.p, sizeof(int32_t) * 2 * nj, hipMemcpyDeviceToHost));
        std::vector<DevJob> dj(nj);
        HIP_TRY(hipMemcpy(dj.data(), b->dJobs.p, sizeof(DevJob) * nj, hipMemcpyDeviceToHost));
        int64_t off = 0;
        b->lastReruns++;
        for (int32_t j = 0; j < nj; j++) {
            bool over = false;
            for (int l = 0; l < 2; l++) {
                if (b->hCounts[2 * j + l] > b->jobs[j].cap[l]) {
                    over = true;
                    b->lastOvfMissing += b->hCounts[2 * j + l] - b->jobs[j].cap[l];
                }
                b->jobs[j].cap[l] = std::max(b->jobs[j].cap[l], b->hCounts[2 * j + l]);
                dj[j].cap[l] = b->jobs[j].cap[l];
            }
            b->lastOvfJobs += over;
            b->jobs[j].outOff = off;
            dj[j].outOff = off;
            off += (int64_t)dj[j].cap[0] + dj[j].cap[1];
        }
        HIP_TRY(hipMemcpy(b->dJobs.p, dj.data(), sizeof(DevJob) * nj, hipMemcpyHostToDevice));
        if ((st = b->dOut.alloc(std::max<int64_t>(off, 1)))) return st;
        b->kp.out = b->dOut.p;
        b->outTotal = off;
        if ((st = alloc_lists(b))) return st;
        drop_graph(b);
        return batch_run_impl(b, plain);
    }
    if (lists_mode(b)) {
        // lists longer than their device buffers: grow them and assemble the lists again
        // (the kernels' records stay valid)
        int4 tot;
        int64_t recs = 0;
        HIP_TRY(hipMemcpy(&tot, b->dListTotals.p, sizeof(int4), hipMemcpyDeviceToHost));
        if (b->ref)
            recs = (int64_t)tot.x + tot.y + tot.z;
        else
            HIP_TRY(hipMemcpy(&recs, b->dSlotBase.p + b->order.size(), sizeof(int64_t), hipMemcpyDeviceToHost));
        b->records = recs;
        const int64_t need[3] = {tot.x, tot.y, tot.z};
        bool grow = false;
        for (int l = 0; l < 3; l++)
            if (need[l] > b->listCap[l]) {
                if ((st = alloc_list(b, l, need[l] + need[l] / 16 + 1024))) return st;
                grow = true;
            }
        if (grow) {
            if (b->graphLists) (void)hipGraphExecDestroy(b->graphLists);
            b->graphLists = nullptr;
            if ((st = enqueue_lists(b))) return st;
            HIP_TRY(hipStreamSynchronize(s));
        }
    }
    b->ran = true;
    return MP_OK;
}

extern "C" mp_status mp_batch_stats(const mp_batch *b, int64_t *jobs, int64_t *recordCapacity, int64_t *records) {
    if (!b) {
        mp::set_error("mp_batch_stats: null batch");
        return MP_EINVAL;
    }
    if (jobs) *jobs = (int64_t)b->jobs.size();
    if (recordCapacity) *recordCapacity = b->outTotal;
    if (records) *records = b->records;
    return MP_OK;
}

extern "C" mp_status mp_batch_prep_ms(const mp_batch *b, double *ms, int32_t n) {
    if (!b || (n > 0 && !ms)) {
        mp::set_error("mp_batch_prep_ms: null argument");
        return MP_EINVAL;
    }
    for (int32_t i = 0; i < n && i < 8; i++) ms[i] = b->prepMs[i];
    return MP_OK;
}

extern "C" mp_status mp_batch_rerun_stats(const mp_batch *b, int32_t *reruns, int64_t *overflowJobs,
                                           int64_t *missingRecords) {
    if (!b) {
        mp::set_error("mp_batch_rerun_stats: null batch");
        return MP_EINVAL;
    }
    if (reruns) *reruns = b->lastReruns;
    if (overflowJobs) *overflowJobs = b->lastOvfJobs;
    if (missingRecords) *missingRecords = b->lastOvfMissing;
    return MP_OK;
}

extern "C" mp_status mp_batch_set_model(mp_batch *b, const mp_model *m) {
    if (!b || !m) {
        mp::set_error("mp_batch_set_model: null argument");
        return MP_EINVAL;
    }
    if (b->ref) {
        mp::set_error("mp_batch_set_model: not available for reference-arithmetic batches");
        return MP_EUNSUPPORTED;
    }
    if (m->mach[1].nStates != b->modelStates) {
        mp::set_error("mp_batch_set_model: the model's state count differs from the batch's");
        return MP_EINVAL;
    }
    std::vector<float> tab;
    KParams kp = b->kp;
    build_tables(m, (b->flags & MP_BATCH_FP64) != 0, tab, kp);
    mp_status st;
    if ((st = b->dTables.upload(tab, b->eng->stream))) return st;
    std::memcpy(b->kp.T, kp.T, sizeof(kp.T));
    std::memcpy(b->kp.startV, kp.startV, sizeof(kp.startV));
    std::memcpy(b->kp.endV, kp.endV, sizeof(kp.endV));
    std::memcpy(b->kp.T5, kp.T5, sizeof(kp.T5));
    std::memcpy(b->kp.startV5, kp.startV5, sizeof(kp.startV5));
    std::memcpy(b->kp.endV5, kp.endV5, sizeof(kp.endV5));
    b->kp.tables = b->dTables.p;
    drop_graph(b);
    b->ran = false;
    b->listsReady = false;
    return MP_OK;
}

extern "C" mp_status mp_batch_autotune(mp_batch *b, int32_t repeats, int32_t *narrowThreads, int32_t *wideWaves,
                                       float *kernelMs) {
    if (!b) {
        mp::set_error("mp_batch_autotune: null batch");
        return MP_EINVAL;
    }
    if (b->ref) {
        mp::set_error("mp_batch_autotune: not available for reference-arithmetic batches");
        return MP_EUNSUPPORTED;
    }
    repeats = std::max(1, std::min(repeats, 16));
    // hardware tiling only: the results do not depend on it (jobs come from a queue, each job's
    // records land in its own region)
    std::vector<int32_t> nt = {512, 256};
    std::vector<int32_t> ww = {0};
    if (b->kp.nStates == 5) {
        ww.clear();
        for (int w : {wide5_waves(b->wideC), 4, 2, 1})
            if (w <= wide5_waves(b->wideC) && std::find(ww.begin(), ww.end(), w) == ww.end()) ww.push_back(w);
    }
    float best = 1e30f;
    int32_t bestT = b->narrowThreads, bestW = b->wide5Waves;
    mp_status st;
    for (int32_t t : nt)
        for (int32_t w : ww) {
            b->narrowThreads = t;
            b->narrowBlocks = b->narrowWaves * 64 / t;
            b->wide5Waves = w;
            drop_graph(b);
            float ms = 1e30f;
            for (int r = 0; r < repeats; r++) {
                if ((st = batch_run_impl(b, false))) return st;
                ms = std::min(ms, (float)b->kernelMs);
            }
            if (ms < best) {
                best = ms;
                bestT = t;
                bestW = w;
            }
        }
    b->narrowThreads = bestT;
    b->narrowBlocks = b->narrowWaves * 64 / bestT;
    b->wide5Waves = bestW;
    drop_graph(b);
    b->eng->tuned[b->tuneKey] = {bestT, bestW};
    if ((st = batch_run_impl(b, false))) return st;
    if (narrowThreads) *narrowThreads = bestT;
    if (wideWaves) *wideWaves = bestW;
    if (kernelMs) *kernelMs = best;
    return MP_OK;
}

extern "C" mp_status mp_batch_run(mp_batch *b) {
    if (!b) {
        mp::set_error("mp_batch_run: null batch");
        return MP_EINVAL;
    }
    b->lastReruns = 0;
    b->lastOvfJobs = b->lastOvfMissing = 0;
    return batch_run_impl(b, false);
}

extern "C" mp_status mp_batch_run_plain(mp_batch *b) {
    if (!b) {
        mp::set_error("mp_batch_run_plain: null batch");
        return MP_EINVAL;
    }
    b->lastReruns = 0;
    b->lastOvfJobs = b->lastOvfMissing = 0;
    return batch_run_impl(b, true);
}

extern "C" double mp_batch_kernel_ms(const mp_batch *b) { return b ? b->kernelMs : 0.0; }
extern "C" double mp_batch_list_ms(const mp_batch *b) { return b ? b->listMs : 0.0; }
extern "C" int64_t mp_batch_cells(const mp_batch *b) { return b ? b->cells : 0; }
extern "C" int64_t mp_batch_npairs(const mp_batch *b) { return b ? b->npairs : 0; }

// The run assembled the lists and the log-likelihoods on the device
// (mp_lists.hpp); bring the per-pair list offsets and log-likelihoods to the
// host once per run.
static mp_status materialize(mp_batch *b) {
    if (b->listsReady) return MP_OK;
    const int64_t np = b->npairs;
    HIP_TRY(hipSetDevice(b->eng->device));
    hipStream_t s = b->eng->stream;
    std::vector<int64_t> starts(3 * (np + 1), 0);
    b->pairLogProb.assign(np, 0.0);
    b->batchLogProb = 0.0;
    if (lists_mode(b))
        HIP_TRY(hipMemcpyAsync(starts.data(), b->dStarts.p, sizeof(int64_t) * starts.size(), hipMemcpyDeviceToHost,
                               s));
    if (np > 0) {
        HIP_TRY(hipMemcpyAsync(b->pairLogProb.data(), b->dPairLL.p, sizeof(double) * np, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(&b->batchLogProb, b->dBatchLL.p, sizeof(double), hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    for (int l = 0; l < 3; l++) b->listStart[l].assign(starts.begin() + l * (np + 1), starts.begin() + (l + 1) * (np + 1));
    if (b->listStart[0][np] >= INT32_MAX || b->listStart[1][np] >= INT32_MAX || b->listStart[2][np] >= INT32_MAX) {
        mp::set_error("batch too large: a list holds more than 2^31 tuples (split it into several batches)");
        return MP_EUNSUPPORTED;
    }
    b->listsReady = true;
    return MP_OK;
}

// 64-bit device list entries (mpk::put_tuple, pk = bits of x + 1) to the C ABI's tuples.
static void unpack_tuples(const uint64_t *src, mp_aligned_pair *dst, size_t n, int pk) {
    const uint64_t mx = ((uint64_t)1 << pk) - 1;
    for (size_t i = 0; i < n; i++) {
        const uint64_t v = src[i];
        dst[i].weight = (int64_t)(v & 0xffffff);
        dst[i].x = (int32_t)((v >> 24) & mx) - 1;
        dst[i].y = (int32_t)(v >> (24 + pk)) - 1;
    }
}

// First n tuples of list l (device) to host memory.
static mp_status copy_list(mp_batch *b, int32_t l, mp_aligned_pair *out, int64_t n) {
    static_assert(sizeof(mpk::ListTuple) == sizeof(mp_aligned_pair), "tuple layout");
    if (n <= 0) return MP_OK;
    if (!out) {
        mp::set_error("mp_batch_fetch_list: null output");
        return MP_EINVAL;
    }
    HIP_TRY(hipSetDevice(b->eng->device));
    mp_engine *e = b->eng;
    const int pk = b->packBits;
    const size_t esz = pk ? sizeof(uint64_t) : sizeof(mp_aligned_pair);  // bytes per device entry
    const size_t bytes = esz * (size_t)n;
    if (bytes < ((size_t)8 << 20)) {
        if (!pk) {
            HIP_TRY(hipMemcpy(out, b->dList[l].p, bytes, hipMemcpyDeviceToHost));
        } else {
            std::vector<uint64_t> tmp((size_t)n);
            HIP_TRY(hipMemcpy(tmp.data(), b->dList[l].p, bytes, hipMemcpyDeviceToHost));
            unpack_tuples(tmp.data(), out, (size_t)n, pk);
        }
        return MP_OK;
    }
    // Large lists: DMA into a pinned double buffer in 64 MB chunks; while the
    // next chunk is in flight, host threads copy the previous one out.
    if (!e->stage[0]) {
        for (int k = 0; k < 2; k++) {
            HIP_TRY(hipHostMalloc(&e->stage[k], mp_engine::kStageBytes, hipHostMallocDefault));
            HIP_TRY(hipEventCreateWithFlags(&e->stageEv[k], hipEventDisableTiming));
        }
    }
    const char *src = reinterpret_cast<const char *>(b->dList[l].p);
    char *dst = reinterpret_cast<char *>(out);
    const size_t chunk = mp_engine::kStageBytes;
    const size_t nchunk = (bytes + chunk - 1) / chunk;
    const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto drain = [&](size_t k) {  // copy (or unpack) chunk k out of its staging buffer (after its DMA)
        const size_t off = k * chunk, len = std::min(chunk, bytes - off);
        const char *sb = reinterpret_cast<const char *>(e->stage[k & 1]);
        if (pk) {
            const size_t cnt = len / sizeof(uint64_t), part = (cnt + nth - 1) / nth;
            const uint64_t *sv = reinterpret_cast<const uint64_t *>(sb);
            mp::parallel_run(nth, [&](int t) {
                if ((size_t)t * part < cnt)
                    unpack_tuples(sv + t * part, out + off / sizeof(uint64_t) + t * part, std::min(part, cnt - t * part),
                                  pk);
            });
            return;
        }
        const size_t part = (len + nth - 1) / nth;
        mp::parallel_run(nth, [&](int t) {
            if ((size_t)t * part < len) std::memcpy(dst + off + t * part, sb + t * part, std::min(part, len - t * part));
        });
    };
    for (size_t k = 0; k < nchunk; k++) {
        const size_t off = k * chunk, len = std::min(chunk, bytes - off);
        HIP_TRY(hipMemcpyAsync(e->stage[k & 1], src + off, len, hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipEventRecord(e->stageEv[k & 1], e->stream));
        if (k > 0) {
            HIP_TRY(hipEventSynchronize(e->stageEv[(k - 1) & 1]));
            drain(k - 1);
        }
    }
    HIP_TRY(hipEventSynchronize(e->stageEv[(nchunk - 1) & 1]));
    drain(nchunk - 1);
    return MP_OK;
}

extern "C" mp_status mp_batch_fetch_counts(mp_batch *b, int64_t *counts, double *logProb) {
    if (!b || !counts) {
        mp::set_error("mp_batch_fetch_counts: null argument");
        return MP_EINVAL;
    }
    if (!b->ran) {
        mp::set_error("mp_batch_fetch_counts: batch has not run");
        return MP_EINVAL;
    }
    mp_status st = materialize(b);
    if (st) return st;
    for (int64_t i = 0; i < b->npairs; i++) {
        for (int l = 0; l < 3; l++) counts[3 * i + l] = b->listStart[l][i + 1] - b->listStart[l][i];
        if (logProb) logProb[i] = b->pairLogProb[i];
    }
    return MP_OK;
}

extern "C" mp_status mp_batch_fetch_total_log_prob(mp_batch *b, double *total) {
    if (!b || !total) {
        mp::set_error("mp_batch_fetch_total_log_prob: null argument");
        return MP_EINVAL;
    }
    if (!b->ran) {
        mp::set_error("mp_batch_fetch_total_log_prob: batch has not run");
        return MP_EINVAL;
    }
    mp_status st = materialize(b);
    if (st) return st;
    *total = b->batchLogProb;
    return MP_OK;
}

extern "C" mp_status mp_batch_fetch_list(mp_batch *b, int32_t list, mp_aligned_pair *out) {
    if (!b || list < 0 || list > 2) {
        mp::set_error("mp_batch_fetch_list: bad argument");
        return MP_EINVAL;
    }
    if (!b->ran) {
        mp::set_error("mp_batch_fetch_list: batch has not run");
        return MP_EINVAL;
    }
    mp_status st = materialize(b);
    if (st) return st;
    return copy_list(b, list, out, b->listStart[list][b->npairs]);
}

extern "C" mp_status mp_batch_poa_augment(mp_batch *b, mp_poa *const *poas, int64_t nPoas, const int64_t *pairPoa,
                                          const int64_t *pairReadNo, const mp_rle_string *reads,
                                          int32_t compareRepeatCounts, int32_t threads) {
    if (!b || !poas || nPoas < 0 || (b->npairs && (!pairPoa || !reads))) {
        mp::set_error("mp_batch_poa_augment: null argument");
        return MP_EINVAL;
    }
    if (!b->ran || b->mode != MP_MODE_POSTERIORS) {
        mp::set_error("mp_batch_poa_augment: needs a posterior-mode batch that has run");
        return MP_EINVAL;
    }
    mp_status st = materialize(b);
    if (st) return st;
    // the three lists straight into the engine's pinned buffers (one DMA each, no staging copy)
    mp_engine *e = b->eng;
    HIP_TRY(hipSetDevice(e->device));
    const mp_aligned_pair *lists[3];
    const int64_t *starts[3];
    for (int l = 0; l < 3; l++) {
        const size_t bytes = sizeof(mp_aligned_pair) * (size_t)std::max<int64_t>(b->listStart[l][b->npairs], 1);
        if (bytes > e->listHostBytes[l]) {
            const size_t want = std::max(bytes, e->listHostBytes[l] + e->listHostBytes[l] / 4);
            if (e->listHost[l]) HIP_TRY(hipHostFree(e->listHost[l]));
            e->listHost[l] = nullptr;
            e->listHostBytes[l] = 0;
            if (hipHostMalloc(&e->listHost[l], want, hipHostMallocDefault) != hipSuccess) {
                e->listHost[l] = nullptr;
                mp::set_error("mp_batch_poa_augment: pinned host allocation failed");
                return MP_ENOMEM;
            }
            e->listHostBytes[l] = want;
        }
        const int64_t n = b->listStart[l][b->npairs];
        if (n > 0 && b->packBits) {  // 64-bit entries: staged and unpacked by host threads
            mp_status cs = copy_list(b, l, static_cast<mp_aligned_pair *>(e->listHost[l]), n);
            if (cs) return cs;
        } else if (n > 0) {
            HIP_TRY(hipMemcpyAsync(e->listHost[l], b->dList[l].p, sizeof(mp_aligned_pair) * (size_t)n,
                                   hipMemcpyDeviceToHost, e->stream));
        }
        lists[l] = static_cast<const mp_aligned_pair *>(e->listHost[l]);
        starts[l] = b->listStart[l].data();
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    return mp::poa_augment_pairs(poas, nPoas, pairPoa, pairReadNo, reads, b->pairStrand.data(), b->npairs, lists,
                                 starts, compareRepeatCounts != 0, threads);
}

extern "C" mp_status mp_batch_fetch_expectations(mp_batch *b, double *trans, double *likelihood) {
    if (!b || !trans) {
        mp::set_error("mp_batch_fetch_expectations: null argument");
        return MP_EINVAL;
    }
    if (!b->ran || b->mode != MP_MODE_EXPECTATIONS) {
        mp::set_error("mp_batch_fetch_expectations: batch has not run in expectation mode");
        return MP_EINVAL;
    }
    const int32_t nj = (int32_t)b->jobs.size();
    std::vector<double> ex((size_t)std::max(nj, 1) * 10, 0.0);
    if (nj) {
        HIP_TRY(hipSetDevice(b->eng->device));
        HIP_TRY(hipMemcpy(ex.data(), b->dExpect.p, sizeof(double) * 10 * nj, hipMemcpyDeviceToHost));
    }
    for (int64_t i = 0; i < b->npairs; i++) {
        double t[10] = {0};
        for (int32_t s = 0; s < b->pairNJobs[i]; s++) {
            const int32_t q = b->order[b->pairFirstJob[i] + s];
            if (q < 0) continue;
            for (int k = 0; k < 10; k++) t[k] += ex[(size_t)q * 10 + k];
        }
        for (int k = 0; k < 9; k++) trans[9 * i + k] = t[k];
        if (likelihood) likelihood[i] = t[9];
    }
    return MP_OK;
}

extern "C" mp_status mp_batch_fetch_emission_counts(mp_batch *b, double *match16, double *gapX4, double *gapY4,
                                                    double *repeat) {
    if (!b) {
        mp::set_error("mp_batch_fetch_emission_counts: null batch");
        return MP_EINVAL;
    }
    if (!b->ran || b->mode != MP_MODE_EXPECTATIONS) {
        mp::set_error("mp_batch_fetch_emission_counts: batch has not run in expectation mode");
        return MP_EINVAL;
    }
    std::vector<double> h(kTabFloats, 0.0);
    if (!b->jobs.empty()) {
        HIP_TRY(hipSetDevice(b->eng->device));
        HIP_TRY(hipMemcpy(h.data(), b->dEmit.p, sizeof(double) * kTabFloats, hipMemcpyDeviceToHost));
    }
    // table entries -> forward-strand parameters: the reverse strand reads the
    // reverse complement (M'[i][j] = M[3-i][3-j], G'[i] = G[3-i]); N entries are
    // constants, not parameters; run-length rows are already strand-resolved
    if (match16)
        for (int i = 0; i < 16; i++) match16[i] = 0.0;
    if (gapX4)
        for (int i = 0; i < 4; i++) gapX4[i] = 0.0;
    if (gapY4)
        for (int i = 0; i < 4; i++) gapY4[i] = 0.0;
    for (int st = 0; st < 2; st++)
        for (int i = 0; i < 4; i++) {
            const int fi = st ? i : 3 - i;
            if (gapX4) gapX4[fi] += h[kTabGX + st * 5 + i];
            if (gapY4) gapY4[fi] += h[kTabGY + st * 5 + i];
            for (int j = 0; j < 4; j++) {
                const int fj = st ? j : 3 - j;
                if (match16) match16[fi * 4 + fj] += h[kTabE + st * 25 + i * 5 + j];
            }
        }
    if (repeat)
        for (int i = 0; i < 4 * kRL * kRL; i++) repeat[i] = h[kTabRL + i];
    return MP_OK;
}

extern "C" mp_status mp_pairhmm_forward_log_prob(mp_engine *e, const mp_model *m, const mp_pair *pair,
                                                 const mp_band_params *p, double *logProb) {
    if (!pair || !logProb) {
        mp::set_error("mp_pairhmm_forward_log_prob: null argument");
        return MP_EINVAL;
    }
    mp_batch *b = nullptr;
    mp_status st = mp_batch_create(e, m, pair, 1, p, MP_MODE_FORWARD_ONLY, 0, &b);
    if (st) return st;
    int64_t counts[3];
    if (!(st = mp_batch_run_plain(b))) st = mp_batch_fetch_counts(b, counts, logProb);
    mp_batch_free(b);
    return st;
}

// computeForwardProbability for one pair with no device allocation: the
// encoded tables, the job and its window, the band entries, the symbols, the
// counters and the value ring of the forward-only kernel live in this device's
// `scratch`; the same kernel as mp_pairhmm_forward_log_prob runs on the
// engine's stream.  Calls on one device share the scratch (and the engine's
// stream): calls that overlap on one device race.
extern "C" mp_status mp_pairhmm_forward_log_prob_noalloc(mp_engine *e, const mp_model *m, const mp_pair *pair,
                                                         const mp_band_params *p, double *logProb) {
    if (!e || !m || !pair || !p || !logProb) {
        mp::set_error("mp_pairhmm_forward_log_prob_noalloc: null argument");
        return MP_EINVAL;
    }
    const mp_pair &pr = *pair;
    if (pr.lX < 0 || pr.lY < 0 || (pr.lX > 0 && !pr.sX) || (pr.lY > 0 && !pr.sY) || pr.nAnchors < 0 ||
        (pr.nAnchors > 0 && !pr.anchors)) {
        mp::set_error("mp_pairhmm_forward_log_prob_noalloc: bad sequence or anchor arguments");
        return MP_EINVAL;
    }
    if (m->mach[1].nStates != 3) {
        mp::set_error("mp_pairhmm_forward_log_prob_noalloc: three-state machines only");
        return MP_EUNSUPPORTED;
    }
    if (p->diagonalExpansion < 0 || (p->diagonalExpansion & 1)) {
        mp::set_error("mp_pairhmm_forward_log_prob_noalloc: diagonalExpansion must be even and non-negative");
        return MP_EINVAL;
    }
    if (!mp::anchors_valid(pr.anchors, pr.nAnchors, pr.lX, pr.lY)) {
        mp::set_error("mp_pairhmm_forward_log_prob_noalloc: anchors must be strictly increasing, inside the "
                      "sequences, with even non-negative expansion");
        return MP_EANCHORS;
    }
    const int64_t D = pr.lX + pr.lY;
    if (D == 0) {
        *logProb = 0.0;
        return MP_OK;
    }
    // scratch layout (256-byte aligned regions)
    auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const int64_t padEnd = (D / kK + 2) * kK;
    const int64_t nsym = kSymPad + 2 + pr.lX + pr.lY + kSymPad;
    const size_t oTab = 0, oJob = up(oTab + sizeof(float) * kTabFloats), oWin = up(oJob + sizeof(DevJob)),
                 oCnt = up(oWin + sizeof(DevWin)), oRing = up(oCnt + 64),
                 oBand = up(oRing + sizeof(float4) * (size_t)(4 * kK) * 64),
                 oSym = up(oBand + sizeof(int32_t) * (size_t)padEnd), oEnd = oSym + sizeof(uint32_t) * (size_t)nsym;
    if (D >= kMaxDiagonals || oEnd > kNoAllocScratchBytes) {
        mp::set_error("mp_pairhmm_forward_log_prob_noalloc: the pair needs " + std::to_string(oEnd) +
                      " bytes of device scratch; the allocation-free entry holds " +
                      std::to_string(kNoAllocScratchBytes) + " (use mp_pairhmm_forward_log_prob)");
        return MP_EUNSUPPORTED;
    }
    // host staging, reused across calls on this thread
    thread_local std::vector<uint32_t> sym;
    thread_local std::vector<int32_t> band;
    thread_local std::vector<float> tab;
    thread_local std::vector<DevWin> wj;
    thread_local PlanScratch sc;
    const int strand = pr.strand ? 1 : 0;
    sym.assign((size_t)nsym, enc_x(4, 0));
    sym[0] = kZeroSym;
    const int64_t baseX = kSymPad + 1, baseY = baseX + pr.lX + 1;
    sym[kSymPad] = enc_x(4, strand);
    for (int64_t x = 0; x < pr.lX; x++) sym[baseX + x] = enc_x(pr.sX[x], strand);
    sym[baseY - 1] = enc_y(4, strand);
    for (int64_t y = 0; y < pr.lY; y++) sym[baseY + y] = enc_y(pr.sY[y], strand);
    HostJob hj{};
    hj.lX = (int32_t)pr.lX;
    hj.lY = (int32_t)pr.lY;
    hj.symX = baseX;
    hj.symY = baseY;
    hj.flags = strand | (pr.raggedLeft ? 2 : 0) | (pr.raggedRight ? 4 : 0);
    band.resize((size_t)padEnd);
    if (!plan_job(hj, JobAnc{pr.anchors, pr.nAnchors, 0, 0}, *p, MP_MODE_FORWARD_ONLY, false, band.data(), padEnd, sc,
                  wj)) {
        mp::set_error("mp_pairhmm_forward_log_prob_noalloc: invalid band (anchors inconsistent with the sequences)");
        return MP_EANCHORS;
    }
    if (hj.cls > 2) {
        mp::set_error("mp_pairhmm_forward_log_prob_noalloc: band width " + std::to_string(hj.wmax) +
                      " exceeds the 64 cells of the allocation-free entry (use mp_pairhmm_forward_log_prob)");
        return MP_EUNSUPPORTED;
    }
    KParams kp;
    std::memset(&kp, 0, sizeof(kp));
    build_tables(m, false, tab, kp);
    DevJob dj{};
    dj.symX = hj.symX;
    dj.symY = hj.symY;
    dj.bandOff = 0;
    dj.lX = hj.lX;
    dj.lY = hj.lY;
    dj.winOff = 0;
    dj.nWin = 1;
    dj.flags = hj.flags;
    unsigned char *sp = e->scratch;
    kp.tables = reinterpret_cast<const float *>(sp + oTab);
    kp.jobs = reinterpret_cast<const DevJob *>(sp + oJob);
    kp.win = reinterpret_cast<const DevWin *>(sp + oWin);
    int32_t *cnt = reinterpret_cast<int32_t *>(sp + oCnt);  // queue[4], error, counts[2], pad, logTotal
    kp.queue = cnt;
    kp.error = cnt + 4;
    kp.counts = cnt + 5;
    kp.logTotal = reinterpret_cast<int2 *>(cnt + 8);
    kp.ring = reinterpret_cast<float *>(sp + oRing);
    kp.ringLen = 4 * kK;
    kp.band = reinterpret_cast<const int *>(sp + oBand);
    kp.sym = reinterpret_cast<const uint32_t *>(sp + oSym);
    kp.mode = MP_MODE_FORWARD_ONLY;
    kp.threshold = (float)p->threshold;
    for (int c = 0; c < kNumClasses; c++) kp.jobBegin[c] = kp.jobEnd[c] = 1;
    kp.jobBegin[hj.cls] = 0;
    kp.wideBegin = kp.wideEnd = 1;
    HIP_TRY(hipSetDevice(e->device));
    hipStream_t s = e->stream;
    HIP_TRY(hipMemcpyAsync(sp + oTab, tab.data(), sizeof(float) * kTabFloats, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(sp + oJob, &dj, sizeof(DevJob), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(sp + oWin, wj.data(), sizeof(DevWin), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(sp + oBand, band.data(), sizeof(int32_t) * (size_t)padEnd, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(sp + oSym, sym.data(), sizeof(uint32_t) * (size_t)nsym, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, s, cnt, 0, 10);  // counters, error, total
    hipLaunchKernelGGL(pairhmm_seq_kernel, dim3(1), dim3(64), kSeqLdsBytes, s, kp);
    HIP_TRY(hipGetLastError());
    int32_t back[10];
    HIP_TRY(hipMemcpyAsync(back, cnt, sizeof(back), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (back[4]) {
        mp::set_error("mp_pairhmm_forward_log_prob_noalloc: kernel error code " + std::to_string(back[4]));
        return MP_EHIP;
    }
    float t;
    std::memcpy(&t, &back[8], sizeof(float));
    *logProb = std::log((double)t) + (double)back[9] * 0.69314718055994530942;
    return MP_OK;
}

extern "C" mp_status mp_pairhmm_expectations(mp_engine *e, const mp_model *m, const mp_pair *pair,
                                             const mp_band_params *p, double *trans9, double *likelihood) {
    if (!pair || !trans9) {
        mp::set_error("mp_pairhmm_expectations: null argument");
        return MP_EINVAL;
    }
    mp_batch *b = nullptr;
    mp_status st = mp_batch_create(e, m, pair, 1, p, MP_MODE_EXPECTATIONS, 0, &b);
    if (st) return st;
    if (!(st = mp_batch_run_plain(b))) st = mp_batch_fetch_expectations(b, trans9, likelihood);
    mp_batch_free(b);
    return st;
}

extern "C" void mp_batch_free(mp_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->eng->device);
    drop_graph(b);
    delete b;
}

extern "C" mp_status mp_pairhmm_posteriors(mp_engine *e, const mp_model *m, const mp_pair *pair,
                                           const mp_band_params *p, int32_t mode, mp_aligned_pair *matches,
                                           int64_t matchCap, int64_t *matchCount, mp_aligned_pair *gapX,
                                           int64_t gapXCap, int64_t *gapXCount, mp_aligned_pair *gapY,
                                           int64_t gapYCap, int64_t *gapYCount, double *logProb) {
    if (!pair || !matchCount) {
        mp::set_error("mp_pairhmm_posteriors: null argument");
        return MP_EINVAL;
    }
    mp_batch *b = nullptr;
    mp_status st = mp_batch_create(e, m, pair, 1, p, mode, 0, &b);
    if (st) return st;
    if ((st = mp_batch_run_plain(b))) {
        mp_batch_free(b);
        return st;
    }
    int64_t counts[3];
    double lp = 0.0;
    if ((st = mp_batch_fetch_counts(b, counts, &lp))) {
        mp_batch_free(b);
        return st;
    }
    mp_aligned_pair *dst[3] = {matches, gapX, gapY};
    int64_t caps[3] = {matchCap, gapXCap, gapYCap};
    int64_t *cnt[3] = {matchCount, gapXCount, gapYCount};
    for (int l = 0; l < 3; l++) {
        if (cnt[l]) *cnt[l] = counts[l];
        if (!dst[l] || counts[l] == 0) continue;
        if ((st = copy_list(b, l, dst[l], std::min(counts[l], caps[l])))) {
            mp_batch_free(b);
            return st;
        }
    }
    if (logProb) *logProb = lp;
    mp_batch_free(b);
    return MP_OK;
}
