// mp_split.hpp -- posterior kernel with concurrent forward and backward
// sweeps (gfx950).
//
// Reference algorithm: getPosteriorProbsWithBanding (pairwiseAligner.c:687-808)
// with stateMachine3_cellCalculate (stateMachine.c:559-581), RLE emissions
// (stateMachine.c:702-736) and posterior emission (pairwiseAligner.c:586-664).
//
// Each job (one read/reference sub-problem) is held by a row of SEG lanes in
// a forward wavefront and the same row of its partner backward wavefront in
// the same workgroup.  Per traceback window the forward row walks up from the
// previous window's end and the backward row walks down from tb at the same
// time; they meet at the crossing diagonal cx, where the window total is
// formed from the cut between cx and cx+1:
//     total = sum_{cells of cx+1} in_F . (e . B)  +  sum_{cells of cx+2} inM_F . eM . B_M
// (every path crosses that cut by exactly one gap move cx -> cx+1 or one
// match move cx-1 -> cx+1 or cx -> cx+2; in_F are the forward inputs gathered
// from cx and cx-1, e . B the backward sends of cx+1 and cx+2).  In exact
// arithmetic this equals the reference's window total.  After the crossing
// the forward row emits posteriors for (cx, from] and the backward row for
// (lo, cx]; both continue their recursions.  A window's critical path is one
// sweep instead of two.
//
// Checkpoints instead of stored values: before the crossing each sweep
// stores, for every 8-diagonal cell it enters, its state on entry (the sends
// of the two diagonals before the cell: 16 bytes per lane) in an HBM ring.
// After the crossing a row rebuilds its partner's values of the current cell
// from that checkpoint, running the partner's recursion over the cell with the
// same operations, rescale points and start vectors as the partner sweep, so
// the rebuilt values are the partner's values bit for bit; the posterior steps
// then read them from registers.  HBM carries 2 bytes per cell and lane each
// way instead of 12.
//
// Anti-diagonals are walked in grid-aligned macro steps of 8 (one "cell" of
// the grid d = 8g..8g+7); every per-step operand (band entry, encoded
// symbols) and the partner's checkpoint are registers loaded one macro step
// ahead (the band two).  Macro steps that cover all 8 anti-diagonals of every
// running row with one-lane neighbour shifts (the common case) run a
// branch-free body; while every row runs they follow each other in an inner
// loop.
//
// Rows of 16 lanes (the common case) hold cell x of an anti-diagonal on lane
// x mod 16 ("x-aligned"); when every three consecutive anti-diagonals of the
// band fit in 16 consecutive x (checked on the host) each neighbour input is
// a fixed one-lane rotation inside the row.  Rows of 32 and 64 lanes hold
// cell k on lane k and shift by band geometry.  Cells outside the band read
// a zero-emission symbol, so they send zero.
//
// Output: one record per cell with any posterior >= threshold:
// (pM, pX, pY, (d << 11) | cell); the host orders the records of each
// anti-diagonal.
#pragma once

#include "mp_device.hpp"

namespace mpk {

struct SplitShared {
    float xv[kMaxPairs][2][2][4][64];            // crossing values [pair][slot][side][term][lane]
    int xexp[kMaxPairs][2][2][4];                // [pair][slot][side][row]
    int xflag[kNumClasses][kMaxPairs][2][4];     // crossings published [cls][pair][side][row]
    int jobId[kNumClasses][kMaxPairs][4];        // job published by the forward row
    int jobSeq[kNumClasses][kMaxPairs][4];       // jobs published
    int doneB[kNumClasses][kMaxPairs][4];        // jobs finished by the backward row
    float vec[2][2][3];                          // [start / end][plain / ragged][state]: the boundary vectors
};
constexpr int kSplitLdsBytes =
    ((kTabFloats * 4 + (int)sizeof(SplitShared)) > kTabReach ? (kTabFloats * 4 + (int)sizeof(SplitShared))
                                                              : kTabReach);
// Checkpoint ring: [wave slot][cell][64 lanes] of 4 floats (the sweep's sends
// on entering the cell), so that one cell's 64 lanes store and load 1024
// contiguous bytes.
constexpr int kRingCellBytes = 64 * 16;

enum { S_FETCH = 0, S_RUN = 1, S_XWAIT = 2, S_JWAIT = 3, S_EXIT = 4 };

__device__ __forceinline__ int lds_acquire(int *p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_release(int *p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

struct F3 {
    float x, y, z;
};

// Neighbour inputs.  Cell k of anti-diagonal d is (x, y) = (xL + k, d - xL - k).
// Forward: X from (x-1, y) and Y from (x, y-1) on d-1, M from (x-1, y-1) on d-2.
// Backward: X from (x+1, y) and Y from (x, y+1) on d+1, M from (x+1, y+1) on d+2.
// xL1 / xL2 are the band starts of those diagonals; the fast forms assume
// the one-lane shifts that hold on every cell flagged fast by the host.
template <int SEG>
__device__ __forceinline__ void fwd_in_fast(int xL, int xL1, int xL2, float oX1, float oY1, float oM2, int k,
                                            float &inX, float &inY, float &inM) {
    const bool same1 = xL == xL1;  // X from lane k-1, Y from lane k; else X from k, Y from k+1
    const int dm = xL - xL2;       // M from lane k-1 (0), k (1), k+1 (2)
    const float xs = row_shr<SEG>(oX1, k), ys = row_shl<SEG>(oY1, k);
    const float ml = row_shr<SEG>(oM2, k), mr = row_shl<SEG>(oM2, k);
    inX = same1 ? xs : oX1;
    inY = same1 ? oY1 : ys;
    inM = dm == 0 ? ml : (dm == 2 ? mr : oM2);
}
__device__ __forceinline__ void fwd_in_slow(int xL, int b1, int b2, float oX1, float oY1, float oM2, bool act,
                                            int rowBase, int k, float &inX, float &inY, float &inM) {
    const int xL1 = band_xL(b1), xL2 = band_xL(b2);
    inX = gather(oX1, k + xL - xL1 - 1, band_wm1_narrow(b1), act, rowBase, k);
    inY = gather(oY1, k + xL - xL1, band_wm1_narrow(b1), act, rowBase, k);
    inM = gather(oM2, k + xL - xL2 - 1, band_wm1_narrow(b2), act, rowBase, k);
}
template <int SEG>
__device__ __forceinline__ void bwd_in_fast(int xL, int xL1, int xL2, float oX1, float oY1, float oM2, int k,
                                            float &inX, float &inY, float &inM) {
    const bool same1 = xL == xL1;  // X from lane k+1, Y from lane k; else X from k, Y from k-1
    const int dm = xL - xL2;       // M from lane k-1 (-2), k (-1), k+1 (0)
    const float xs = row_shl<SEG>(oX1, k), ys = row_shr<SEG>(oY1, k);
    const float ml = row_shr<SEG>(oM2, k), mr = row_shl<SEG>(oM2, k);
    inX = same1 ? xs : oX1;
    inY = same1 ? oY1 : ys;
    inM = dm == -2 ? ml : (dm == 0 ? mr : oM2);
}
__device__ __forceinline__ void bwd_in_slow(int xL, int b1, int b2, float oX1, float oY1, float oM2, bool act,
                                            int rowBase, int k, float &inX, float &inY, float &inM) {
    const int xL1 = band_xL(b1), xL2 = band_xL(b2);
    inX = gather(oX1, k + xL - xL1 + 1, band_wm1_narrow(b1), act, rowBase, k);
    inY = gather(oY1, k + xL - xL1, band_wm1_narrow(b1), act, rowBase, k);
    inM = gather(oM2, k + xL - xL2 + 1, band_wm1_narrow(b2), act, rowBase, k);
}

// BWD: the wavefront's role (compiled separately, so each role's code holds
// only its own sweep, rebuild and posterior bodies).
template <int SEG, bool BWD>
__device__ __forceinline__ void split_class(const KParams &kp, const float *tab, SplitShared *sh, int cls) {
    constexpr int K = kK;
    constexpr bool XAL = SEG == 16;  // x-aligned lanes
    // Every parameter is read once here: the loop's acquire loads would otherwise
    // force the compiler to re-read the parameter block after each of them.
    const DevJob *const gJobs = kp.jobs;
    const DevWin *const gWin = kp.win;
    int4 *const gOut = kp.out;
    int32_t *const gCounts = kp.counts;
    int32_t *const gWinEnd = kp.winEnd;
    int2 *const gLogTotal = kp.logTotal;
    int2 *const gChain = kp.chain;
    const int gTbd = kp.tbd;
    int32_t *const gQueue = kp.queue + cls;
    int32_t *const gError = kp.error;
    const int gJobBegin = kp.jobBegin[cls], gJobEnd = kp.jobEnd[cls];
    const int32_t *const gJobMap = kp.jobMap;
    // (the boundary vectors, used at sweep starts only, are read from LDS)
    const float(*const sv)[3] = sh->vec[0];
    const float(*const ev)[3] = sh->vec[1];
    const int lane = threadIdx.x & 63;
    const int k = lane & (SEG - 1);
    const int row = lane / SEG;
    const int rowBase = lane & ~(SEG - 1);
    const int wib = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int pr = wib >> 1;
    constexpr bool bwd = BWD;
    const size_t slot = (size_t)blockIdx.x * (blockDim.x >> 6) + wib;
    const int RC = kp.ringLen / K;
    char *myRing = reinterpret_cast<char *>(kp.ring) + slot * (size_t)RC * kRingCellBytes;
    const char *peerRing = reinterpret_cast<const char *>(kp.ring) + (slot ^ 1) * (size_t)RC * kRingCellBytes;
    int *myExp = kp.ringExp + slot * (size_t)RC * 4;
    const int *peerExp = kp.ringExp + (slot ^ 1) * (size_t)RC * 4;
    const uint32_t laneRing = (uint32_t)lane * 16u;
    // masked ring stores go through a buffer resource (out-of-range offsets are dropped)
    const __amdgpu_buffer_rsrc_t ringRs = __builtin_amdgcn_make_buffer_rsrc(myRing, 0, RC * kRingCellBytes, kBufWord3);
    const char *symB = reinterpret_cast<const char *>(kp.sym);
    const char *bandB = reinterpret_cast<const char *>(kp.band);
    const float T00 = kp.T[0], T01 = kp.T[1], T02 = kp.T[2], T10 = kp.T[3], T11 = kp.T[4], T12 = kp.T[5],
                T20 = kp.T[6], T21 = kp.T[7], T22 = kp.T[8];
    // wave-uniform thresholds (scalar registers): a record needs pM >= thr or,
    // with indels, pX or pY >= thr
    const float thr = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(kp.threshold)));
    const float thrXY = __int_as_float(
        __builtin_amdgcn_readfirstlane(kp.withIndels ? __float_as_int(kp.threshold) : 0x7f800000));
    int *xflagMine = &sh->xflag[cls][pr][bwd ? 1 : 0][row];
    int *xflagPeer = &sh->xflag[cls][pr][bwd ? 0 : 1][row];
    int *jobSeqP = &sh->jobSeq[cls][pr][row];
    int *jobIdP = &sh->jobId[cls][pr][row];
    int *doneBP = &sh->doneB[cls][pr][row];

    int phase = bwd ? S_JWAIT : S_FETCH;
    int jn = 0, xn = 0;
    int job = -1, D = 0, nWin = 0, winOff = 0, jflags = 0, cap = 0, win = 0;
    uint32_t xrow = 0, yrow = 0, bandRow = 0;  // byte offsets: x/y symbol bases, the job's band
    int64_t outBase = 0;
    int tb = 0, lo = 0, from = 0, cx = 0;
    int capD = -1;  // restarted window job: the previous window's traceback anti-diagonal (chain value)
    int fst = 0;    // the job's forward start anti-diagonal (a cell boundary)
    int d = 0, gs = 0, S = 0, Stot = 0, cnt = 0, cpos = 0;
    bool post = false, init = false, reload = true;
    float totalInv = 0.f;
    float oX1 = 0.f, oY1 = 0.f, oM1 = 0.f, oM2 = 0.f;
    int bm1 = kEmptyBand, bm2 = kEmptyBand;  // band entries of the previous two processed diagonals
    int bC[K], bN[K];
    uint32_t sxC[K], syC[K];
    float4 ckC = make_float4(0.f, 0.f, 0.f, 0.f), ckN = ckC;  // partner checkpoints: current and next cell
    int eC = 0;
#pragma unroll
    for (int j = 0; j < K; j++) {
        bC[j] = bN[j] = 0;
        sxC[j] = syC[j] = 0u;
    }

    // cell index of this lane on an anti-diagonal starting at xL
    auto cellOf = [&](int xL) { return XAL ? ((k - xL) & 15) : k; };
    auto bandCell = [&](int c, int *dst) {
        const int4 *p = reinterpret_cast<const int4 *>(bandB + bandRow + (uint32_t)c * 32u);
        const int4 a = p[0], b = p[1];
        dst[0] = a.x; dst[1] = a.y; dst[2] = a.z; dst[3] = a.w;
        dst[4] = b.x; dst[5] = b.y; dst[6] = b.z; dst[7] = b.w;
    };
    // symbol byte offsets of this lane's cell on anti-diagonal dj (band entry b);
    // cells outside the band read the zero symbol at offset 0
    auto symAddr = [&](int dj, int b, uint32_t &xa, uint32_t &ya) {
        const int xL = band_xL(b), kx = cellOf(xL);
        const bool in = kx <= band_wm1_narrow(b);
        const uint32_t x = (uint32_t)(xL + kx);
        xa = in ? xrow + (x << 2) : 0u;
        ya = in ? yrow + (((uint32_t)dj - x) << 2) : 0u;
    };
    auto symCell = [&](int c, const int *bb, uint32_t *sx, uint32_t *sy) {
#pragma unroll
        for (int j = 0; j < K; j++) {
            uint32_t xa, ya;
            symAddr(c * K + j, bb[j], xa, ya);
            sx[j] = *reinterpret_cast<const uint32_t *>(symB + xa);
            sy[j] = *reinterpret_cast<const uint32_t *>(symB + ya);
        }
    };
    auto peerCkpt = [&](int cp) {
        return *reinterpret_cast<const float4 *>(peerRing + ((uint32_t)cp * kRingCellBytes + laneRing));
    };
    auto resetExchange = [&]() {
        oX1 = oY1 = oM1 = oM2 = 0.f;
        bm1 = bm2 = kEmptyBand;
    };
    auto loadWin = [&]() {
        const DevWin w = gWin[winOff + win];
        tb = w.tb;
        lo = w.lo;
        from = w.from;
        cx = w.cx;
    };
    auto startBwdWindow = [&]() {
        d = tb;
        gs = tb & ~(K - 1);
        S = 0;
        init = true;
        post = false;
        reload = true;
        resetExchange();
    };
    auto setupJob = [&]() {
        const DevJob jb = gJobs[job];
        xrow = (uint32_t)(jb.symX - 1) * 4u;
        yrow = (uint32_t)(jb.symY - 1) * 4u;
        bandRow = (uint32_t)jb.bandOff * 4u;
        D = jb.lX + jb.lY;
        winOff = jb.winOff;
        nWin = jb.nWin;
        jflags = jb.flags;
        cap = bwd ? jb.cap[1] : jb.cap[0];
        outBase = jb.outOff + (bwd ? jb.cap[0] : 0);
        fst = jb.fstart;
        cnt = 0;
        win = 0;
        loadWin();
        if (!bwd) {
            d = jb.fstart;
            gs = d;
            S = 0;
            init = true;
            post = false;
            reload = true;
            resetExchange();
            const bool restart = (jflags & kJobRestart) != 0;
            capD = restart ? lo + gTbd + 1 : -1;
        } else {
            startBwdWindow();
        }
        phase = S_RUN;
    };
    // Crossing values of this side: forward = inputs of cx+1 (X, Y, M) and the
    // match input of cx+2; backward = sends of cx+1 (X, Y, M) and of cx+2 (M).
    auto publishCrossing = [&]() {
        float v0, v1, v2, v3;
        if (!bwd) {
            if constexpr (XAL) {
                // No masking: lanes outside the bands of cx+1 / cx+2 meet zero
                // backward sends.  (A predicated DPP would also read 0 from the
                // disabled source lanes.)
                v0 = dpp<kRowRor1, false>(oX1);
                v1 = oY1;
                v2 = dpp<kRowRor1, false>(oM2);
                v3 = dpp<kRowRor1, false>(oM1);
            } else {
                const int *bandJ = reinterpret_cast<const int *>(bandB + bandRow);
                const int b1 = bandJ[cx + 1], b2 = bandJ[cx + 2];
                const int xL1 = band_xL(b1), xL2 = band_xL(b2);
                const bool a1 = k <= band_wm1_narrow(b1), a2 = k <= band_wm1_narrow(b2) && cx + 2 <= D;
                const int xc = band_xL(bm1), xcm = band_xL(bm2), wc = band_wm1_narrow(bm1), wcm = band_wm1_narrow(bm2);
                v0 = gather(oX1, k + xL1 - xc - 1, wc, a1, rowBase, k);
                v1 = gather(oY1, k + xL1 - xc, wc, a1, rowBase, k);
                v2 = gather(oM2, k + xL1 - xcm - 1, wcm, a1, rowBase, k);
                v3 = gather(oM1, k + xL2 - xc - 1, wc, a2, rowBase, k);
            }
        } else {
            v0 = oX1;
            v1 = oY1;
            v2 = oM1;
            v3 = oM2;
        }
        const int sl = xn & 1, side = bwd ? 1 : 0;
        sh->xv[pr][sl][side][0][lane] = v0;
        sh->xv[pr][sl][side][1][lane] = v1;
        sh->xv[pr][sl][side][2][lane] = v2;
        sh->xv[pr][sl][side][3][lane] = v3;
        if (k == 0) sh->xexp[pr][sl][side][row] = S;
        wait_vm();  // this side's ring values are in memory before the flag
        if (k == 0) lds_release(xflagMine, xn + 1);
        phase = S_XWAIT;
    };
    auto bwdWindowEnd = [&]() {
        if (k == 0) gWinEnd[2 * (winOff + win) + 1] = cnt;
        if (win + 1 == nWin) {
            if (k == 0) {
                gCounts[2 * job + 1] = cnt;
                lds_release(doneBP, jn);
            }
            phase = S_JWAIT;
        } else {
            win += 1;
            loadWin();
            startBwdWindow();
        }
    };

    int idle = 0;
    for (;;) {
        // ---------------- per-row protocol ----------------
        // Forward rows take jobs as a gang: once no row of the wavefront holds
        // a job, every fetching row takes the next job of the queue in one
        // claim.  Neighbouring jobs of the queue have about the same band area,
        // so the rows of a wavefront reach their crossings (where the posterior
        // steps begin) at about the same macro step, and the wavefront runs the
        // cheaper body without posteriors for the first half of its jobs.
        const uint64_t fetchRows = __ballot(phase == S_FETCH && k == 0);
        const bool gangReady = __all(phase == S_FETCH || phase == S_EXIT);
        if (phase == S_FETCH && gangReady) {  // forward rows only
            int claim = 0;
            if (lane == __builtin_ctzll(fetchRows)) claim = atomicAdd(gQueue, __popcll(fetchRows));
            claim = __builtin_amdgcn_readfirstlane(claim);
            const int jnew = claim + __popcll(fetchRows & ((1ull << rowBase) - 1ull));
            job = gJobBegin + jnew;
            if (job >= gJobEnd) job = -1;
            else if (gJobMap) job = gJobMap[job];
            if (k == 0) {
                *jobIdP = job;
                lds_release(jobSeqP, jn + 1);
            }
            jn += 1;
            if (job < 0) phase = S_EXIT;
            else setupJob();
        } else if (phase == S_JWAIT) {
            if (!bwd) {
                if (lds_acquire(doneBP) >= jn) phase = S_FETCH;
            } else if (lds_acquire(jobSeqP) >= jn + 1) {
                job = *jobIdP;
                jn += 1;
                if (job < 0) phase = S_EXIT;
                else setupJob();
            }
        } else if (phase == S_XWAIT) {
            if (lds_acquire(xflagPeer) > xn) {
                const int sl = xn & 1;
                const float t = rsum<SEG>(sh->xv[pr][sl][0][0][lane] * sh->xv[pr][sl][1][0][lane] +
                                          sh->xv[pr][sl][0][1][lane] * sh->xv[pr][sl][1][1][lane] +
                                          sh->xv[pr][sl][0][2][lane] * sh->xv[pr][sl][1][2][lane] +
                                          sh->xv[pr][sl][0][3][lane] * sh->xv[pr][sl][1][3][lane]);
                Stot = sh->xexp[pr][sl][0][row] + sh->xexp[pr][sl][1][row];
                totalInv = t > 0.f ? 1.f / t : 0.f;
                if (!bwd && win + 1 == nWin && k == 0) gLogTotal[job] = make_int2(__float_as_int(t), Stot);
                xn += 1;
                post = true;
                reload = true;
                phase = S_RUN;
                if (!bwd) {
                    d = cx + 1;
                } else {
                    d = cx;
                    if (cx < lo + 1) bwdWindowEnd();
                }
                gs = d & ~(K - 1);
            }
        }
        if (__all(phase == S_EXIT)) break;
        const bool running = phase == S_RUN;
        if (!__any(running)) {
            if (++idle > kSpinLimit || __hip_atomic_load(gError, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                if (lane == 0) atomicExch(gError, 1);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
            continue;
        }
        idle = 0;
        const bool allRunning = __all(running);
        if (!running) continue;
        // While every row of the wavefront runs, macro steps follow each other in
        // this inner loop, their state in the same registers, until some row
        // reaches its stop diagonal and the row protocol takes over; with a row
        // waiting, the protocol runs after every macro step.
        bool reached = false;
        do {
        F3 rC[K];  // the partner's values of this cell, rebuilt from its checkpoint when some row emits
        const int stop = bwd ? (post ? lo + 1 : cx + 1) : (post ? tb : cx);
        // a macro step that cannot reach the row's stop diagonal means a broken
        // invariant: report it instead of walking off the band
        const bool broken = bwd ? (gs + K - 1 < stop || gs < 0) : (gs > stop);
        if (__any(broken)) {
            if (broken) {
                atomicExch(gError, 2);
                phase = S_EXIT;
            }
            break;
        }

        // ---------------- one macro step (8 anti-diagonals of one grid cell) ----------------
        const int cell = gs / K;
        const int lastCell = D / K + 1;  // bands are padded by one extra cell
        const int cN = bwd ? max(cell - 1, 0) : min(cell + 1, lastCell);
        const int cN2 = bwd ? max(cell - 2, 0) : min(cell + 2, lastCell);
        if (reload) {
            cpos = cell % RC;
            bandCell(cell, bC);
            bandCell(cN, bN);
            symCell(cell, bC, sxC, syC);
            ckC = peerCkpt(post ? cpos : 0);
            eC = peerExp[cpos * 4 + row];
            reload = false;
        }
        const int cposN = bwd ? (cpos == 0 ? RC - 1 : cpos - 1) : (cpos + 1 == RC ? 0 : cpos + 1);
        int bN2[K];
        bandCell(cN2, bN2);
        const int eN = peerExp[cposN * 4 + row];
        // (rows not emitting posteriors read one fixed checkpoint slot, a cache hit)
        ckN = peerCkpt(post ? cposN : 0);
        // rolling prefetch: after step j consumes its symbols, the same registers
        // are reloaded for step j of the next cell
        // Loads are never predicated (a predicated load would be copied into its
        // register with a wait for every load in flight): cells outside the next
        // band read the symbol pool's first entry, a cache hit.
        auto prefetchStep = [&](auto J, auto) {
            constexpr int j = decltype(J)::value;
            uint32_t xa, ya;
            symAddr(cN * K + j, bN[j], xa, ya);
            sxC[j] = *reinterpret_cast<const uint32_t *>(symB + xa);
            syC[j] = *reinterpret_cast<const uint32_t *>(symB + ya);
        };

        const bool slowRow = (bC[0] & (bwd ? kSlowB : kSlowF)) != 0;
        // (the macro step holding a restarted job's chain anti-diagonal takes the general body)
        const bool fullRow = !init && !slowRow &&
                             (bwd ? (d == gs + K - 1 && stop <= gs) : (d == gs && stop >= gs + K - 1)) &&
                             (capD < gs || capD >= gs + K);
        const bool full = __all(fullRow);
        const bool anyPost = __any(post);
        const bool slowWave = __any(slowRow);
        const uint32_t cbase = (uint32_t)cpos * kRingCellBytes + laneRing;
        int *expSlot = myExp + cpos * 4 + row;

        // This sweep's checkpoint on entering the cell: read by the partner after
        // the crossing (and, going forward, by the next window's backward row
        // above `from`; a window job's warm-up at or below lo is never read).
        {
            const bool enter = bwd ? (d == gs + K - 1) & !post
                                   : (d == gs) & (!post | (gs + K - 1 > from)) & (gs + K - 1 > lo);
            buf_store4(ringRs, enter ? cbase : kBufOff, oX1, oY1, oM1, oM2);
        }

        // The partner's values of this cell, rebuilt from its checkpoint for the
        // rows emitting posteriors: the partner's own recursion, rescale points
        // (the forward at the cell's first diagonal, the backward at its last
        // unless the sweep starts in the cell) and start vectors, so the values
        // are the ones the partner computed.  FR: every emitting row needs the
        // whole cell and no sweep starts in it.
        auto rebuildBackward = [&](auto FRc) {  // forward rows: the backward sweep of this window
            constexpr bool FR = decltype(FRc)::value;
            const bool initCell = tb <= gs + K - 1;  // the backward sweep starts at tb, in this cell
            float uX1 = ckC.x, uY1 = ckC.y, uM1 = ckC.z, uM2 = ckC.w;
            int ub1 = bN[0], ub2 = bN[1];
            if (!FR && initCell) {
                uX1 = uY1 = uM1 = uM2 = 0.f;
                ub1 = ub2 = kEmptyBand;
            }
            const int hiB = min(gs + K - 1, tb), loB = max(gs, d);
            const bool rg = tb == D && ((jflags >> 2) & 1);
            const bool slowB = !XAL && __any(post && (bC[0] & kSlowB) != 0);
            static_for<0, K>([&](auto J) {
                constexpr int j = K - 1 - decltype(J)::value;
                const int dj = gs + j;
                const Emis e = emissions(tab, sxC[j], syC[j]);
                if (!FR && !(post && dj <= hiB && dj >= loB)) return;
                const int b = bC[j];
                const int xL = band_xL(b);
                const int kx = cellOf(xL);
                const bool act = kx <= band_wm1_narrow(b);
                float inX, inY, inM;
                if constexpr (XAL) {
                    inX = dpp<kRowRor15, true>(uX1);
                    inY = uY1;
                    inM = dpp<kRowRor15, true>(uM2);
                } else if (!slowB) {
                    bwd_in_fast<SEG>(xL, band_xL(ub1), band_xL(ub2), uX1, uY1, uM2, k, inX, inY, inM);
                } else {
                    bwd_in_slow(xL, ub1, ub2, uX1, uY1, uM2, act, rowBase, k, inX, inY, inM);
                }
                float q0 = fmaf(T02, inY, fmaf(T01, inX, T00 * inM));
                float q1 = fmaf(T12, inY, fmaf(T11, inX, T10 * inM));
                float q2 = fmaf(T22, inY, fmaf(T21, inX, T20 * inM));
                if (!FR && initCell && dj == tb) {
                    q0 = ev[rg][0];
                    q1 = ev[rg][1];
                    q2 = ev[rg][2];
                }
                if (j == K - 1 && (FR || !initCell)) {
                    const float m = rmaxpos<SEG>(fmaxf(q0, fmaxf(q1, q2)));
                    const int e2 = __builtin_amdgcn_frexp_expf(m);
                    q0 = ldexpf(q0, -e2);
                    q1 = ldexpf(q1, -e2);
                    q2 = ldexpf(q2, -e2);
                    uM1 = ldexpf(uM1, -e2);
                }
                rC[j] = F3{q0, q1, q2};
                uM2 = uM1;
                uX1 = e.eX * q1;
                uY1 = e.eY * q2;
                uM1 = e.eM * q0;
                ub2 = ub1;
                ub1 = b;
            });
        };
        auto rebuildForward = [&](auto FRc) {  // backward rows: the job's forward sweep
            constexpr bool FR = decltype(FRc)::value;
            const bool initCell = fst == gs;  // the forward sweep starts at this cell's first diagonal
            float uX1 = ckC.x, uY1 = ckC.y, uM1 = ckC.z, uM2 = ckC.w;
            int ub1 = bN[K - 1], ub2 = bN[K - 2];
            if (!FR && initCell) {
                uX1 = uY1 = uM1 = uM2 = 0.f;
                ub1 = ub2 = kEmptyBand;
            }
            const int hiF = min(gs + K - 1, d);
            const bool restart = (jflags & kJobRestart) != 0, rg = (jflags >> 1) & 1;
            const bool slowF = !XAL && __any(post && (bC[0] & kSlowF) != 0);
            static_for<0, K>([&](auto J) {
                constexpr int j = decltype(J)::value;
                const int dj = gs + j;
                const Emis e = emissions(tab, sxC[j], syC[j]);
                if (!FR && !(post && dj <= hiF)) return;
                const int b = bC[j];
                const int xL = band_xL(b);
                const int kx = cellOf(xL);
                const bool act = kx <= band_wm1_narrow(b);
                float inX, inY, inM;
                if constexpr (XAL) {
                    inX = dpp<kRowRor1, true>(uX1);
                    inY = uY1;
                    inM = dpp<kRowRor1, true>(uM2);
                } else if (!slowF) {
                    fwd_in_fast<SEG>(xL, band_xL(ub1), band_xL(ub2), uX1, uY1, uM2, k, inX, inY, inM);
                } else {
                    fwd_in_slow(xL, ub1, ub2, uX1, uY1, uM2, act, rowBase, k, inX, inY, inM);
                }
                float p0 = e.eM * inM, p1 = e.eX * inX, p2 = e.eY * inY;
                if (!FR && j == 0 && initCell) {
                    p0 = act ? (restart ? 1.f : sv[rg][0]) : 0.f;
                    p1 = act ? (restart ? 1.f : sv[rg][1]) : 0.f;
                    p2 = act ? (restart ? 1.f : sv[rg][2]) : 0.f;
                }
                if (j == 0) {
                    const float m = rmaxpos<SEG>(fmaxf(p0, fmaxf(p1, p2)));
                    const int e2 = __builtin_amdgcn_frexp_expf(m);
                    p0 = ldexpf(p0, -e2);
                    p1 = ldexpf(p1, -e2);
                    p2 = ldexpf(p2, -e2);
                    uM1 = ldexpf(uM1, -e2);
                }
                rC[j] = F3{p0, p1, p2};
                const float q0 = fmaf(T20, p2, fmaf(T10, p1, T00 * p0));
                const float q1 = fmaf(T21, p2, fmaf(T11, p1, T01 * p0));
                const float q2 = fmaf(T22, p2, fmaf(T12, p1, T02 * p0));
                uM2 = uM1;
                uX1 = q1;
                uY1 = q2;
                uM1 = q0;
                ub2 = ub1;
                ub1 = b;
            });
        };
        if (anyPost) {
            using T_ = std::true_type;
            using F_ = std::false_type;
            if (!bwd) {
                if (__all(!post || ((d == gs) & (tb > gs + K - 1)))) rebuildBackward(T_{});
                else rebuildBackward(F_{});
            } else {
                if (__all(!post || ((d == gs + K - 1) & (fst != gs)))) rebuildForward(T_{});
                else rebuildForward(F_{});
            }
        }

        // One step.  FULL: every running row processes all 8 anti-diagonals with
        // fast shifts (no per-step conditions); POST: some row emits posteriors.
        auto fwdStep = [&](auto J, auto FULLc, auto POSTc, float &sc) {
            constexpr int j = decltype(J)::value;
            constexpr bool FULL = decltype(FULLc)::value, POST = decltype(POSTc)::value;
            const int dj = gs + j;
            if (!FULL && !(dj >= d && dj <= stop)) {
                prefetchStep(J, POSTc);
                return;
            }
            const int b = bC[j];
            const int xL = band_xL(b);
            const int kx = cellOf(xL);
            const bool act = kx <= band_wm1_narrow(b);
            // (bodies with posteriors follow a rebuild, which looked the match emissions up)
            const Emis e = emissions(tab, sxC[j], syC[j]);
            float inX, inY, inM;
            if constexpr (XAL) {
                inX = dpp<kRowRor1, true>(oX1);
                inY = oY1;
                inM = dpp<kRowRor1, true>(oM2);
            } else if (FULL || !slowWave) {
                fwd_in_fast<SEG>(xL, band_xL(bm1), band_xL(bm2), oX1, oY1, oM2, k, inX, inY, inM);
            } else {
                fwd_in_slow(xL, bm1, bm2, oX1, oY1, oM2, act, rowBase, k, inX, inY, inM);
            }
            float p0 = e.eM * inM, p1 = e.eX * inX, p2 = e.eY * inY;
            // (cells outside the band have zero emissions, hence p = 0)
            if (!FULL && j == 0 && init) {
                // the start vector; a restarted window job starts from a flat vector (kWarmup before lo)
                const bool restart = (jflags & kJobRestart) != 0, rg = (jflags >> 1) & 1;
                p0 = act ? (restart ? 1.f : sv[rg][0]) : 0.f;
                p1 = act ? (restart ? 1.f : sv[rg][1]) : 0.f;
                p2 = act ? (restart ? 1.f : sv[rg][2]) : 0.f;
            }
            if (j == 0 && (FULL || dj == d)) {
                // exact power-of-two rescale at the start of every grid cell
                const float m = rmaxpos<SEG>(fmaxf(p0, fmaxf(p1, p2)));
                const int e2 = __builtin_amdgcn_frexp_expf(m);
                p0 = ldexpf(p0, -e2);
                p1 = ldexpf(p1, -e2);
                p2 = ldexpf(p2, -e2);
                oM1 = ldexpf(oM1, -e2);
                S += e2;
                *expSlot = S;
            }
            if constexpr (!FULL) {
                // chain value of a restarted window job: F . endV at the previous window's tb
                const bool capHere = dj == capD;
                if (__any(capHere)) {
                    const float u = rsum<SEG>(fmaf(ev[0][2], p2, fmaf(ev[0][1], p1, ev[0][0] * p0)));
                    if (capHere && k == 0) gChain[job] = make_int2(__float_as_int(u), S);
                }
            }
            if constexpr (POST) if (FULL || anyPost) {  // (the general body also runs without a rebuild)
                if (!FULL || j == 0) sc = ldexpf(totalInv, S + eC - Stot);
                const F3 pb = rC[j];
                const float pM = p0 * pb.x * sc, pX = p1 * pb.y * sc, pY = p2 * pb.z * sc;
                const float pXY = __uint_as_float(max(__float_as_uint(pX), __float_as_uint(pY)));
                const bool emit = act & post & (dj <= from) & ((pM >= thr) | (pXY >= thrXY));
                const uint64_t m = __ballot(emit);
                const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                const int first = rbcast<SEG>(below, rowBase, 0);
                const int last = rbcast<SEG>(below + (emit ? 1 : 0), rowBase, SEG - 1);
                const int idx = cnt + (below - first);
                if (emit & (idx < cap))
                    gOut[outBase + idx] = make_int4(__float_as_int(pM), __float_as_int(pX), __float_as_int(pY),
                                                    (int)rec_code(dj, kx));
                cnt += last - first;
            }
            if (!FULL && dj == stop) reached = true;
            prefetchStep(J, POSTc);
            const float q0 = fmaf(T20, p2, fmaf(T10, p1, T00 * p0));
            const float q1 = fmaf(T21, p2, fmaf(T11, p1, T01 * p0));
            const float q2 = fmaf(T22, p2, fmaf(T12, p1, T02 * p0));
            oM2 = oM1;
            oX1 = q1;
            oY1 = q2;
            oM1 = q0;
            bm2 = bm1;
            bm1 = b;
        };
        auto bwdStep = [&](auto J, auto FULLc, auto POSTc, float &sc) {
            constexpr int j = K - 1 - decltype(J)::value;
            constexpr bool FULL = decltype(FULLc)::value, POST = decltype(POSTc)::value;
            const int dj = gs + j;
            if (!FULL && !(dj <= d && dj >= stop)) {
                prefetchStep(std::integral_constant<int, j>{}, POSTc);
                return;
            }
            const int b = bC[j];
            const int xL = band_xL(b);
            const int kx = cellOf(xL);
            const bool act = kx <= band_wm1_narrow(b);
            // (bodies with posteriors follow a rebuild, which looked the match emissions up)
            const Emis e = emissions(tab, sxC[j], syC[j]);
            float inX, inY, inM;
            if constexpr (XAL) {
                inX = dpp<kRowRor15, true>(oX1);
                inY = oY1;
                inM = dpp<kRowRor15, true>(oM2);
            } else if (FULL || !slowWave) {
                bwd_in_fast<SEG>(xL, band_xL(bm1), band_xL(bm2), oX1, oY1, oM2, k, inX, inY, inM);
            } else {
                bwd_in_slow(xL, bm1, bm2, oX1, oY1, oM2, act, rowBase, k, inX, inY, inM);
            }
            // B = T in
            float q0 = fmaf(T02, inY, fmaf(T01, inX, T00 * inM));
            float q1 = fmaf(T12, inY, fmaf(T11, inX, T10 * inM));
            float q2 = fmaf(T22, inY, fmaf(T21, inX, T20 * inM));
            if (!FULL && init && dj == d) {
                // the end vector (the ragged one at the matrix end of a ragged-right job)
                const bool rg = tb == D && ((jflags >> 2) & 1);
                q0 = ev[rg][0];
                q1 = ev[rg][1];
                q2 = ev[rg][2];
            }
            // (cells outside the band keep q but send e . q = 0; their stored q is never read)
            if (j == K - 1 && (FULL || (dj == d && !init))) {
                const float m = rmaxpos<SEG>(fmaxf(q0, fmaxf(q1, q2)));
                const int e2 = __builtin_amdgcn_frexp_expf(m);
                q0 = ldexpf(q0, -e2);
                q1 = ldexpf(q1, -e2);
                q2 = ldexpf(q2, -e2);
                oM1 = ldexpf(oM1, -e2);
                S += e2;
            }
            // the cell's exponent is final after its first processed diagonal
            if (FULL ? j == K - 1 : dj == d) *expSlot = S;
            if constexpr (POST) if (FULL || anyPost) {  // (the general body also runs without a rebuild)
                if (!FULL || j == K - 1) sc = ldexpf(totalInv, eC + S - Stot);
                const F3 pf = rC[j];
                const float pM = pf.x * q0 * sc, pX = pf.y * q1 * sc, pY = pf.z * q2 * sc;
                const float pXY = __uint_as_float(max(__float_as_uint(pX), __float_as_uint(pY)));
                const bool emit = act & post & (dj <= from) & ((pM >= thr) | (pXY >= thrXY));
                const uint64_t m = __ballot(emit);
                const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                const int first = rbcast<SEG>(below, rowBase, 0);
                const int last = rbcast<SEG>(below + (emit ? 1 : 0), rowBase, SEG - 1);
                const int idx = cnt + (below - first);
                if (emit & (idx < cap))
                    gOut[outBase + idx] = make_int4(__float_as_int(pM), __float_as_int(pX), __float_as_int(pY),
                                                    (int)rec_code(dj, kx));
                cnt += last - first;
            }
            if (!FULL && dj == stop) reached = true;
            prefetchStep(std::integral_constant<int, j>{}, POSTc);
            // send u = e (.) B
            oM2 = oM1;
            oX1 = e.eX * q1;
            oY1 = e.eY * q2;
            oM1 = e.eM * q0;
            bm2 = bm1;
            bm1 = b;
        };
        using T_ = std::true_type;
        using F_ = std::false_type;
        float sc = 0.f;
        if (!bwd) {
            if (full && anyPost) static_for<0, K>([&](auto J) { fwdStep(J, T_{}, T_{}, sc); });
            else if (full) static_for<0, K>([&](auto J) { fwdStep(J, T_{}, F_{}, sc); });
            else static_for<0, K>([&](auto J) { fwdStep(J, F_{}, T_{}, sc); });
            if (full) reached = stop == gs + K - 1;
        } else {
            if (full && anyPost) static_for<0, K>([&](auto J) { bwdStep(J, T_{}, T_{}, sc); });
            else if (full) static_for<0, K>([&](auto J) { bwdStep(J, T_{}, F_{}, sc); });
            else static_for<0, K>([&](auto J) { bwdStep(J, F_{}, T_{}, sc); });
            if (full) reached = stop == gs;
        }
        init = false;
#pragma unroll
        for (int j = 0; j < K; j++) {
            bC[j] = bN[j];
            bN[j] = bN2[j];
        }
        eC = eN;
        ckC = ckN;
        if (!reached) {
            gs += bwd ? -K : K;
            d = bwd ? gs + K - 1 : gs;
            cpos = cposN;
        }
        } while (allRunning && !__any(reached));

        if (reached) {
            if (!post) {
                publishCrossing();
            } else if (!bwd) {
                // forward window end at tb
                if (k == 0) gWinEnd[2 * (winOff + win)] = cnt;
                if (win + 1 == nWin) {
                    if (k == 0) gCounts[2 * job] = cnt;
                    phase = S_JWAIT;
                } else {
                    win += 1;
                    const int next = tb + 1;
                    loadWin();
                    post = false;
                    d = next;
                    gs = d & ~(K - 1);
                    reload = true;
                    if (cx < d) publishCrossing();  // crossing right at the window start
                }
            } else {
                bwdWindowEnd();
            }
        }
    }
}

__global__ __launch_bounds__(512) void pairhmm_split_kernel(KParams kp) {
    extern __shared__ float lds[];
    for (int i = threadIdx.x; i < kTabFloats; i += blockDim.x) lds[i] = kp.tables[i];
    SplitShared *sh = reinterpret_cast<SplitShared *>(lds + kTabFloats);
    int *z = reinterpret_cast<int *>(sh->xflag);
    const int nz = (int)((offsetof(SplitShared, vec) - offsetof(SplitShared, xflag)) / sizeof(int));
    for (int i = threadIdx.x; i < nz; i += blockDim.x) z[i] = 0;
    if (threadIdx.x < 6) {
        sh->vec[0][threadIdx.x / 3][threadIdx.x % 3] = kp.startV[threadIdx.x / 3][threadIdx.x % 3];
        sh->vec[1][threadIdx.x / 3][threadIdx.x % 3] = kp.endV[threadIdx.x / 3][threadIdx.x % 3];
    }
    __syncthreads();
    // wavefronts 2p and 2p+1 of a workgroup are the forward and backward rows of partner pair p
    const bool bwd = ((threadIdx.x >> 6) & 1) != 0;
    if (bwd) split_class<64, true>(kp, lds, sh, 2);
    else split_class<64, false>(kp, lds, sh, 2);
    if (__hip_atomic_load(kp.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    if (bwd) split_class<32, true>(kp, lds, sh, 1);
    else split_class<32, false>(kp, lds, sh, 1);
    if (__hip_atomic_load(kp.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    if (bwd) split_class<16, true>(kp, lds, sh, 0);
    else split_class<16, false>(kp, lds, sh, 0);
}

}  // namespace mpk
