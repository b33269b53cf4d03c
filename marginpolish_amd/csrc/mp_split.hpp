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
// the forward row emits posteriors for (cx, from] reading the backward values
// its partner left in HBM, and the backward row emits (lo, cx] reading the
// forward values; both continue their recursions.  A window's critical path
// is one sweep instead of two.
//
// Output: one record per cell with any posterior >= threshold:
// (pM, pX, pY, (d << 11) | cell).  The forward side writes each anti-diagonal's
// cells in descending order, the backward side in ascending order, so that
// the host can lay the records out in the reference's list order.
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
};
constexpr int kSplitLdsBytes =
    ((kTabFloats * 4 + (int)sizeof(SplitShared)) > kTabReach ? (kTabFloats * 4 + (int)sizeof(SplitShared))
                                                              : kTabReach);

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

template <int SEG>
__device__ void split_class(const KParams &kp, const float *tab, SplitShared *sh, int cls) {
    constexpr int K = kK;
    constexpr int kEmpty = INT32_MIN;  // band entry of a missing anti-diagonal (no cell can match)
    const int lane = threadIdx.x & 63;
    const int k = lane & (SEG - 1);
    const int row = lane / SEG;
    const int rowBase = lane & ~(SEG - 1);
    const int wib = threadIdx.x >> 6;
    const int pr = wib >> 1;
    const bool bwd = (wib & 1) != 0;  // wave-uniform role
    const size_t slot = (size_t)blockIdx.x * (blockDim.x >> 6) + wib;
    const int R = kp.ringLen;
    const int RC = R / K;
    // ring: [slot][R anti-diagonals][64 lanes] of 3 floats; exponents: [slot][R/8 cells][4 rows]
    F3 *myRing = reinterpret_cast<F3 *>(kp.ring) + slot * (size_t)R * 64 + lane;
    const F3 *peerRing = reinterpret_cast<const F3 *>(kp.ring) + (slot ^ 1) * (size_t)R * 64 + lane;
    int *myExp = kp.ringExp + slot * (size_t)RC * 4 + row;
    const int *peerExp = kp.ringExp + (slot ^ 1) * (size_t)RC * 4 + row;
    const float T00 = kp.T[0], T01 = kp.T[1], T02 = kp.T[2], T10 = kp.T[3], T11 = kp.T[4], T12 = kp.T[5],
                T20 = kp.T[6], T21 = kp.T[7], T22 = kp.T[8];
    const float thr = kp.threshold;
    const bool withIndels = kp.withIndels != 0;
    int *xflagMine = &sh->xflag[cls][pr][bwd ? 1 : 0][row];
    int *xflagPeer = &sh->xflag[cls][pr][bwd ? 0 : 1][row];
    int *jobSeqP = &sh->jobSeq[cls][pr][row];
    int *jobIdP = &sh->jobId[cls][pr][row];
    int *doneBP = &sh->doneB[cls][pr][row];

    int phase = bwd ? S_JWAIT : S_FETCH;
    int jn = 0, xn = 0;
    int job = -1, D = 0, lastCell = 0, nWin = 0, winOff = 0, jflags = 0, cap = 0, win = 0;
    int64_t symX = 0, symY = 0, bandOff = 0, outBase = 0;
    int tb = 0, lo = 0, from = 0, cx = 0;
    int d = 0, gs = 0, S = 0, Stot = 0, cnt = 0, rpos = 0, cpos = 0;
    bool post = false, init = false, reload = true;
    float totalInv = 0.f;
    float iv0 = 0.f, iv1 = 0.f, iv2 = 0.f;
    float oX1 = 0.f, oY1 = 0.f, oM1 = 0.f, oM2 = 0.f;
    int Lm1 = 0, Lm2 = 0, bm1 = kEmpty, bm2 = kEmpty;
    int bC[K], bN[K];
    uint32_t sxC[K], syC[K];
    F3 rC[K];
    int eC = 0;
#pragma unroll
    for (int j = 0; j < K; j++) {
        bC[j] = bN[j] = 0;
        sxC[j] = syC[j] = 0u;
        rC[j] = F3{0.f, 0.f, 0.f};
    }

    auto resetExchange = [&]() {
        oX1 = oY1 = oM1 = oM2 = 0.f;
        bm1 = bm2 = kEmpty;
        Lm1 = Lm2 = band_L(kEmpty);
    };
    auto loadWin = [&]() {
        const DevWin w = kp.win[winOff + win];
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
        const bool rg = tb == D && ((jflags >> 2) & 1);
        iv0 = rg ? kp.endV[1][0] : kp.endV[0][0];
        iv1 = rg ? kp.endV[1][1] : kp.endV[0][1];
        iv2 = rg ? kp.endV[1][2] : kp.endV[0][2];
    };
    auto setupJob = [&]() {
        const DevJob jb = kp.jobs[job];
        symX = jb.symX;
        symY = jb.symY;
        bandOff = jb.bandOff;
        D = jb.lX + jb.lY;
        lastCell = D / K + 1;  // bands are padded by one extra cell
        winOff = jb.winOff;
        nWin = jb.nWin;
        jflags = jb.flags;
        cap = bwd ? jb.cap[1] : jb.cap[0];
        outBase = jb.outOff + (bwd ? jb.cap[0] : 0);
        cnt = 0;
        win = 0;
        loadWin();
        if (!bwd) {
            d = 0;
            gs = 0;
            S = 0;
            init = true;
            post = false;
            reload = true;
            resetExchange();
            const bool rg = (jflags >> 1) & 1;
            iv0 = rg ? kp.startV[1][0] : kp.startV[0][0];
            iv1 = rg ? kp.startV[1][1] : kp.startV[0][1];
            iv2 = rg ? kp.startV[1][2] : kp.startV[0][2];
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
            const int *bandJ = kp.band + bandOff;
            const int b1 = bandJ[cx + 1], b2 = bandJ[cx + 2];
            const int L1 = band_L(b1), L2 = band_L(b2);
            const bool a1 = k <= band_wm1_narrow(b1), a2 = k <= band_wm1_narrow(b2) && cx + 2 <= D;
            const int Rm1 = bm1 == kEmpty ? Lm1 - 2 : Lm1 + 2 * band_wm1_narrow(bm1);
            const int Rm2 = bm2 == kEmpty ? Lm2 - 2 : Lm2 + 2 * band_wm1_narrow(bm2);
            v0 = gather(oX1, L1 + 2 * k - 1, Lm1, Rm1, a1, rowBase, k);
            v1 = gather(oY1, L1 + 2 * k + 1, Lm1, Rm1, a1, rowBase, k);
            v2 = gather(oM2, L1 + 2 * k, Lm2, Rm2, a1, rowBase, k);
            v3 = gather(oM1, L2 + 2 * k, Lm1, Rm1, a2, rowBase, k);
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
        if (k == 0) kp.winEnd[2 * (winOff + win) + 1] = cnt;
        if (win + 1 == nWin) {
            if (k == 0) {
                kp.counts[2 * job + 1] = cnt;
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
        if (phase == S_FETCH) {  // forward rows only
            int jnew = 0;
            if (k == 0) jnew = atomicAdd(&kp.queue[cls], 1);
            jnew = rbcast<SEG>(jnew, rowBase, 0);
            job = kp.jobBegin[cls] + jnew;
            if (job >= kp.jobEnd[cls]) job = -1;
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
                if (!bwd && tb == D && k == 0)
                    kp.logProb[job] = log((double)t) + (double)Stot * 0.69314718055994530942;
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
            if (++idle > kSpinLimit || __hip_atomic_load(kp.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                if (lane == 0) atomicExch(kp.error, 1);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
            continue;
        }
        idle = 0;
        if (!running) continue;
        // a macro step that cannot reach the row's stop diagonal means a broken
        // invariant: report it instead of walking off the band
        {
            const int stopNow = bwd ? (post ? lo + 1 : cx + 1) : (post ? tb : cx);
            if (bwd ? (gs + K - 1 < stopNow || gs < 0) : (gs > stopNow)) {
                atomicExch(kp.error, 2);
                phase = S_EXIT;
                continue;
            }
        }

        // ---------------- one macro step (8 anti-diagonals of one grid cell) ----------------
        const int cell = gs / K;
        const int *bandJ = kp.band + bandOff;
        const int cN = bwd ? max(cell - 1, 0) : min(cell + 1, lastCell);
        const int cN2 = bwd ? max(cell - 2, 0) : min(cell + 2, lastCell);
        if (reload) {
            rpos = gs % R;
            cpos = cell % RC;
            const int4 *b4 = reinterpret_cast<const int4 *>(bandJ + (size_t)cell * K);
            const int4 *n4 = reinterpret_cast<const int4 *>(bandJ + (size_t)cN * K);
            const int4 c0 = b4[0], c1 = b4[1], n0 = n4[0], n1 = n4[1];
            bC[0] = c0.x; bC[1] = c0.y; bC[2] = c0.z; bC[3] = c0.w;
            bC[4] = c1.x; bC[5] = c1.y; bC[6] = c1.z; bC[7] = c1.w;
            bN[0] = n0.x; bN[1] = n0.y; bN[2] = n0.z; bN[3] = n0.w;
            bN[4] = n1.x; bN[5] = n1.y; bN[6] = n1.z; bN[7] = n1.w;
#pragma unroll
            for (int j = 0; j < K; j++) {
                const int dj = gs + j, xmy = band_L(bC[j]) + 2 * k;
                sxC[j] = kp.sym[symX + ((dj + xmy) >> 1) - 1];
                syC[j] = kp.sym[symY + ((dj - xmy) >> 1) - 1];
                rC[j] = peerRing[(size_t)(rpos + j) * 64];
            }
            eC = peerExp[cpos * 4];
            reload = false;
        }
        // prefetch: band two cells ahead; symbols, partner values and partner
        // exponent of the next cell
        const int rposN = bwd ? (rpos == 0 ? R - K : rpos - K) : (rpos + K == R ? 0 : rpos + K);
        const int cposN = bwd ? (cpos == 0 ? RC - 1 : cpos - 1) : (cpos + 1 == RC ? 0 : cpos + 1);
        int bN2[K];
        uint32_t sxN[K], syN[K];
        F3 rN[K];
        int eN;
        {
            const int4 *n4 = reinterpret_cast<const int4 *>(bandJ + (size_t)cN2 * K);
            const int4 n0 = n4[0], n1 = n4[1];
            bN2[0] = n0.x; bN2[1] = n0.y; bN2[2] = n0.z; bN2[3] = n0.w;
            bN2[4] = n1.x; bN2[5] = n1.y; bN2[6] = n1.z; bN2[7] = n1.w;
#pragma unroll
            for (int j = 0; j < K; j++) {
                const int dj = cN * K + j, xmy = band_L(bN[j]) + 2 * k;
                sxN[j] = kp.sym[symX + ((dj + xmy) >> 1) - 1];
                syN[j] = kp.sym[symY + ((dj - xmy) >> 1) - 1];
                rN[j] = peerRing[(size_t)(rposN + j) * 64];
            }
            eN = peerExp[cposN * 4];
        }

        const bool slow = __any((bC[0] & (bwd ? kSlowB : kSlowF)) != 0);
        F3 *ringCell = myRing + (size_t)rpos * 64;
        bool reached = false;

        if (!bwd) {
            const int stop = post ? tb : cx;
#pragma unroll
            for (int j = 0; j < K; j++) {
                const int dj = gs + j;
                if (dj >= d && dj <= stop) {
                    const int b = bC[j];
                    const int L0 = band_L(b);
                    const bool act = k <= band_wm1_narrow(b);
                    const Emis e = emissions(tab, sxC[j], syC[j]);
                    float inX, inY, inM;
                    if (!slow) {
                        const int dL1 = L0 - Lm1, dL2 = L0 - Lm2;
                        const float xs = row_shr<SEG>(oX1, k), ys = row_shl<SEG>(oY1, k);
                        const float ml = row_shr<SEG>(oM2, k), mr = row_shl<SEG>(oM2, k);
                        inX = dL1 < 0 ? xs : oX1;
                        inY = dL1 > 0 ? ys : oY1;
                        inM = dL2 < 0 ? ml : (dL2 > 0 ? mr : oM2);
                    } else {
                        const int xmy = L0 + 2 * k;
                        const int Rm1 = bm1 == kEmpty ? Lm1 - 2 : Lm1 + 2 * band_wm1_narrow(bm1);
                        const int Rm2 = bm2 == kEmpty ? Lm2 - 2 : Lm2 + 2 * band_wm1_narrow(bm2);
                        inX = gather(oX1, xmy - 1, Lm1, Rm1, act, rowBase, k);
                        inY = gather(oY1, xmy + 1, Lm1, Rm1, act, rowBase, k);
                        inM = gather(oM2, xmy, Lm2, Rm2, act, rowBase, k);
                    }
                    float p0 = e.eM * inM, p1 = e.eX * inX, p2 = e.eY * inY;
                    if (j == 0 && init) {
                        p0 = iv0;
                        p1 = iv1;
                        p2 = iv2;
                    }
                    if (!act) p0 = p1 = p2 = 0.f;
                    if (j == 0 && dj == d) {
                        // exact power-of-two rescale at the start of every grid cell
                        const float m = rmax<SEG>(fmaxf(p0, fmaxf(p1, p2)));
                        const int e2 = m > 0.f ? __builtin_amdgcn_frexp_expf(m) : 0;
                        p0 = ldexpf(p0, -e2);
                        p1 = ldexpf(p1, -e2);
                        p2 = ldexpf(p2, -e2);
                        oM1 = ldexpf(oM1, -e2);
                        S += e2;
                        if (k == 0) myExp[cpos * 4] = S;
                    }
                    ringCell[(size_t)j * 64] = F3{p0, p1, p2};
                    if (post && dj <= from) {
                        const F3 bb = rC[j];
                        const float sc = ldexpf(totalInv, S + eC - Stot);
                        const float pM = p0 * bb.x * sc, pX = p1 * bb.y * sc, pY = p2 * bb.z * sc;
                        const bool emit = act && (pM >= thr || (withIndels && (pX >= thr || pY >= thr)));
                        const uint64_t m = __ballot(emit);
                        const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                        const int first = rbcast<SEG>(below, rowBase, 0);
                        const int last = rbcast<SEG>(below + (emit ? 1 : 0), rowBase, SEG - 1);
                        const int idx = cnt + (last - below - 1);  // descending cell order
                        if (emit && idx < cap)
                            kp.out[outBase + idx] = make_int4(__float_as_int(pM), __float_as_int(pX),
                                                              __float_as_int(pY), (int)rec_code(dj, k));
                        cnt += last - first;
                    }
                    if (dj == stop) reached = true;
                    const float q0 = T00 * p0 + T10 * p1 + T20 * p2;
                    const float q1 = T01 * p0 + T11 * p1 + T21 * p2;
                    const float q2 = T02 * p0 + T12 * p1 + T22 * p2;
                    oM2 = oM1;
                    oX1 = q1;
                    oY1 = q2;
                    oM1 = q0;
                    Lm2 = Lm1;
                    bm2 = bm1;
                    Lm1 = L0;
                    bm1 = b;
                }
            }
        } else {
            const int stop = post ? lo + 1 : cx + 1;
#pragma unroll
            for (int j = K - 1; j >= 0; j--) {
                const int dj = gs + j;
                if (dj <= d && dj >= stop) {
                    const int b = bC[j];
                    const int L0 = band_L(b);
                    const bool act = k <= band_wm1_narrow(b);
                    const Emis e = emissions(tab, sxC[j], syC[j]);
                    float inX, inY, inM;
                    // successors: X at (d+1, xmy+1), Y at (d+1, xmy-1), M at (d+2, xmy)
                    if (!slow) {
                        const int dL1 = L0 - Lm1, dL2 = L0 - Lm2;
                        const float xs = row_shl<SEG>(oX1, k), ys = row_shr<SEG>(oY1, k);
                        const float ml = row_shr<SEG>(oM2, k), mr = row_shl<SEG>(oM2, k);
                        inX = dL1 > 0 ? xs : oX1;
                        inY = dL1 < 0 ? ys : oY1;
                        inM = dL2 < 0 ? ml : (dL2 > 0 ? mr : oM2);
                    } else {
                        const int xmy = L0 + 2 * k;
                        const int Rm1 = bm1 == kEmpty ? Lm1 - 2 : Lm1 + 2 * band_wm1_narrow(bm1);
                        const int Rm2 = bm2 == kEmpty ? Lm2 - 2 : Lm2 + 2 * band_wm1_narrow(bm2);
                        inX = gather(oX1, xmy + 1, Lm1, Rm1, act, rowBase, k);
                        inY = gather(oY1, xmy - 1, Lm1, Rm1, act, rowBase, k);
                        inM = gather(oM2, xmy, Lm2, Rm2, act, rowBase, k);
                    }
                    // B = T in
                    float q0 = T00 * inM + T01 * inX + T02 * inY;
                    float q1 = T10 * inM + T11 * inX + T12 * inY;
                    float q2 = T20 * inM + T21 * inX + T22 * inY;
                    if (init && dj == d) {
                        q0 = iv0;
                        q1 = iv1;
                        q2 = iv2;
                    }
                    if (!act) q0 = q1 = q2 = 0.f;
                    if (j == K - 1 && dj == d && !init) {
                        const float m = rmax<SEG>(fmaxf(q0, fmaxf(q1, q2)));
                        const int e2 = m > 0.f ? __builtin_amdgcn_frexp_expf(m) : 0;
                        q0 = ldexpf(q0, -e2);
                        q1 = ldexpf(q1, -e2);
                        q2 = ldexpf(q2, -e2);
                        oM1 = ldexpf(oM1, -e2);
                        S += e2;
                    }
                    if (!post) {
                        if (dj == d && k == 0) myExp[cpos * 4] = S;
                        ringCell[(size_t)j * 64] = F3{q0, q1, q2};
                    } else if (dj <= from) {
                        const F3 f = rC[j];
                        const float sc = ldexpf(totalInv, eC + S - Stot);
                        const float pM = f.x * q0 * sc, pX = f.y * q1 * sc, pY = f.z * q2 * sc;
                        const bool emit = act && (pM >= thr || (withIndels && (pX >= thr || pY >= thr)));
                        const uint64_t m = __ballot(emit);
                        const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                        const int first = rbcast<SEG>(below, rowBase, 0);
                        const int last = rbcast<SEG>(below + (emit ? 1 : 0), rowBase, SEG - 1);
                        const int idx = cnt + (below - first);  // ascending cell order
                        if (emit && idx < cap)
                            kp.out[outBase + idx] = make_int4(__float_as_int(pM), __float_as_int(pX),
                                                              __float_as_int(pY), (int)rec_code(dj, k));
                        cnt += last - first;
                    }
                    if (dj == stop) reached = true;
                    // send u = e (.) B
                    oM2 = oM1;
                    oX1 = e.eX * q1;
                    oY1 = e.eY * q2;
                    oM1 = e.eM * q0;
                    Lm2 = Lm1;
                    bm2 = bm1;
                    Lm1 = L0;
                    bm1 = b;
                }
            }
        }
        init = false;
#pragma unroll
        for (int j = 0; j < K; j++) {
            bC[j] = bN[j];
            bN[j] = bN2[j];
            sxC[j] = sxN[j];
            syC[j] = syN[j];
            rC[j] = rN[j];
        }
        eC = eN;

        if (reached) {
            if (!post) {
                publishCrossing();
            } else if (!bwd) {
                // forward window end at tb
                if (k == 0) kp.winEnd[2 * (winOff + win)] = cnt;
                if (win + 1 == nWin) {
                    if (k == 0) kp.counts[2 * job] = cnt;
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
        } else {
            gs += bwd ? -K : K;
            d = bwd ? gs + K - 1 : gs;
            rpos = rposN;
            cpos = cposN;
        }
    }
}

__global__ __launch_bounds__(512) void pairhmm_split_kernel(KParams kp) {
    extern __shared__ float lds[];
    for (int i = threadIdx.x; i < kTabFloats; i += blockDim.x) lds[i] = kp.tables[i];
    SplitShared *sh = reinterpret_cast<SplitShared *>(lds + kTabFloats);
    int *z = reinterpret_cast<int *>(sh->xflag);
    const int nz = (int)((sizeof(SplitShared) - offsetof(SplitShared, xflag)) / sizeof(int));
    for (int i = threadIdx.x; i < nz; i += blockDim.x) z[i] = 0;
    __syncthreads();
    split_class<64>(kp, lds, sh, 2);
    if (__hip_atomic_load(kp.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    split_class<32>(kp, lds, sh, 1);
    if (__hip_atomic_load(kp.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    split_class<16>(kp, lds, sh, 0);
}

}  // namespace mpk
