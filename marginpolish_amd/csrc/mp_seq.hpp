// mp_seq.hpp -- forward-only likelihood and expected transition counts
// (gfx950): one job per row of SEG lanes, forward sweep then (expectation
// mode) backward sweep in the same row.
//
// Reference: computeForwardProbability (pairwiseAligner.c:813-865) and
// getExpectationsUsingAnchors with diagonalCalculationExpectations
// (pairwiseAligner.c:347-369, 666-677, 1151-1164).
#pragma once

#include "mp_split.hpp"

namespace mpk {

constexpr int kSeqLdsBytes = kTabReach;

enum { Q_FETCH = 0, Q_FWD = 1, Q_BWD = 2, Q_DONE = 3 };

// Anti-diagonals are walked in grid-aligned macro steps of 8 (ascending
// forward, descending backward), every per-step operand a statically indexed
// register loaded one macro step ahead.  Forward values of the current
// traceback window are kept in an HBM ring (float4: 3 states + the cell's
// exponent) for the backward sweep.
template <int SEG>
__device__ __forceinline__ void seq_class(const KParams &kp, const float *tab, float *hist, int cls) {
    constexpr int K = kK;
    constexpr int kEmpty = kEmptyBand;
    const int lane = threadIdx.x & 63;
    const int k = lane & (SEG - 1);
    const int rowBase = lane & ~(SEG - 1);
    const int rowInWave = lane / SEG;
    const size_t waveSlot = (size_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int R = kp.ringLen;
    float4 *ringRow = reinterpret_cast<float4 *>(kp.ring) + waveSlot * (size_t)R * 64 + (size_t)rowInWave * R * SEG + k;
    const bool expMode = kp.mode == MP_MODE_EXPECTATIONS;
    const float T00 = kp.T[0], T01 = kp.T[1], T02 = kp.T[2], T10 = kp.T[3], T11 = kp.T[4], T12 = kp.T[5],
                T20 = kp.T[6], T21 = kp.T[7], T22 = kp.T[8];

    int phase = Q_FETCH;
    int job = 0, d = 0, gs = 0, win = 0, nWin = 0, winOff = 0, flags = 0;
    int tb = 0, lo = 0, from = 0, D = 0, lastCell = 0;
    int64_t symX = 0, symY = 0, bandOff = 0;
    int S = 0, SFtb = 0;
    float totalInv = 0.f;
    double likeAcc = 0.0;
    bool init = false, reload = true;
    float iv0 = 0.f, iv1 = 0.f, iv2 = 0.f;
    float oX1 = 0.f, oY1 = 0.f, oM1 = 0.f, oM2 = 0.f;
    int bm1 = kEmpty, bm2 = kEmpty;
    float fX1 = 0.f, fY1 = 0.f, fM1 = 0.f, fM2 = 0.f;
    int fbm1 = kEmpty, fbm2 = kEmpty, fS = 0, fTb = 0;
    int bC[K], bN[K];
    uint32_t sxC[K], syC[K];
    float4 rC[K];
#pragma unroll
    for (int j = 0; j < K; j++) {
        bC[j] = bN[j] = 0;
        sxC[j] = syC[j] = 0u;
        rC[j] = make_float4(0.f, 0.f, 0.f, 0.f);
    }

    for (;;) {
        if (phase == Q_FETCH) {
            int jn = 0;
            if (k == 0) jn = atomicAdd(&kp.queue[cls], 1);
            jn = rbcast<SEG>(jn, rowBase, 0);
            job = kp.jobBegin[cls] + jn;
            if (job >= kp.jobEnd[cls]) {
                phase = Q_DONE;
            } else {
                const DevJob jb = kp.jobs[job];
                symX = jb.symX;
                symY = jb.symY;
                bandOff = jb.bandOff;
                D = jb.lX + jb.lY;
                lastCell = D / K + 1;
                winOff = jb.winOff;
                nWin = jb.nWin;
                flags = jb.flags;
                const DevWin w0 = kp.win[winOff];
                win = 0;
                tb = w0.tb;
                lo = w0.lo;
                from = w0.from;
                likeAcc = 0.0;
                d = 0;
                gs = 0;
                S = 0;
                init = true;
                reload = true;
                oX1 = oY1 = oM1 = oM2 = 0.f;
                bm1 = bm2 = kEmpty;
                const bool rg = (flags >> 1) & 1;
                iv0 = rg ? kp.startV[1][0] : kp.startV[0][0];
                iv1 = rg ? kp.startV[1][1] : kp.startV[0][1];
                iv2 = rg ? kp.startV[1][2] : kp.startV[0][2];
                phase = Q_FWD;
            }
        }
        if (__all(phase == Q_DONE)) break;
        if (phase != Q_FWD && phase != Q_BWD) continue;
        const bool fwd = phase == Q_FWD;
        const int cell = gs / K;
        const int *bandJ = kp.band + bandOff;
        const int cN = min(max(cell + (fwd ? 1 : -1), 0), lastCell);
        const int cN2 = min(max(cell + (fwd ? 2 : -2), 0), lastCell);

        if (reload) {
            const int4 *b4 = reinterpret_cast<const int4 *>(bandJ + (size_t)cell * K);
            const int4 *n4 = reinterpret_cast<const int4 *>(bandJ + (size_t)cN * K);
            const int4 c0 = b4[0], c1 = b4[1], n0 = n4[0], n1 = n4[1];
            bC[0] = c0.x; bC[1] = c0.y; bC[2] = c0.z; bC[3] = c0.w;
            bC[4] = c1.x; bC[5] = c1.y; bC[6] = c1.z; bC[7] = c1.w;
            bN[0] = n0.x; bN[1] = n0.y; bN[2] = n0.z; bN[3] = n0.w;
            bN[4] = n1.x; bN[5] = n1.y; bN[6] = n1.z; bN[7] = n1.w;
            const float4 *rb = ringRow + (size_t)((gs % R) * SEG);
#pragma unroll
            for (int j = 0; j < K; j++) {
                const int dj = gs + j, x = band_xL(bC[j]) + k;
                sxC[j] = kp.sym[symX + x - 1];
                syC[j] = kp.sym[symY + (dj - x) - 1];
                rC[j] = rb[(size_t)j * SEG];
            }
            reload = false;
        }
        int bN2[K];
        uint32_t sxN[K], syN[K];
        float4 rN[K];
        {
            const int4 *n4 = reinterpret_cast<const int4 *>(bandJ + (size_t)cN2 * K);
            const int4 n0 = n4[0], n1 = n4[1];
            bN2[0] = n0.x; bN2[1] = n0.y; bN2[2] = n0.z; bN2[3] = n0.w;
            bN2[4] = n1.x; bN2[5] = n1.y; bN2[6] = n1.z; bN2[7] = n1.w;
            const float4 *rb = ringRow + (size_t)(((cN * K) % R) * SEG);
#pragma unroll
            for (int j = 0; j < K; j++) {
                const int dj = cN * K + j, x = band_xL(bN[j]) + k;
                sxN[j] = kp.sym[symX + x - 1];
                syN[j] = kp.sym[symY + (dj - x) - 1];
                rN[j] = rb[(size_t)j * SEG];
            }
        }

        const bool atEnd = tb == D;
        const bool ragEnd = atEnd && ((flags >> 2) & 1);
        const bool slow = __any((bC[0] & (fwd ? kSlowF : kSlowB)) != 0);
        bool switchBwd = false, winDone = false;
        float sv0 = 0.f, sv1 = 0.f, sv2 = 0.f;
        float e00 = 0.f, e01 = 0.f, e02 = 0.f, e10 = 0.f, e11 = 0.f, e12 = 0.f, e20 = 0.f, e21 = 0.f, e22 = 0.f;
        float4 *ringCell = ringRow + (size_t)((gs % R) * SEG);

        if (fwd) {
#pragma unroll
            for (int j = 0; j < K; j++) {
                const int dj = gs + j;
                if (dj >= d && dj <= tb) {
                    const int b = bC[j];
                    const int xL = band_xL(b);
                    const bool act = k <= band_wm1_narrow(b);
                    const Emis e = emissions(tab, sxC[j], syC[j]);
                    float inX, inY, inM;
                    if (!slow) fwd_in_fast<SEG>(xL, band_xL(bm1), band_xL(bm2), oX1, oY1, oM2, k, inX, inY, inM);
                    else fwd_in_slow(xL, bm1, bm2, oX1, oY1, oM2, act, rowBase, k, inX, inY, inM);
                    float p0 = e.eM * inM, p1 = e.eX * inX, p2 = e.eY * inY;
                    if (j == 0 && init) {
                        p0 = iv0;
                        p1 = iv1;
                        p2 = iv2;
                    }
                    if (!act) p0 = p1 = p2 = 0.f;
                    if (j == 0 && dj == d) {
                        const float m = rmaxpos<SEG>(fmaxf(p0, fmaxf(p1, p2)));
                        const int e2 = __builtin_amdgcn_frexp_expf(m);
                        p0 = ldexpf(p0, -e2);
                        p1 = ldexpf(p1, -e2);
                        p2 = ldexpf(p2, -e2);
                        oM1 = ldexpf(oM1, -e2);
                        S += e2;
                    }
                    const float q0 = fmaf(T20, p2, fmaf(T10, p1, T00 * p0));
                    const float q1 = fmaf(T21, p2, fmaf(T11, p1, T01 * p0));
                    const float q2 = fmaf(T22, p2, fmaf(T12, p1, T02 * p0));
                    if (expMode) ringCell[(size_t)j * SEG] = make_float4(p0, p1, p2, __int_as_float(S));
                    if (dj == tb) {
                        sv0 = p0;
                        sv1 = p1;
                        sv2 = p2;
                        switchBwd = true;
                    }
                    oM2 = oM1;
                    oX1 = q1;
                    oY1 = q2;
                    oM1 = q0;
                    bm2 = bm1;
                    bm1 = b;
                }
            }
        } else {
#pragma unroll
            for (int j = K - 1; j >= 0; j--) {
                const int dj = gs + j;
                if (dj <= d && dj >= lo) {
                    const int b = bC[j];
                    const int xL = band_xL(b);
                    const bool act = k <= band_wm1_narrow(b);
                    const int Sin = S;
                    const Emis e = emissions(tab, sxC[j], syC[j]);
                    float inX, inY, inM;
                    if (!slow) bwd_in_fast<SEG>(xL, band_xL(bm1), band_xL(bm2), oX1, oY1, oM2, k, inX, inY, inM);
                    else bwd_in_slow(xL, bm1, bm2, oX1, oY1, oM2, act, rowBase, k, inX, inY, inM);
                    float q0 = fmaf(T02, inY, fmaf(T01, inX, T00 * inM));
                    float q1 = fmaf(T12, inY, fmaf(T11, inX, T10 * inM));
                    float q2 = fmaf(T22, inY, fmaf(T21, inX, T20 * inM));
                    if (init && dj == d) {
                        q0 = iv0;
                        q1 = iv1;
                        q2 = iv2;
                    }
                    if (!act) q0 = q1 = q2 = 0.f;
                    if (j == K - 1 && dj == d && !init) {
                        const float m = rmaxpos<SEG>(fmaxf(q0, fmaxf(q1, q2)));
                        const int e2 = __builtin_amdgcn_frexp_expf(m);
                        q0 = ldexpf(q0, -e2);
                        q1 = ldexpf(q1, -e2);
                        q2 = ldexpf(q2, -e2);
                        oM1 = ldexpf(oM1, -e2);
                        S += e2;
                    }
                    if (act) {
                        // expected transition counts, pushed from this cell (the predecessor)
                        // into successors on posterior diagonals (diagonalCalculationExpectations)
                        const float4 f = rC[j];
                        const float sc = ldexpf(totalInv, __float_as_int(f.w) + Sin - SFtb);
                        const bool okXY = dj + 1 > lo && dj + 1 <= from, okM = dj + 2 > lo && dj + 2 <= from;
                        const float tX = okXY ? inX : 0.f, tY = okXY ? inY : 0.f, tM = okM ? inM : 0.f;
                        const float a0 = f.x * sc, a1 = f.y * sc, a2 = f.z * sc;
                        e00 += a0 * (T00 * tM); e01 += a0 * (T01 * tX); e02 += a0 * (T02 * tY);
                        e10 += a1 * (T10 * tM); e11 += a1 * (T11 * tX); e12 += a1 * (T12 * tY);
                        e20 += a2 * (T20 * tM); e21 += a2 * (T21 * tX); e22 += a2 * (T22 * tY);
                        // expected emission counts: this cell's state posteriors, binned by
                        // the emission-table entry each state reads here
                        if (dj > lo && dj <= from) {
                            const float scq = ldexpf(totalInv, __float_as_int(f.w) + S - SFtb);
                            const uint32_t sx = sxC[j], sy = syC[j], sxy = sx + sy;
                            const float pM = f.x * q0 * scq;
                            atomicAdd(&hist[kTabE + (((sxy >> 16) & 0xffu) >> 2)], pM);
                            atomicAdd(&hist[kTabRL + (min(sxy & 0xffffu, kRLLastByte) >> 2)], pM);
                            atomicAdd(&hist[(sx >> 24) >> 2], f.y * q1 * scq);
                            atomicAdd(&hist[(sy >> 24) >> 2], f.z * q2 * scq);
                        }
                    }
                    if (dj == lo) winDone = true;
                    oM2 = oM1;
                    oX1 = e.eX * q1;
                    oY1 = e.eY * q2;
                    oM1 = e.eM * q0;
                    bm2 = bm1;
                    bm1 = b;
                }
            }
            const float ev[9] = {e00, e01, e02, e10, e11, e12, e20, e21, e22};
            for (int i = 0; i < 9; i++) {
                const float v = rsum<SEG>(ev[i]);
                if (k == 0 && v != 0.f) atomicAdd(&kp.expect[(size_t)job * 10 + i], (double)v);
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

        if (switchBwd) {
            // window total sum(F(tb) . end)
            const float e0 = ragEnd ? kp.endV[1][0] : kp.endV[0][0];
            const float e1 = ragEnd ? kp.endV[1][1] : kp.endV[0][1];
            const float e2 = ragEnd ? kp.endV[1][2] : kp.endV[0][2];
            const float t = rsum<SEG>(sv0 * e0 + sv1 * e1 + sv2 * e2);
            SFtb = S;
            totalInv = t > 0.f ? 1.f / t : 0.f;
            const double lt = log((double)t) + (double)SFtb * 0.69314718055994530942;
            if (atEnd && k == 0) kp.logTotal[job] = make_int2(__float_as_int(t), SFtb);
            // getExpectations adds the window total once per posterior diagonal (pairwiseAligner.c:674)
            likeAcc += (double)(from - lo) * lt;
            if (!expMode) {
                if (k == 0) kp.counts[2 * job] = kp.counts[2 * job + 1] = 0;
                phase = Q_FETCH;
            } else {
                iv0 = e0;
                iv1 = e1;
                iv2 = e2;
                fX1 = oX1; fY1 = oY1; fM1 = oM1; fM2 = oM2;
                fbm1 = bm1; fbm2 = bm2; fS = S; fTb = tb;
                oX1 = oY1 = oM1 = oM2 = 0.f;
                bm1 = bm2 = kEmpty;
                S = 0;
                init = true;
                reload = true;
                d = tb;
                gs = tb & ~(K - 1);
                phase = Q_BWD;
            }
        } else if (winDone) {
            if (win + 1 == nWin) {
                if (k == 0) {
                    kp.counts[2 * job] = kp.counts[2 * job + 1] = 0;
                    kp.expect[(size_t)job * 10 + 9] = likeAcc;
                }
                phase = Q_FETCH;
            } else {
                win += 1;
                const DevWin w1 = kp.win[winOff + win];
                d = fTb + 1;
                gs = d & ~(K - 1);
                tb = w1.tb;
                lo = w1.lo;
                from = w1.from;
                oX1 = fX1; oY1 = fY1; oM1 = fM1; oM2 = fM2;
                bm1 = fbm1; bm2 = fbm2; S = fS;
                reload = true;
                phase = Q_FWD;
            }
        } else {
            gs += fwd ? K : -K;
            d = fwd ? gs : gs + K - 1;
        }
    }
}

// Expectation mode adds an FP32 histogram of expected emission counts over
// the emission-table entries (one per workgroup, in LDS after the table),
// added to kp.emitCount in FP64 when the workgroup ends.
constexpr int kSeqLdsBytesExp = kTabReach + kTabFloats * 4;

__global__ __launch_bounds__(512) void pairhmm_seq_kernel(KParams kp) {
    extern __shared__ float lds[];
    float *hist = lds + kTabReach / 4;
    const bool expMode = kp.mode == MP_MODE_EXPECTATIONS;
    for (int i = threadIdx.x; i < kTabFloats; i += blockDim.x) {
        lds[i] = kp.tables[i];
        if (expMode) hist[i] = 0.f;
    }
    __syncthreads();
    seq_class<64>(kp, lds, hist, 2);
    seq_class<32>(kp, lds, hist, 1);
    seq_class<16>(kp, lds, hist, 0);
    if (expMode) {
        __syncthreads();
        for (int i = threadIdx.x; i < kTabFloats; i += blockDim.x)
            if (hist[i] != 0.f) atomicAdd(&kp.emitCount[i], (double)hist[i]);
    }
}

}  // namespace mpk
