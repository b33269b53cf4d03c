// mp_pairhmm.hip -- banded pair-HMM forward-backward + posterior extraction
// for gfx950 (MI355X), and the C-ABI batch driver around it.
//
// Reference algorithm: getPosteriorProbsWithBanding (pairwiseAligner.c:687-808)
// with stateMachine3_cellCalculate (stateMachine.c:559-581), RLE emissions
// (stateMachine.c:702-736) and posterior emission (pairwiseAligner.c:586-664).
//
// Kernel design
//  * One band anti-diagonal per step; one lane per band cell.  A job (one
//    read/reference sub-problem) owns a "row" of seg lanes (16, 32 or 64, the
//    smallest power of two >= the job's widest anti-diagonal), so a 64-lane
//    wavefront runs 64/seg independent jobs.
//  * Anti-diagonals are walked in grid-aligned macro steps of 8 (one "cell"
//    of the grid d = 8g..8g+7), forward ascending and backward descending.
//    The 8 steps are unrolled and every per-step operand (packed band entry,
//    symbols, forward values read back by the backward sweep) is a statically
//    indexed register loaded one macro step ahead (the band two), so no wait
//    inside a macro step covers a memory operation of that step.
//  * Probability space: forward f = e (.) in, sends q = T^t f; backward
//    B = T in, sends u = e (.) B.  The three neighbour inputs (gap-X from
//    xmy-1 / xmy+1 on the previous step, gap-Y from xmy+1 / xmy-1, match from
//    two steps back) are one-lane DPP wave shifts when the band edge moved by
//    one (the common case), ds_bpermute with a range check otherwise.
//  * FP32 scaled arithmetic: at the start of every grid cell the row maximum
//    is folded to [0.5,1) by an exact power-of-two rescale (integer exponent,
//    exact bookkeeping).  Per traceback window the total is sum(F(tb) . end)
//    (the reference's cut total is constant across a window in exact
//    arithmetic); log-likelihoods are formed in FP64.
//  * Forward values of the current traceback window live in an HBM ring
//    (float4 per lane per anti-diagonal: 3 states + the cell's exponent) and
//    are read back by the backward sweep; posteriors >= threshold are
//    compacted per row with ballot/popcount prefixes and written as
//    (weight, x-1, y-1) tuples in the reference's per-window emission order.
//  * Persistent launch: one workgroup per CU, every row pulls jobs from a
//    per-class atomic queue ordered by band area (largest first).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "mp_internal.hpp"

namespace {

// ---------------------------------------------------------------------------
// Device-side layout
// ---------------------------------------------------------------------------

// Per-strand emission table, copied to LDS: E[5][5] match (N rows/cols =
// 1/16), GX[5], GY[5] (N = 1/4), then the run-length factor
// RL[5][51][51] = exp(2.3025 * log10 P(observed | underlying)) by x base.
constexpr int kOffE = 0, kOffGX = 25, kOffGY = 30, kOffRL = 40;
constexpr int kTabStride = 13056;  // >= 40 + 5*51*51, multiple of 64
constexpr int kRingStride = 256;   // floats per ring diagonal per wave: float4 x 64 lanes
constexpr int kNumClasses = 3;     // SEG 16, 32, 64
constexpr int kSymPad = 128;       // N symbols around the symbol pool
constexpr int kWideMaxC = 32;      // wide kernel: up to 64 * 32 = 2048 cells per anti-diagonal
constexpr int kBlock = 512;
constexpr int kK = 8;              // anti-diagonals per macro step (grid cell)

struct DevJob {
    int64_t symX, symY, bandOff, outOff;
    int32_t lX, lY, winOff, nWin;
    int32_t cap[3];
    int32_t flags;  // bit0 strand, bit1 ragged left, bit2 ragged right
};

struct DevWin {
    int32_t tb, lo, from, pad;
};

struct KParams {
    const DevJob *jobs;
    const uint16_t *sym;
    const int *band;  // packed (xmyL << 12) | (width - 1), cells of kK padded
    const DevWin *win;
    int4 *out;
    int32_t *counts;
    double *logProb;
    float *ring;
    const float *tables;
    int32_t *queue;  // kNumClasses + 1 counters (the last for wide jobs)
    float *ringWide;
    int32_t ringWideLen, wideC, wideBegin, wideEnd;
    int32_t jobBegin[kNumClasses], jobEnd[kNumClasses];
    int32_t ringLen;
    int32_t withIndels;
    int32_t mode;        // MP_MODE_*
    double *expect;      // per job: 9 expected transition counts + likelihood (expectation mode)
    int32_t rle;
    float threshold;
    float T[9];          // probability transitions, [from*3 + to], M=0 X=1 Y=2
    float startV[2][3];  // [ragged][state]
    float endV[2][3];
};

__device__ __forceinline__ float bperm(float v, int srcLane) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(srcLane << 2, __float_as_int(v)));
}

enum { PH_FETCH = 0, PH_FWD = 1, PH_BWD = 2, PH_DONE = 3 };

__device__ __forceinline__ float shr1(float v) {  // lane i <- lane i-1 (0 into lane 0)
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float shl1(float v) {  // lane i <- lane i+1 (0 into lane 63)
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, true));
}

__device__ __forceinline__ float row_max(float v, int seg) {
    for (int o = seg >> 1; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ float row_sum(float v, int seg) {
    for (int o = seg >> 1; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Band entries are packed as (xmyL << 12) | (width - 1).
__device__ __forceinline__ int band_L(int b) { return b >> 12; }
__device__ __forceinline__ int band_R(int b) { return (b >> 12) + 2 * (b & 4095); }

// Neighbour input of a cell from the exchange register v of a previous step:
// the source lane is k + s.  Fast path (|s| <= 1): one DPP shift; the lanes
// that would read across a row boundary read 0.  General path: ds_bpermute
// with an explicit band-range check.
struct Xchg {
    int k, seg, rowBase;
    __device__ __forceinline__ float fast(float v, int s) const {
        const float l = shr1(v), r = shl1(v);
        const float a = (k == 0) ? 0.f : l;
        const float b = (k == seg - 1) ? 0.f : r;
        return s == 0 ? v : (s < 0 ? a : b);
    }
    __device__ __forceinline__ float slow(float v, int srcXmy, int Lp, int Rp, bool act) const {
        const bool ok = act && srcXmy >= Lp && srcXmy <= Rp;
        const float r = bperm(v, rowBase + (ok ? ((srcXmy - Lp) >> 1) : k));
        return ok ? r : 0.f;
    }
};

// One job per row of seg lanes (16, 32 or 64); the anti-diagonals are walked
// in grid-aligned macro steps of kK (8).  A forward macro step visits the 8
// diagonals of one grid cell in ascending order, a backward one in descending
// order; both bodies are unrolled so every per-step operand (band, symbols,
// ring entries) is a statically indexed register loaded one macro step ahead
// (the band two).
__device__ void run_class(const KParams &kp, const float *lds, int cls, int seg) {
    constexpr int K = kK;
    const int lane = threadIdx.x & 63;
    const int k = lane & (seg - 1);
    const int rowBase = lane & ~(seg - 1);
    const int rowInWave = lane / seg;
    const size_t waveSlot = (size_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int R = kp.ringLen;  // multiple of K
    float4 *ringRow = reinterpret_cast<float4 *>(kp.ring) + waveSlot * (size_t)R * 64 + (size_t)rowInWave * R * seg + k;
    const uint64_t rowMask = (seg == 64 ? ~0ull : ((1ull << seg) - 1ull)) << rowBase;
    const uint64_t belowInRow = rowMask & ((1ull << lane) - 1ull);
    const Xchg xc{k, seg, rowBase};

    int phase = PH_FETCH;
    int job = 0, d = 0, gs = 0, win = 0, nWin = 0;
    int tb = 0, lo = 0, from = 0, D = 0, lastCell = 0;
    int64_t symX = 0, symY = 0, bandOff = 0, outOff = 0;
    int capM = 0, capX = 0, capY = 0, flags = 0, winOff = 0;
    int cntM = 0, cntX = 0, cntY = 0;
    int S = 0, SFtb = 0;
    float totalInv = 0.f;
    double likeAcc = 0.0;
    bool init = false, reload = true;
    float iv0 = 0.f, iv1 = 0.f, iv2 = 0.f;
    int Lm1 = 1, Rm1 = -1, Lm2 = 1, Rm2 = -1;
    float oX1 = 0.f, oY1 = 0.f, oM1 = 0.f, oM2 = 0.f;
    int fLm1 = 1, fRm1 = -1, fLm2 = 1, fRm2 = -1, fS = 0, fTb = 0;
    const bool expMode = kp.mode == MP_MODE_EXPECTATIONS, fwdOnly = kp.mode == MP_MODE_FORWARD_ONLY;
    const bool postMode = kp.mode == MP_MODE_POSTERIORS || kp.mode == MP_MODE_MATCH_ONLY;
    const int stopOff = expMode ? 0 : 1;  // backward sweep ends at lo + stopOff
    float fX1 = 0.f, fY1 = 0.f, fM1 = 0.f, fM2 = 0.f;
    const float T00 = kp.T[0], T01 = kp.T[1], T02 = kp.T[2], T10 = kp.T[3], T11 = kp.T[4], T12 = kp.T[5],
                T20 = kp.T[6], T21 = kp.T[7], T22 = kp.T[8];
    int bC[K], bN[K];
    unsigned sxC[K], syC[K];
    float4 rC[K];
#pragma unroll
    for (int j = 0; j < K; j++) {
        bC[j] = bN[j] = 0;
        sxC[j] = syC[j] = 4u;
        rC[j] = make_float4(0.f, 0.f, 0.f, 0.f);
    }

    for (;;) {
        if (phase == PH_FETCH) {
            int jn = 0;
            if (k == 0) jn = atomicAdd(&kp.queue[cls], 1);
            jn = __shfl(jn, rowBase, 64);
            job = kp.jobBegin[cls] + jn;
            if (job >= kp.jobEnd[cls]) {
                phase = PH_DONE;
            } else {
                const DevJob jb = kp.jobs[job];
                symX = jb.symX; symY = jb.symY; bandOff = jb.bandOff; outOff = jb.outOff;
                D = jb.lX + jb.lY;
                lastCell = D / K;
                winOff = jb.winOff; nWin = jb.nWin;
                capM = jb.cap[0]; capX = jb.cap[1]; capY = jb.cap[2];
                flags = jb.flags;
                const DevWin w0 = kp.win[winOff];
                win = 0; tb = w0.tb; lo = w0.lo; from = w0.from;
                cntM = cntX = cntY = 0;
                likeAcc = 0.0;
                d = 0; gs = 0; S = 0; init = true; reload = true;
                Lm1 = 1; Rm1 = -1; Lm2 = 1; Rm2 = -1;
                oX1 = oY1 = oM1 = oM2 = 0.f;
                const bool rg = (flags >> 1) & 1;
                iv0 = rg ? kp.startV[1][0] : kp.startV[0][0];
                iv1 = rg ? kp.startV[1][1] : kp.startV[0][1];
                iv2 = rg ? kp.startV[1][2] : kp.startV[0][2];
                phase = PH_FWD;
            }
        }
        if (__all(phase == PH_DONE)) break;
        if (phase != PH_FWD && phase != PH_BWD) continue;
        const bool fwd = phase == PH_FWD;
        const int step = fwd ? K : -K;
        const int cell = gs / K;
        const int *bandJ = kp.band + bandOff;

        if (reload) {
            const int cN = min(max(cell + (fwd ? 1 : -1), 0), lastCell);
            const int4 *b4 = reinterpret_cast<const int4 *>(bandJ + (size_t)cell * K);
            const int4 *n4 = reinterpret_cast<const int4 *>(bandJ + (size_t)cN * K);
            const int4 c0 = b4[0], c1 = b4[1], n0 = n4[0], n1 = n4[1];
            bC[0] = c0.x; bC[1] = c0.y; bC[2] = c0.z; bC[3] = c0.w; bC[4] = c1.x; bC[5] = c1.y; bC[6] = c1.z; bC[7] = c1.w;
            bN[0] = n0.x; bN[1] = n0.y; bN[2] = n0.z; bN[3] = n0.w; bN[4] = n1.x; bN[5] = n1.y; bN[6] = n1.z; bN[7] = n1.w;
#pragma unroll
            for (int j = 0; j < K; j++) {
                const int dj = gs + j, xmy = band_L(bC[j]) + 2 * k;
                sxC[j] = kp.sym[symX + ((dj + xmy) >> 1) - 1];
                syC[j] = kp.sym[symY + ((dj - xmy) >> 1) - 1];
            }
            if (!fwd) {
                const float4 *rb = ringRow + (size_t)((gs % R) * seg);
#pragma unroll
                for (int j = 0; j < K; j++) rC[j] = rb[(size_t)j * seg];
            }
            reload = false;
        }

        // Prefetch: band of the cell two macro steps ahead, symbols and (backward)
        // ring entries of the next cell.
        int bN2[K];
        unsigned sxN[K], syN[K];
        float4 rN[K];
        {
            const int c2 = min(max(cell + 2 * (fwd ? 1 : -1), 0), lastCell);
            const int4 *n4 = reinterpret_cast<const int4 *>(bandJ + (size_t)c2 * K);
            const int4 n0 = n4[0], n1 = n4[1];
            bN2[0] = n0.x; bN2[1] = n0.y; bN2[2] = n0.z; bN2[3] = n0.w;
            bN2[4] = n1.x; bN2[5] = n1.y; bN2[6] = n1.z; bN2[7] = n1.w;
            const int c1 = min(max(cell + (fwd ? 1 : -1), 0), lastCell);
#pragma unroll
            for (int j = 0; j < K; j++) {
                const int dj = c1 * K + j, xmy = band_L(bN[j]) + 2 * k;
                sxN[j] = kp.sym[symX + ((dj + xmy) >> 1) - 1];
                syN[j] = kp.sym[symY + ((dj - xmy) >> 1) - 1];
            }
            if (!fwd) {
                const float4 *rb = ringRow + (size_t)(((c1 * K) % R) * seg);
#pragma unroll
                for (int j = 0; j < K; j++) rN[j] = rb[(size_t)j * seg];
            } else {
#pragma unroll
                for (int j = 0; j < K; j++) rN[j] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
        }

        const float *tab = lds + ((flags & 1) ? kTabStride : 0);
        const bool atEnd = tb == D;
        const bool ragEnd = atEnd && ((flags >> 2) & 1);
        bool switchBwd = false, winDone = false;
        float sv0 = 0.f, sv1 = 0.f, sv2 = 0.f;
        float e00 = 0.f, e01 = 0.f, e02 = 0.f, e10 = 0.f, e11 = 0.f, e12 = 0.f, e20 = 0.f, e21 = 0.f, e22 = 0.f;
        float4 *ringCell = ringRow + (size_t)((gs % R) * seg);

        if (fwd) {
#pragma unroll
            for (int j = 0; j < K; j++) {
                const int dj = gs + j;
                if (dj >= d && dj <= tb) {
                    const int L0 = band_L(bC[j]), R0 = band_R(bC[j]);
                    const int xmy = L0 + 2 * k;
                    const bool act = xmy <= R0;
                    const unsigned sx = sxC[j], sy = syC[j];
                    const unsigned xb = min(sx & 255u, 4u), yb = min(sy & 255u, 4u);
                    const float eX = tab[kOffGX + xb], eY = tab[kOffGY + yb];
                    float eM = tab[kOffE + xb * 5 + yb];
                    if (kp.rle) eM *= tab[kOffRL + xb * 2601 + min(sx >> 8, 50u) * 51 + min(sy >> 8, 50u)];
                    const int dL1 = L0 - Lm1, dL2 = L0 - Lm2;
                    float inX, inY, inM;
                    if ((dL1 == 1 || dL1 == -1) && dL2 >= -2 && dL2 <= 2) {
                        inX = xc.fast(oX1, (dL1 - 1) >> 1);
                        inY = xc.fast(oY1, (dL1 + 1) >> 1);
                        inM = xc.fast(oM2, dL2 >> 1);
                    } else {
                        inX = xc.slow(oX1, xmy - 1, Lm1, Rm1, act);
                        inY = xc.slow(oY1, xmy + 1, Lm1, Rm1, act);
                        inM = xc.slow(oM2, xmy, Lm2, Rm2, act);
                    }
                    float p0 = eM * inM, p1 = eX * inX, p2 = eY * inY;
                    if (j == 0 && init) { p0 = iv0; p1 = iv1; p2 = iv2; }
                    if (!act) p0 = p1 = p2 = 0.f;
                    if (j == 0 && dj == d) {
                        // exact power-of-two rescale at the start of each grid cell
                        const float m = row_max(fmaxf(p0, fmaxf(p1, p2)), seg);
                        const int e2 = m > 0.f ? __builtin_amdgcn_frexp_expf(m) : 0;
                        p0 = ldexpf(p0, -e2); p1 = ldexpf(p1, -e2); p2 = ldexpf(p2, -e2);
                        oM1 = ldexpf(oM1, -e2);
                        S += e2;
                    }
                    // send q = T^t p
                    const float q0 = T00 * p0 + T10 * p1 + T20 * p2;
                    const float q1 = T01 * p0 + T11 * p1 + T21 * p2;
                    const float q2 = T02 * p0 + T12 * p1 + T22 * p2;
                    if (!fwdOnly) ringCell[(size_t)j * seg] = make_float4(p0, p1, p2, __int_as_float(S));
                    if (dj == tb) {
                        sv0 = p0; sv1 = p1; sv2 = p2;
                        switchBwd = true;
                    }
                    oM2 = oM1; oX1 = q1; oY1 = q2; oM1 = q0;
                    Lm2 = Lm1; Rm2 = Rm1; Lm1 = L0; Rm1 = R0;
                }
            }
        } else {
#pragma unroll
            for (int j = K - 1; j >= 0; j--) {
                const int dj = gs + j;
                if (dj <= d && dj >= lo + stopOff) {
                    const int L0 = band_L(bC[j]), R0 = band_R(bC[j]);
                    const int xmy = L0 + 2 * k;
                    const bool act = xmy <= R0;
                    const int x = (dj + xmy) >> 1, y = (dj - xmy) >> 1;
                    const int Sin = S;
                    const unsigned sx = sxC[j], sy = syC[j];
                    const unsigned xb = min(sx & 255u, 4u), yb = min(sy & 255u, 4u);
                    const float eX = tab[kOffGX + xb], eY = tab[kOffGY + yb];
                    float eM = tab[kOffE + xb * 5 + yb];
                    if (kp.rle) eM *= tab[kOffRL + xb * 2601 + min(sx >> 8, 50u) * 51 + min(sy >> 8, 50u)];
                    // successors: X at (d+1, xmy+1), Y at (d+1, xmy-1), M at (d+2, xmy)
                    const int dL1 = L0 - Lm1, dL2 = L0 - Lm2;
                    float inX, inY, inM;
                    if ((dL1 == 1 || dL1 == -1) && dL2 >= -2 && dL2 <= 2) {
                        inX = xc.fast(oX1, (dL1 + 1) >> 1);
                        inY = xc.fast(oY1, (dL1 - 1) >> 1);
                        inM = xc.fast(oM2, dL2 >> 1);
                    } else {
                        inX = xc.slow(oX1, xmy + 1, Lm1, Rm1, act);
                        inY = xc.slow(oY1, xmy - 1, Lm1, Rm1, act);
                        inM = xc.slow(oM2, xmy, Lm2, Rm2, act);
                    }
                    // B = T in  (B_s = sum_t T[s][t] in_t)
                    float q0 = T00 * inM + T01 * inX + T02 * inY;
                    float q1 = T10 * inM + T11 * inX + T12 * inY;
                    float q2 = T20 * inM + T21 * inX + T22 * inY;
                    if (init && dj == d) { q0 = iv0; q1 = iv1; q2 = iv2; }
                    if (!act) q0 = q1 = q2 = 0.f;
                    if (j == K - 1 && dj == d && !init) {
                        const float m = row_max(fmaxf(q0, fmaxf(q1, q2)), seg);
                        const int e2 = m > 0.f ? __builtin_amdgcn_frexp_expf(m) : 0;
                        q0 = ldexpf(q0, -e2); q1 = ldexpf(q1, -e2); q2 = ldexpf(q2, -e2);
                        oM1 = ldexpf(oM1, -e2);
                        S += e2;
                    }
                    if (expMode && act) {
                        // expected transition counts, pushed from this cell (the predecessor) into
                        // successors on posterior diagonals (diagonalCalculationExpectations)
                        const float4 f = rC[j];
                        const float sc = ldexpf(totalInv, __float_as_int(f.w) + Sin - SFtb);
                        const bool okXY = dj + 1 > lo && dj + 1 <= from, okM = dj + 2 > lo && dj + 2 <= from;
                        const float tX = okXY ? inX : 0.f, tY = okXY ? inY : 0.f, tM = okM ? inM : 0.f;
                        const float a0 = f.x * sc, a1 = f.y * sc, a2 = f.z * sc;
                        e00 += a0 * (T00 * tM); e01 += a0 * (T01 * tX); e02 += a0 * (T02 * tY);
                        e10 += a1 * (T10 * tM); e11 += a1 * (T11 * tX); e12 += a1 * (T12 * tY);
                        e20 += a2 * (T20 * tM); e21 += a2 * (T21 * tX); e22 += a2 * (T22 * tY);
                    }
                    if (postMode && dj <= from) {
                        const float4 f = rC[j];
                        const float sc = ldexpf(totalInv, __float_as_int(f.w) + S - SFtb);
                        const float pM = f.x * q0 * sc, pX = f.y * q1 * sc, pY = f.z * q2 * sc;
                        const float thr = kp.threshold;
                        const bool okM = act && x > 0 && y > 0 && pM >= thr;
                        const bool okX = kp.withIndels && act && x > 0 && pX >= thr;
                        const bool okY = kp.withIndels && act && y > 0 && pY >= thr;
                        const uint64_t bM = __ballot(okM), bX = __ballot(okX), bY = __ballot(okY);
                        if (okM) {
                            const int idx = cntM + __popcll(bM & belowInRow);
                            if (idx < capM)
                                kp.out[outOff + idx] = make_int4((int)floorf(fminf(pM, 1.f) * 1e7f), 0, x - 1, y - 1);
                        }
                        if (okX) {
                            const int idx = cntX + __popcll(bX & belowInRow);
                            if (idx < capX)
                                kp.out[outOff + capM + idx] =
                                    make_int4((int)floorf(fminf(pX, 1.f) * 1e7f), 0, x - 1, y - 1);
                        }
                        if (okY) {
                            const int idx = cntY + __popcll(bY & belowInRow);
                            if (idx < capY)
                                kp.out[outOff + capM + capX + idx] =
                                    make_int4((int)floorf(fminf(pY, 1.f) * 1e7f), 0, x - 1, y - 1);
                        }
                        cntM += __popcll(bM & rowMask);
                        cntX += __popcll(bX & rowMask);
                        cntY += __popcll(bY & rowMask);
                    }
                    if (dj == lo + stopOff) winDone = true;
                    // send u = e (.) B
                    oM2 = oM1; oX1 = eX * q1; oY1 = eY * q2; oM1 = eM * q0;
                    Lm2 = Lm1; Rm2 = Rm1; Lm1 = L0; Rm1 = R0;
                }
            }
        }
        if (expMode && !fwd) {
            const float ev[9] = {e00, e01, e02, e10, e11, e12, e20, e21, e22};
            for (int i = 0; i < 9; i++) {
                const float v = row_sum(ev[i], seg);
                if (k == 0 && v != 0.f) atomicAdd(&kp.expect[(size_t)job * 10 + i], (double)v);
            }
        }
        init = false;
#pragma unroll
        for (int j = 0; j < K; j++) {
            bC[j] = bN[j]; bN[j] = bN2[j];
            sxC[j] = sxN[j]; syC[j] = syN[j];
            rC[j] = rN[j];
        }

        if (switchBwd) {
            // Window total sum(F(tb) . end); the backward sweep starts at tb.
            const float e0 = ragEnd ? kp.endV[1][0] : kp.endV[0][0];
            const float e1 = ragEnd ? kp.endV[1][1] : kp.endV[0][1];
            const float e2 = ragEnd ? kp.endV[1][2] : kp.endV[0][2];
            const float t = row_sum(sv0 * e0 + sv1 * e1 + sv2 * e2, seg);
            SFtb = S;
            totalInv = t > 0.f ? 1.f / t : 0.f;
            const double lt = log((double)t) + (double)SFtb * 0.69314718055994530942;
            if (atEnd && k == 0) kp.logProb[job] = lt;
            // getExpectations adds the window total once per posterior diagonal (pairwiseAligner.c:674)
            likeAcc += (double)(from - lo) * lt;
            iv0 = e0; iv1 = e1; iv2 = e2;
            fX1 = oX1; fY1 = oY1; fM1 = oM1; fM2 = oM2;
            fLm1 = Lm1; fRm1 = Rm1; fLm2 = Lm2; fRm2 = Rm2; fS = S; fTb = tb;
            oX1 = oY1 = oM1 = oM2 = 0.f;
            Lm1 = 1; Rm1 = -1; Lm2 = 1; Rm2 = -1;
            S = 0; init = true; reload = true;
            d = tb; gs = tb & ~(K - 1);
            phase = PH_BWD;
            if (fwdOnly) {
                if (k == 0) kp.counts[3 * job + 0] = kp.counts[3 * job + 1] = kp.counts[3 * job + 2] = 0;
                phase = PH_FETCH;
            }
        } else if (winDone) {
            if (win + 1 == nWin) {
                if (k == 0) {
                    kp.counts[3 * job + 0] = cntM;
                    kp.counts[3 * job + 1] = cntX;
                    kp.counts[3 * job + 2] = cntY;
                    if (expMode) kp.expect[(size_t)job * 10 + 9] = likeAcc;
                }
                phase = PH_FETCH;
            } else {
                win += 1;
                const DevWin w1 = kp.win[winOff + win];
                d = fTb + 1; gs = d & ~(K - 1);
                tb = w1.tb; lo = w1.lo; from = w1.from;
                oX1 = fX1; oY1 = fY1; oM1 = fM1; oM2 = fM2;
                Lm1 = fLm1; Rm1 = fRm1; Lm2 = fLm2; Rm2 = fRm2; S = fS;
                reload = true;
                phase = PH_FWD;
            }
        } else {
            gs += step;
            d = fwd ? gs : gs + K - 1;
        }
    }
}

// ---------------------------------------------------------------------------
// Wide bands (more than 64 cells on some anti-diagonal: anchor-free calls,
// large band half-widths).  One job per wavefront; cell i = lane + 64c for
// c < C.  The (X, Y, M) outputs of the last three anti-diagonals live in LDS
// and are read by band coordinate; one anti-diagonal per step.  Scaling is
// lazy: the exponent chosen at a grid-cell start comes from the previous
// anti-diagonal's maximum, and the two-back match input is corrected by the
// same exponent on the next step.
// ---------------------------------------------------------------------------

struct WideCtx {
    const KParams &kp;
    float *xb;  // [3 slots][3 states][W]
    int W, C, lane;
    int64_t symX, symY, bandOff;
    const float *tab;
    __device__ float *slot(int d, int st) const { return xb + ((size_t)(((d % 3) + 3) % 3) * 3 + st) * W; }
    __device__ void band(int d, int &L, int &R) const {
        const int b = kp.band[bandOff + d];
        L = band_L(b);
        R = band_R(b);
    }
};

__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float wave_sum(float v) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_sum_d(double v) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ __launch_bounds__(64) void pairhmm_wide_kernel(KParams kp) {
    extern __shared__ float xbuf[];
    const int lane = threadIdx.x;
    const int C = kp.wideC, W = 64 * C;
    float4 *ring = reinterpret_cast<float4 *>(kp.ringWide) + (size_t)blockIdx.x * kp.ringWideLen * W;
    const int R = kp.ringWideLen;
    const uint64_t below = (1ull << lane) - 1ull;
    const bool expMode = kp.mode == MP_MODE_EXPECTATIONS, fwdOnly = kp.mode == MP_MODE_FORWARD_ONLY;
    const bool postMode = kp.mode <= MP_MODE_MATCH_ONLY;
    const int stopOff = expMode ? 0 : 1;
    const float T[9] = {kp.T[0], kp.T[1], kp.T[2], kp.T[3], kp.T[4], kp.T[5], kp.T[6], kp.T[7], kp.T[8]};

    for (;;) {
        int jn = 0;
        if (lane == 0) jn = atomicAdd(&kp.queue[3], 1);
        jn = __shfl(jn, 0, 64);
        const int job = kp.wideBegin + jn;
        if (job >= kp.wideEnd) break;
        const DevJob jb = kp.jobs[job];
        const int D = jb.lX + jb.lY;
        const bool rgL = (jb.flags >> 1) & 1, rgR = (jb.flags >> 2) & 1;
        WideCtx cx{kp, xbuf, W, C, lane, jb.symX, jb.symY, jb.bandOff, kp.tables + ((jb.flags & 1) ? kTabStride : 0)};
        int cntM = 0, cntX = 0, cntY = 0;
        double likeAcc = 0.0;
        double eacc[9];
        for (int i = 0; i < 9; i++) eacc[i] = 0.0;
        int S = 0;          // forward exponent
        float lastMax = 0.75f;
        int pendM = 0;      // exponent correction of the two-back match input
        int dStart = 0;     // first forward anti-diagonal of this window
        for (int i = lane; i < 9 * W; i += 64) xbuf[i] = 0.f;
        __syncthreads();
        for (int win = 0; win < jb.nWin; win++) {
            const DevWin wn = kp.win[jb.winOff + win];
            const bool atEnd = wn.tb == D;
            // ---- forward from dStart to tb ----
            for (int d = dStart; d <= wn.tb; d++) {
                int e2 = 0;
                if (d > 0 && (d & (kK - 1)) == 0) {
                    const float m = wave_max(lastMax);
                    e2 = m > 0.f ? __builtin_amdgcn_frexp_expf(m) : 0;
                }
                const float scl = ldexpf(1.f, -e2), sclM = ldexpf(1.f, -e2 - pendM);
                int L0, R0, L1 = 1, R1 = -1, L2 = 1, R2 = -1;
                cx.band(d, L0, R0);
                if (d >= 1) cx.band(d - 1, L1, R1);
                if (d >= 2) cx.band(d - 2, L2, R2);
                float mx = 0.f;
                for (int c = 0; c < C; c++) {
                    const int i = lane + 64 * c, xmy = L0 + 2 * i;
                    const bool act = xmy <= R0;
                    const int x = (d + xmy) >> 1, y = (d - xmy) >> 1;
                    float p0 = 0.f, p1 = 0.f, p2 = 0.f;
                    if (act) {
                        const unsigned sx = x > 0 ? kp.sym[jb.symX + x - 1] : 4u, sy = y > 0 ? kp.sym[jb.symY + y - 1] : 4u;
                        const unsigned xb = min(sx & 255u, 4u), yb = min(sy & 255u, 4u);
                        float eM = cx.tab[kOffE + xb * 5 + yb];
                        if (kp.rle) eM *= cx.tab[kOffRL + xb * 2601 + min(sx >> 8, 50u) * 51 + min(sy >> 8, 50u)];
                        if (d == 0) {
                            p0 = rgL ? kp.startV[1][0] : kp.startV[0][0];
                            p1 = rgL ? kp.startV[1][1] : kp.startV[0][1];
                            p2 = rgL ? kp.startV[1][2] : kp.startV[0][2];
                        } else {
                            const int sX = xmy - 1, sY = xmy + 1;
                            const float inX = (sX >= L1 && sX <= R1) ? cx.slot(d - 1, 1)[(sX - L1) >> 1] : 0.f;
                            const float inY = (sY >= L1 && sY <= R1) ? cx.slot(d - 1, 2)[(sY - L1) >> 1] : 0.f;
                            const float inM = (xmy >= L2 && xmy <= R2) ? cx.slot(d - 2, 0)[(xmy - L2) >> 1] : 0.f;
                            p0 = eM * inM * sclM;
                            p1 = cx.tab[kOffGX + xb] * inX * scl;
                            p2 = cx.tab[kOffGY + yb] * inY * scl;
                        }
                    }
                    mx = fmaxf(mx, fmaxf(p0, fmaxf(p1, p2)));
                    cx.slot(d, 0)[i] = T[0] * p0 + T[3] * p1 + T[6] * p2;
                    cx.slot(d, 1)[i] = T[1] * p0 + T[4] * p1 + T[7] * p2;
                    cx.slot(d, 2)[i] = T[2] * p0 + T[5] * p1 + T[8] * p2;
                    if (!fwdOnly) ring[(size_t)(d % R) * W + i] = make_float4(p0, p1, p2, __int_as_float(S + e2));
                    if (d == wn.tb) {
                        const float e0 = rgR && atEnd ? kp.endV[1][0] : kp.endV[0][0];
                        const float e1 = rgR && atEnd ? kp.endV[1][1] : kp.endV[0][1];
                        const float e2v = rgR && atEnd ? kp.endV[1][2] : kp.endV[0][2];
                        lastMax = p0 * e0 + p1 * e1 + p2 * e2v + (c == 0 ? 0.f : lastMax);  // partial total
                    }
                }
                S += e2;
                pendM = e2;
                if (d != wn.tb) lastMax = mx;
                __syncthreads();
            }
            const float t = wave_sum(lastMax);
            const int SFtb = S;
            const float totalInv = t > 0.f ? 1.f / t : 0.f;
            const double lt = log((double)t) + (double)SFtb * 0.69314718055994530942;
            if (atEnd && lane == 0) kp.logProb[job] = lt;
            likeAcc += (double)(wn.from - wn.lo) * lt;
            if (fwdOnly) break;
            // ---- backward from tb down to lo + stopOff ----
            for (int i = lane; i < 9 * W; i += 64) xbuf[i] = 0.f;
            __syncthreads();
            int SB = 0, pendMB = 0;
            float lastMaxB = 0.75f;
            for (int d = wn.tb; d >= wn.lo + stopOff; d--) {
                int e2 = 0;
                if (d < wn.tb && ((d + 1) & (kK - 1)) == 0) {
                    const float m = wave_max(lastMaxB);
                    e2 = m > 0.f ? __builtin_amdgcn_frexp_expf(m) : 0;
                }
                const float scl = ldexpf(1.f, -e2);
                int L0, R0, L1 = 1, R1 = -1, L2 = 1, R2 = -1;
                cx.band(d, L0, R0);
                if (d + 1 <= wn.tb) cx.band(d + 1, L1, R1);
                if (d + 2 <= wn.tb) cx.band(d + 2, L2, R2);
                float mx = 0.f;
                float ep[9];
                for (int q = 0; q < 9; q++) ep[q] = 0.f;
                for (int c = 0; c < C; c++) {
                    const int i = lane + 64 * c, xmy = L0 + 2 * i;
                    const bool act = xmy <= R0;
                    const int x = (d + xmy) >> 1, y = (d - xmy) >> 1;
                    float q0 = 0.f, q1 = 0.f, q2 = 0.f, eX = 0.f, eY = 0.f, eM = 0.f, inX = 0.f, inY = 0.f, inM = 0.f;
                    if (act) {
                        const unsigned sx = x > 0 ? kp.sym[jb.symX + x - 1] : 4u, sy = y > 0 ? kp.sym[jb.symY + y - 1] : 4u;
                        const unsigned xb = min(sx & 255u, 4u), yb = min(sy & 255u, 4u);
                        eX = cx.tab[kOffGX + xb];
                        eY = cx.tab[kOffGY + yb];
                        eM = cx.tab[kOffE + xb * 5 + yb];
                        if (kp.rle) eM *= cx.tab[kOffRL + xb * 2601 + min(sx >> 8, 50u) * 51 + min(sy >> 8, 50u)];
                        if (d == wn.tb) {
                            const bool rg = atEnd && rgR;
                            q0 = rg ? kp.endV[1][0] : kp.endV[0][0];
                            q1 = rg ? kp.endV[1][1] : kp.endV[0][1];
                            q2 = rg ? kp.endV[1][2] : kp.endV[0][2];
                        } else {
                            const int sX = xmy + 1, sY = xmy - 1;
                            inX = (sX >= L1 && sX <= R1) ? cx.slot(d + 1, 1)[(sX - L1) >> 1] : 0.f;
                            inY = (sY >= L1 && sY <= R1) ? cx.slot(d + 1, 2)[(sY - L1) >> 1] : 0.f;
                            inM = (xmy >= L2 && xmy <= R2) ? cx.slot(d + 2, 0)[(xmy - L2) >> 1] * ldexpf(1.f, -pendMB) : 0.f;
                            q0 = (T[0] * inM + T[1] * inX + T[2] * inY) * scl;
                            q1 = (T[3] * inM + T[4] * inX + T[5] * inY) * scl;
                            q2 = (T[6] * inM + T[7] * inX + T[8] * inY) * scl;
                        }
                    }
                    mx = fmaxf(mx, fmaxf(q0, fmaxf(q1, q2)));
                    const float4 f = act ? ring[(size_t)(d % R) * W + i] : make_float4(0.f, 0.f, 0.f, 0.f);
                    if (expMode && act) {
                        const float sc = ldexpf(totalInv, __float_as_int(f.w) + SB - SFtb);
                        const bool okXY = d + 1 > wn.lo && d + 1 <= wn.from, okM = d + 2 > wn.lo && d + 2 <= wn.from;
                        const float tX = okXY ? inX : 0.f, tY = okXY ? inY : 0.f, tM = okM ? inM : 0.f;
                        const float a0 = f.x * sc, a1 = f.y * sc, a2 = f.z * sc;
                        ep[0] += a0 * T[0] * tM; ep[1] += a0 * T[1] * tX; ep[2] += a0 * T[2] * tY;
                        ep[3] += a1 * T[3] * tM; ep[4] += a1 * T[4] * tX; ep[5] += a1 * T[5] * tY;
                        ep[6] += a2 * T[6] * tM; ep[7] += a2 * T[7] * tX; ep[8] += a2 * T[8] * tY;
                    }
                    if (postMode && d <= wn.from) {
                        const float sc = ldexpf(totalInv, __float_as_int(f.w) + SB + e2 - SFtb);
                        const float pM = f.x * q0 * sc, pX = f.y * q1 * sc, pY = f.z * q2 * sc;
                        const float thr = kp.threshold;
                        const bool okM = act && x > 0 && y > 0 && pM >= thr;
                        const bool okX = kp.withIndels && act && x > 0 && pX >= thr;
                        const bool okY = kp.withIndels && act && y > 0 && pY >= thr;
                        const uint64_t bM = __ballot(okM), bX = __ballot(okX), bY = __ballot(okY);
                        if (okM) {
                            const int idx = cntM + __popcll(bM & below);
                            if (idx < jb.cap[0])
                                kp.out[jb.outOff + idx] = make_int4((int)floorf(fminf(pM, 1.f) * 1e7f), 0, x - 1, y - 1);
                        }
                        if (okX) {
                            const int idx = cntX + __popcll(bX & below);
                            if (idx < jb.cap[1])
                                kp.out[jb.outOff + jb.cap[0] + idx] =
                                    make_int4((int)floorf(fminf(pX, 1.f) * 1e7f), 0, x - 1, y - 1);
                        }
                        if (okY) {
                            const int idx = cntY + __popcll(bY & below);
                            if (idx < jb.cap[2])
                                kp.out[jb.outOff + jb.cap[0] + jb.cap[1] + idx] =
                                    make_int4((int)floorf(fminf(pY, 1.f) * 1e7f), 0, x - 1, y - 1);
                        }
                        cntM += __popcll(bM);
                        cntX += __popcll(bX);
                        cntY += __popcll(bY);
                    }
                    cx.slot(d, 0)[i] = eM * q0;
                    cx.slot(d, 1)[i] = eX * q1;
                    cx.slot(d, 2)[i] = eY * q2;
                }
                if (expMode)
                    for (int q = 0; q < 9; q++) eacc[q] += (double)ep[q];
                SB += e2;
                pendMB = e2;
                lastMaxB = mx;
                __syncthreads();
            }
            // ---- resume the forward sweep after tb: rebuild its exchange state from the ring ----
            if (win + 1 < jb.nWin) {
                const int tb = wn.tb;
                int L0, R0, L1, R1;
                cx.band(tb, L0, R0);
                cx.band(tb - 1, L1, R1);
                for (int c = 0; c < C; c++) {
                    const int i = lane + 64 * c;
                    const float4 f = ring[(size_t)(tb % R) * W + i];
                    const float4 g = ring[(size_t)((tb - 1) % R) * W + i];
                    const bool a0 = L0 + 2 * i <= R0, a1 = L1 + 2 * i <= R1;
                    const float sg = ldexpf(1.f, __float_as_int(g.w) - __float_as_int(f.w));
                    cx.slot(tb, 0)[i] = a0 ? T[0] * f.x + T[3] * f.y + T[6] * f.z : 0.f;
                    cx.slot(tb, 1)[i] = a0 ? T[1] * f.x + T[4] * f.y + T[7] * f.z : 0.f;
                    cx.slot(tb, 2)[i] = a0 ? T[2] * f.x + T[5] * f.y + T[8] * f.z : 0.f;
                    cx.slot(tb - 1, 0)[i] = a1 ? (T[0] * g.x + T[3] * g.y + T[6] * g.z) * sg : 0.f;
                }
                __syncthreads();
                dStart = tb + 1;
                pendM = 0;
                lastMax = 0.75f;
            }
        }
        if (lane == 0) {
            kp.counts[3 * job + 0] = cntM;
            kp.counts[3 * job + 1] = cntX;
            kp.counts[3 * job + 2] = cntY;
        }
        if (expMode) {
            for (int q = 0; q < 9; q++) {
                const double v = wave_sum_d(eacc[q]);
                if (lane == 0) kp.expect[(size_t)job * 10 + q] = v;
            }
            if (lane == 0) kp.expect[(size_t)job * 10 + 9] = likeAcc;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void pairhmm_kernel(KParams kp) {
    extern __shared__ float lds[];
    for (int i = threadIdx.x; i < 2 * kTabStride; i += blockDim.x) lds[i] = kp.tables[i];
    __syncthreads();
    for (int cls = kNumClasses - 1; cls >= 0; cls--) run_class(kp, lds, cls, 16 << cls);
}

__global__ void fill_i32(int32_t *p, int32_t v, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

#define HIP_TRY(call)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            mp::set_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #call); \
            return MP_EHIP;                                                               \
        }                                                                                 \
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
    int32_t cls;
    int64_t cells;
    int32_t cap[3];
    int64_t outOff;
    int32_t ringNeed;
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
    mp_status upload(const std::vector<T> &v, hipStream_t s) {
        mp_status st = alloc(v.size());
        if (st != MP_OK || v.empty()) return st;
        HIP_TRY(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
        return MP_OK;
    }
};

}  // namespace

struct mp_engine {
    int device = 0;
    int cus = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

struct mp_batch {
    mp_engine *eng = nullptr;
    int32_t mode = 0;
    int32_t flags = 0;
    int64_t npairs = 0;
    std::vector<HostJob> jobs;         // in device order
    std::vector<int64_t> pairCrop;     // reference crop offset per pair
    std::vector<int64_t> pairFirstJob; // jobs of a pair are consecutive in split order in `order`
    std::vector<int32_t> order;        // pair-major list of device job indices (split order)
    std::vector<int32_t> pairNJobs;
    int64_t cells = 0;
    int32_t ringLen = 1;
    int32_t classBegin[kNumClasses] = {0, 0, 0}, classEnd[kNumClasses] = {0, 0, 0};
    KParams kp;
    DevBuf<DevJob> dJobs;
    DevBuf<uint16_t> dSym;
    DevBuf<int32_t> dBand;
    DevBuf<DevWin> dWin;
    DevBuf<int4> dOut;
    DevBuf<int32_t> dCounts;
    DevBuf<double> dLogProb;
    DevBuf<float> dRing;
    DevBuf<int32_t> dRingS;
    DevBuf<float> dTables;
    DevBuf<int32_t> dQueue;
    DevBuf<double> dExpect;
    DevBuf<float> dRingWide;
    int32_t wideBegin = 0, wideEnd = 0, wideC = 0, ringWideLen = 0, wideWaves = 0;
    int64_t outTotal = 0;
    int32_t gridRows = 0;
    double kernelMs = 0.0;
    bool ran = false;
    std::vector<int32_t> hCounts;  // fetched
};

namespace {

mp_status launch(mp_batch *b, int grid, int block, const KParams &kp) {
    size_t lds = 2 * kTabStride * sizeof(float);
    hipLaunchKernelGGL(pairhmm_kernel, dim3(grid), dim3(block), lds, b->eng->stream, kp);
    HIP_TRY(hipGetLastError());
    return MP_OK;
}

mp_status launch_wide(mp_batch *b, int grid, const KParams &kp) {
    size_t lds = (size_t)9 * 64 * kp.wideC * sizeof(float);
    hipLaunchKernelGGL(pairhmm_wide_kernel, dim3(grid), dim3(64), lds, b->eng->stream, kp);
    HIP_TRY(hipGetLastError());
    return MP_OK;
}

mp_status reset_queue(mp_batch *b, int32_t *q, int n) {
    hipLaunchKernelGGL(fill_i32, dim3((n + 255) / 256), dim3(256), 0, b->eng->stream, q, 0, n);
    HIP_TRY(hipGetLastError());
    return MP_OK;
}

void build_tables(const mp_model *m, std::vector<float> &tab, KParams &kp) {
    tab.assign(2 * kTabStride, 0.f);
    for (int s = 0; s < 2; s++) {
        const mp::Machine &mc = m->mach[s];
        float *t = tab.data() + s * kTabStride;
        for (int i = 0; i < 5; i++)
            for (int j = 0; j < 5; j++) {
                double lp = (i >= 4 || j >= 4) ? -2.772588722 : mc.match[i * 4 + j];
                t[kOffE + i * 5 + j] = (float)std::exp(lp);
            }
        for (int i = 0; i < 5; i++) {
            t[kOffGX + i] = (float)std::exp(i >= 4 ? -1.386294361 : mc.gapX[i]);
            t[kOffGY + i] = (float)std::exp(i >= 4 ? -1.386294361 : mc.gapY[i]);
        }
        for (int b = 0; b < 5; b++) {
            int bb = b == 4 ? 0 : b;
            int row = s ? bb : 3 - bb;
            for (int u = 0; u < mp::kMaxRepeat; u++)
                for (int o = 0; o < mp::kMaxRepeat; o++) {
                    double l10 = m->repeatLog10[(size_t)row * mp::kMaxRepeat * mp::kMaxRepeat +
                                                (size_t)u * mp::kMaxRepeat + o];
                    t[kOffRL + b * 2601 + u * 51 + o] = (float)std::exp(2.3025 * l10);
                }
        }
    }
    const mp::Machine &f = m->mach[1];
    auto P = [&](int i) { return (float)std::exp(f.trans[i]); };
    // T[from*3 + to]
    kp.T[0] = P(mp::kMatchContinue); kp.T[1] = P(mp::kGapOpenX);     kp.T[2] = P(mp::kGapOpenY);
    kp.T[3] = P(mp::kMatchFromGapX); kp.T[4] = P(mp::kGapExtendX);   kp.T[5] = P(mp::kGapSwitchToY);
    kp.T[6] = P(mp::kMatchFromGapY); kp.T[7] = P(mp::kGapSwitchToX); kp.T[8] = P(mp::kGapExtendY);
    kp.startV[0][0] = 1.f; kp.startV[0][1] = 0.f; kp.startV[0][2] = 0.f;
    kp.startV[1][0] = 0.f; kp.startV[1][1] = 1.f; kp.startV[1][2] = 1.f;
    kp.endV[0][0] = P(mp::kMatchContinue);
    kp.endV[0][1] = P(mp::kMatchFromGapX);
    kp.endV[0][2] = P(mp::kMatchFromGapY);
    kp.endV[1][0] = (float)std::exp((f.trans[mp::kGapOpenX] + f.trans[mp::kGapOpenY]) / 2.0);
    kp.endV[1][1] = P(mp::kGapExtendX);
    kp.endV[1][2] = P(mp::kGapExtendY);
    kp.rle = f.rle;
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
    HIP_TRY(hipFuncSetAttribute((const void *)pairhmm_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                2 * kTabStride * sizeof(float)));
    HIP_TRY(hipFuncSetAttribute((const void *)pairhmm_wide_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                9 * 64 * kWideMaxC * sizeof(float)));
    mp_engine *e = new mp_engine();
    e->device = device;
    e->cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess) {
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
    if (e->ev0) (void)hipEventDestroy(e->ev0);
    if (e->ev1) (void)hipEventDestroy(e->ev1);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

extern "C" int32_t mp_engine_cu_count(const mp_engine *e) { return e ? e->cus : 0; }

extern "C" mp_status mp_batch_create(mp_engine *e, const mp_model *m, const mp_pair *pairs, int64_t npairs,
                                     const mp_band_params *p, int32_t mode, int32_t flags, mp_batch **out) {
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
    *out = nullptr;
    HIP_TRY(hipSetDevice(e->device));
    mp_batch *b = new mp_batch();
    b->eng = e;
    b->mode = mode;
    b->flags = flags;
    b->npairs = npairs;
    b->pairCrop.assign(npairs, 0);
    b->pairNJobs.assign(npairs, 0);
    b->pairFirstJob.assign(npairs, 0);

    // ---- 1. symbols, cropping, split points -> jobs (pair order) ----
    // Symbol pool with N padding at both ends (the kernel reads up to one
    // row width beyond a sequence without clamping).
    std::vector<uint16_t> sym(kSymPad, 4);
    std::vector<HostJob> jobs;
    std::vector<std::vector<int64_t>> jobAnchors;
    std::vector<int64_t> sp, anc;
    for (int64_t i = 0; i < npairs; i++) {
        const mp_pair &pr = pairs[i];
        if (pr.lX < 0 || pr.lY < 0 || (pr.lX > 0 && !pr.sX) || (pr.lY > 0 && !pr.sY) ||
            (pr.nAnchors > 0 && !pr.anchors) || pr.nAnchors < 0 || pr.lX > INT32_MAX / 2 || pr.lY > INT32_MAX / 2) {
            delete b;
            mp::set_error("pair " + std::to_string(i) + ": bad sequence or anchor arguments");
            return MP_EINVAL;
        }
        if (!mp::anchors_valid(pr.anchors, pr.nAnchors, pr.lX, pr.lY)) {
            delete b;
            mp::set_error("pair " + std::to_string(i) + ": anchors must be strictly increasing, inside the "
                          "sequences, with even non-negative expansion");
            return MP_EANCHORS;
        }
        int64_t first = 0, end = pr.lX;
        if ((flags & MP_BATCH_CROP_REFERENCE) && pr.nAnchors > 0) {
            const int64_t *a0 = pr.anchors, *al = pr.anchors + 3 * (pr.nAnchors - 1);
            first = std::max<int64_t>(a0[0] - a0[1], 0);
            end = std::min<int64_t>(1 + al[0] + (pr.lY - al[1]), pr.lX);
        }
        b->pairCrop[i] = first;
        const int64_t lX = end - first, lY = pr.lY;
        anc.assign(pr.anchors ? pr.anchors : nullptr, pr.anchors ? pr.anchors + 3 * pr.nAnchors : nullptr);
        for (int64_t a = 0; a < pr.nAnchors; a++) anc[3 * a] -= first;
        const int64_t baseX = (int64_t)sym.size();
        sym.insert(sym.end(), pr.sX + first, pr.sX + end);
        const int64_t baseY = (int64_t)sym.size();
        sym.insert(sym.end(), pr.sY, pr.sY + lY);
        if (mode == MP_MODE_FORWARD_ONLY) {  // computeForwardProbability does not split
            sp.assign({0, 0, lX, lY});
        } else {
            mp::split_points(anc.data(), pr.nAnchors, lX, lY, p->splitMatrixBiggerThanThis, pr.raggedLeft != 0,
                             pr.raggedRight != 0, sp);
        }
        const int64_t ns = (int64_t)sp.size() / 4;
        b->pairFirstJob[i] = (int64_t)jobs.size();
        int64_t j = 0;
        for (int64_t s = 0; s < ns; s++) {
            const int64_t x1 = sp[4 * s], y1 = sp[4 * s + 1], x2 = sp[4 * s + 2], y2 = sp[4 * s + 3];
            std::vector<int64_t> sub;
            while (j < pr.nAnchors) {
                const int64_t ax = anc[3 * j], ay = anc[3 * j + 1];
                if (ax + ay >= x2 + y2) break;
                sub.push_back(ax - x1);
                sub.push_back(ay - y1);
                sub.push_back(anc[3 * j + 2]);
                j++;
            }
            HostJob hj{};
            hj.pair = i;
            hj.xOff = x1;
            hj.yOff = y1;
            hj.symX = baseX + x1;
            hj.symY = baseY + y1;
            hj.lX = (int32_t)(x2 - x1);
            hj.lY = (int32_t)(y2 - y1);
            const bool rl = pr.raggedLeft || s > 0, rr = pr.raggedRight || s < ns - 1;
            hj.flags = (pr.strand ? 1 : 0) | (rl ? 2 : 0) | (rr ? 4 : 0);
            jobs.push_back(hj);
            jobAnchors.push_back(std::move(sub));
        }
        b->pairNJobs[i] = (int32_t)ns;
    }

    sym.resize(sym.size() + kSymPad, 4);

    // ---- 2. bands, windows, classes (parallel over jobs) ----
    const int64_t nj = (int64_t)jobs.size();
    std::vector<int64_t> bandOff(nj + 1, 0);
    // Each job's band padded to whole cells of kK anti-diagonals (16-byte aligned).
    for (int64_t j = 0; j < nj; j++) bandOff[j + 1] = bandOff[j] + ((jobs[j].lX + jobs[j].lY) / kK + 1) * kK;
    std::vector<int32_t> band((size_t)bandOff[nj]);
    std::vector<std::vector<mp::Window>> wins(nj);
    std::atomic<int> err{0};
    {
        const int nth = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        std::atomic<int64_t> next{0};
        for (int t = 0; t < nth; t++)
            th.emplace_back([&]() {
                std::vector<int32_t> L, R;
                for (;;) {
                    int64_t j = next.fetch_add(1);
                    if (j >= nj) break;
                    HostJob &hj = jobs[j];
                    const int64_t D = hj.lX + hj.lY;
                    L.resize(D + 1);
                    R.resize(D + 1);
                    const std::vector<int64_t> &a = jobAnchors[j];
                    const bool dyn = mode != MP_MODE_FORWARD_ONLY && p->dynamicAnchorExpansion != 0;
                    if (!mp::build_band(a.data(), (int64_t)a.size() / 3, hj.lX, hj.lY, p->diagonalExpansion, dyn,
                                        L.data(), R.data())) {
                        err = 1;
                        continue;
                    }
                    int32_t wmax = 0;
                    int64_t cells = 0;
                    for (int64_t d = 0; d <= D; d++) {
                        int32_t w = (R[d] - L[d]) / 2 + 1;
                        wmax = std::max(wmax, w);
                        cells += w;
                        band[bandOff[j] + d] = (int32_t)((uint32_t)L[d] << 12) | (std::min(w, 4096) - 1);
                    }
                    for (int64_t d = D + 1; d < bandOff[j + 1] - bandOff[j]; d++) band[bandOff[j] + d] = band[bandOff[j] + D];
                    hj.wmax = wmax;
                    hj.cells = D > 0 ? cells : 0;
                    hj.cls = wmax <= 16 ? 0 : (wmax <= 32 ? 1 : (wmax <= 64 ? 2 : 3));
                    hj.bandOff = bandOff[j];
                    if (D > 0 && mode == MP_MODE_FORWARD_ONLY) wins[j].assign(1, mp::Window{(int32_t)D, 0, (int32_t)D});
                    else if (D > 0) mp::build_windows(L.data(), R.data(), D, *p, wins[j]);
                    // live forward entries: (lo, tb] of a window; whole cells plus margin
                    int32_t need = 1;
                    for (auto &w : wins[j]) need = std::max(need, w.tb - w.lo + 2 * kK);
                    hj.ringNeed = (need + kK - 1) / kK * kK;
                }
            });
        for (auto &t : th) t.join();
    }
    if (err) {
        delete b;
        mp::set_error("invalid band (anchors inconsistent with the sequences)");
        return MP_EANCHORS;
    }
    for (int64_t j = 0; j < nj; j++) {
        if (jobs[j].wmax > 64 * kWideMaxC) {
            delete b;
            mp::set_error("pair " + std::to_string(jobs[j].pair) + ": band width " + std::to_string(jobs[j].wmax) +
                          " exceeds the " + std::to_string(64 * kWideMaxC) + "-cell kernel limit");
            return MP_EUNSUPPORTED;
        }
    }

    // ---- 3. device order: by class, then band area descending ----
    std::vector<int32_t> ord;
    for (int64_t j = 0; j < nj; j++)
        if (jobs[j].lX + jobs[j].lY > 0) ord.push_back((int32_t)j);
    std::stable_sort(ord.begin(), ord.end(), [&](int32_t a, int32_t c) {
        if (jobs[a].cls != jobs[c].cls) return jobs[a].cls < jobs[c].cls;
        return jobs[a].cells > jobs[c].cells;
    });
    std::vector<int32_t> devIndex(nj, -1);
    std::vector<DevJob> dj(ord.size());
    std::vector<DevWin> dw;
    int64_t outOff = 0;
    int32_t ringLen = 4 * kK;
    for (size_t q = 0; q < ord.size(); q++) {
        HostJob &hj = jobs[ord[q]];
        devIndex[ord[q]] = (int32_t)q;
        const int64_t D = hj.lX + hj.lY;
        hj.cap[0] = mode <= MP_MODE_MATCH_ONLY ? (int32_t)std::min<int64_t>(2 * D + 64, INT32_MAX / 4) : 0;
        hj.cap[1] = mode == MP_MODE_POSTERIORS ? (int32_t)std::min<int64_t>(D + 64, INT32_MAX / 4) : 0;
        hj.cap[2] = hj.cap[1];
        hj.outOff = outOff;
        outOff += (int64_t)hj.cap[0] + hj.cap[1] + hj.cap[2];
        hj.winOff = (int32_t)dw.size();
        hj.nWin = (int32_t)wins[ord[q]].size();
        for (auto &w : wins[ord[q]]) dw.push_back(DevWin{w.tb, w.lo, w.from, 0});
        if (mode != MP_MODE_FORWARD_ONLY && hj.cls < 3) ringLen = std::max(ringLen, hj.ringNeed);
        if (hj.cls == 3) {
            b->wideC = std::max(b->wideC, (hj.wmax + 63) / 64);
            if (mode != MP_MODE_FORWARD_ONLY) b->ringWideLen = std::max(b->ringWideLen, hj.ringNeed);
        }
        DevJob &d = dj[q];
        d.symX = hj.symX; d.symY = hj.symY; d.bandOff = hj.bandOff; d.outOff = hj.outOff;
        d.lX = hj.lX; d.lY = hj.lY; d.winOff = hj.winOff; d.nWin = hj.nWin;
        d.cap[0] = hj.cap[0]; d.cap[1] = hj.cap[1]; d.cap[2] = hj.cap[2];
        d.flags = hj.flags;
        b->cells += hj.cells;
    }
    for (int c = 0; c < kNumClasses; c++) {
        b->classBegin[c] = (int32_t)ord.size();
        b->classEnd[c] = (int32_t)ord.size();
    }
    for (int c = kNumClasses - 1; c >= 0; c--)
        for (size_t q = 0; q < ord.size(); q++)
            if (jobs[ord[q]].cls == c) {
                b->classBegin[c] = std::min<int32_t>(b->classBegin[c], (int32_t)q);
                b->classEnd[c] = (int32_t)q + 1;
            }
    for (int c = 0; c < kNumClasses; c++)
        if (b->classBegin[c] > b->classEnd[c]) b->classBegin[c] = b->classEnd[c];
    b->wideBegin = b->wideEnd = (int32_t)ord.size();
    for (size_t q = 0; q < ord.size(); q++)
        if (jobs[ord[q]].cls == 3) {
            b->wideBegin = std::min<int32_t>(b->wideBegin, (int32_t)q);
            b->wideEnd = (int32_t)q + 1;
        }
    // pair-major job list in split order (device indices; -1 for empty jobs)
    b->order.resize(nj);
    for (int64_t j = 0; j < nj; j++) b->order[j] = devIndex[j];
    b->jobs.resize(ord.size());
    for (size_t q = 0; q < ord.size(); q++) b->jobs[q] = jobs[ord[q]];
    b->ringLen = ringLen;
    b->outTotal = outOff;

    // ---- 4. upload ----
    hipStream_t s = e->stream;
    std::vector<float> tab;
    KParams &kp = b->kp;
    std::memset(&kp, 0, sizeof(kp));
    build_tables(m, tab, kp);
    mp_status st;
    if ((st = b->dJobs.upload(dj, s)) || (st = b->dSym.upload(sym, s)) || (st = b->dBand.upload(band, s)) ||
        (st = b->dWin.upload(dw, s)) || (st = b->dTables.upload(tab, s)) || (st = b->dOut.alloc(std::max<int64_t>(outOff, 1))) ||
        (st = b->dCounts.alloc(std::max<size_t>(3 * dj.size(), 1))) ||
        (st = b->dLogProb.alloc(std::max<size_t>(dj.size(), 1))) || (st = b->dQueue.alloc(kNumClasses + 1)) ||
        (st = b->dExpect.alloc(mode == MP_MODE_EXPECTATIONS ? std::max<size_t>(10 * dj.size(), 1) : 1))) {
        delete b;
        return st;
    }
    // Ring: one slice per resident wavefront.
    b->gridRows = e->cus;
    const size_t waves = (size_t)e->cus * (kBlock / 64);
    if ((st = b->dRing.alloc(waves * (size_t)ringLen * kRingStride))) {
        delete b;
        return st;
    }
    if (b->wideEnd > b->wideBegin) {
        // wide jobs: one wavefront each, ring of float4 per cell; bounded to ~48 GB
        const size_t perWave = (size_t)std::max(b->ringWideLen, 1) * 64 * b->wideC * sizeof(float4);
        const size_t cap = std::max<size_t>(1, ((size_t)48 << 30) / perWave);
        b->wideWaves = (int32_t)std::min<size_t>({(size_t)(b->wideEnd - b->wideBegin), (size_t)e->cus * 4, cap});
        if ((st = b->dRingWide.alloc((size_t)b->wideWaves * perWave / sizeof(float)))) {
            delete b;
            return st;
        }
    }
    kp.jobs = b->dJobs.p;
    kp.sym = b->dSym.p;
    kp.band = b->dBand.p;
    kp.win = b->dWin.p;
    kp.out = b->dOut.p;
    kp.counts = b->dCounts.p;
    kp.logProb = b->dLogProb.p;
    kp.ring = b->dRing.p;
    kp.ringWide = b->dRingWide.p;
    kp.ringWideLen = std::max(b->ringWideLen, 1);
    kp.wideC = std::max(b->wideC, 1);
    kp.wideBegin = b->wideBegin;
    kp.wideEnd = b->wideEnd;
    kp.tables = b->dTables.p;
    kp.queue = b->dQueue.p;
    kp.ringLen = ringLen;
    kp.withIndels = mode == MP_MODE_POSTERIORS;
    kp.mode = mode;
    kp.expect = b->dExpect.p;
    kp.threshold = (float)p->threshold;
    for (int c = 0; c < kNumClasses; c++) {
        kp.jobBegin[c] = b->classBegin[c];
        kp.jobEnd[c] = b->classEnd[c];
    }
    HIP_TRY(hipStreamSynchronize(s));
    *out = b;
    return MP_OK;
}

static mp_status batch_run_impl(mp_batch *b, bool plain) {
    mp_engine *e = b->eng;
    HIP_TRY(hipSetDevice(e->device));
    hipStream_t s = e->stream;
    const int32_t nj = (int32_t)b->jobs.size();
    if (nj == 0) {
        b->kernelMs = 0.0;
        b->ran = true;
        return MP_OK;
    }
    mp_status st;
    float ms = 0.f;
    if (b->mode == MP_MODE_EXPECTATIONS && (st = reset_queue(b, reinterpret_cast<int32_t *>(b->dExpect.p), 20 * nj)))
        return st;
    const bool anyNarrow = b->classEnd[0] > b->classBegin[0] || b->classEnd[1] > b->classBegin[1] ||
                           b->classEnd[2] > b->classBegin[2];
    if (!plain) {
        if ((st = reset_queue(b, b->dQueue.p, kNumClasses + 1))) return st;
        HIP_TRY(hipEventRecord(e->ev0, s));
        if (anyNarrow && (st = launch(b, e->cus, kBlock, b->kp))) return st;
        if (b->wideEnd > b->wideBegin && (st = launch_wide(b, b->wideWaves, b->kp))) return st;
        HIP_TRY(hipEventRecord(e->ev1, s));
        HIP_TRY(hipEventSynchronize(e->ev1));
        HIP_TRY(hipEventElapsedTime(&ms, e->ev0, e->ev1));
    } else {
        // one launch (one wavefront) per job, in device order
        DevBuf<int32_t> q;
        const int nq = kNumClasses + 1;
        if ((st = q.alloc((size_t)nj * nq))) return st;
        if ((st = reset_queue(b, q.p, nj * nq))) return st;
        HIP_TRY(hipEventRecord(e->ev0, s));
        for (int32_t j = 0; j < nj; j++) {
            KParams kp = b->kp;
            for (int c = 0; c < kNumClasses; c++) kp.jobBegin[c] = kp.jobEnd[c] = j;
            kp.wideBegin = kp.wideEnd = j;
            const int c = b->jobs[j].cls;
            kp.queue = q.p + (size_t)j * nq;
            if (c < 3) {
                kp.jobEnd[c] = j + 1;
                if ((st = launch(b, 1, 64, kp))) return st;
            } else {
                kp.wideEnd = j + 1;
                if ((st = launch_wide(b, 1, kp))) return st;
            }
        }
        HIP_TRY(hipEventRecord(e->ev1, s));
        HIP_TRY(hipEventSynchronize(e->ev1));
        HIP_TRY(hipEventElapsedTime(&ms, e->ev0, e->ev1));
    }
    b->kernelMs = ms;
    // Overflowing output lists: size them exactly and run the batch again.
    b->hCounts.resize((size_t)nj * 3);
    HIP_TRY(hipMemcpy(b->hCounts.data(), b->dCounts.p, sizeof(int32_t) * 3 * nj, hipMemcpyDeviceToHost));
    bool overflow = false;
    for (int32_t j = 0; j < nj && !overflow; j++)
        for (int l = 0; l < 3; l++)
            if (b->hCounts[3 * j + l] > b->jobs[j].cap[l]) overflow = true;
    if (overflow) {
        std::vector<DevJob> dj(nj);
        HIP_TRY(hipMemcpy(dj.data(), b->dJobs.p, sizeof(DevJob) * nj, hipMemcpyDeviceToHost));
        int64_t off = 0;
        for (int32_t j = 0; j < nj; j++) {
            for (int l = 0; l < 3; l++) {
                b->jobs[j].cap[l] = std::max(b->jobs[j].cap[l], b->hCounts[3 * j + l]);
                dj[j].cap[l] = b->jobs[j].cap[l];
            }
            b->jobs[j].outOff = off;
            dj[j].outOff = off;
            off += (int64_t)dj[j].cap[0] + dj[j].cap[1] + dj[j].cap[2];
        }
        HIP_TRY(hipMemcpy(b->dJobs.p, dj.data(), sizeof(DevJob) * nj, hipMemcpyHostToDevice));
        if ((st = b->dOut.alloc(std::max<int64_t>(off, 1)))) return st;
        b->kp.out = b->dOut.p;
        b->outTotal = off;
        return batch_run_impl(b, plain);
    }
    b->ran = true;
    return MP_OK;
}

extern "C" mp_status mp_batch_run(mp_batch *b) {
    if (!b) {
        mp::set_error("mp_batch_run: null batch");
        return MP_EINVAL;
    }
    return batch_run_impl(b, false);
}

extern "C" mp_status mp_batch_run_plain(mp_batch *b) {
    if (!b) {
        mp::set_error("mp_batch_run_plain: null batch");
        return MP_EINVAL;
    }
    return batch_run_impl(b, true);
}

extern "C" double mp_batch_kernel_ms(const mp_batch *b) { return b ? b->kernelMs : 0.0; }
extern "C" int64_t mp_batch_cells(const mp_batch *b) { return b ? b->cells : 0; }
extern "C" int64_t mp_batch_npairs(const mp_batch *b) { return b ? b->npairs : 0; }

extern "C" mp_status mp_batch_fetch_counts(mp_batch *b, int64_t *counts, double *logProb) {
    if (!b || !counts) {
        mp::set_error("mp_batch_fetch_counts: null argument");
        return MP_EINVAL;
    }
    if (!b->ran) {
        mp::set_error("mp_batch_fetch_counts: batch has not run");
        return MP_EINVAL;
    }
    const int32_t nj = (int32_t)b->jobs.size();
    std::vector<double> lp(std::max(nj, 1));
    if (nj) {
        HIP_TRY(hipSetDevice(b->eng->device));
        HIP_TRY(hipMemcpy(lp.data(), b->dLogProb.p, sizeof(double) * nj, hipMemcpyDeviceToHost));
    }
    for (int64_t i = 0; i < b->npairs; i++) {
        int64_t c[3] = {0, 0, 0};
        double l = 0.0;
        for (int32_t s = 0; s < b->pairNJobs[i]; s++) {
            int32_t q = b->order[b->pairFirstJob[i] + s];
            if (q < 0) continue;
            for (int k = 0; k < 3; k++) c[k] += b->hCounts[3 * q + k];
            l += lp[q];
        }
        for (int k = 0; k < 3; k++) counts[3 * i + k] = c[k];
        if (logProb) logProb[i] = l;
    }
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
    const int32_t nj = (int32_t)b->jobs.size();
    if (nj == 0) return MP_OK;
    std::vector<int4> host((size_t)b->outTotal);
    HIP_TRY(hipSetDevice(b->eng->device));
    HIP_TRY(hipMemcpy(host.data(), b->dOut.p, sizeof(int4) * b->outTotal, hipMemcpyDeviceToHost));
    int64_t w = 0;
    for (int64_t i = 0; i < b->npairs; i++) {
        const int64_t crop = b->pairCrop[i];
        for (int32_t s = 0; s < b->pairNJobs[i]; s++) {
            const int32_t q = b->order[b->pairFirstJob[i] + s];
            if (q < 0) continue;
            const HostJob &hj = b->jobs[q];
            int64_t base = hj.outOff;
            for (int l = 0; l < list; l++) base += hj.cap[l];
            const int32_t n = b->hCounts[3 * q + list];
            // the reference appends each split's list reversed (pairwiseAligner.c:1072-1074)
            for (int32_t t = n - 1; t >= 0; t--) {
                const int4 v = host[base + t];
                mp_aligned_pair &o = out[w++];
                o.weight = (int64_t)(uint32_t)v.x | ((int64_t)v.y << 32);
                o.x = (int32_t)(v.z + hj.xOff + crop);
                o.y = (int32_t)(v.w + hj.yOff);
            }
        }
    }
    return MP_OK;
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
        std::vector<mp_aligned_pair> tmp(counts[l]);
        if ((st = mp_batch_fetch_list(b, l, tmp.data()))) {
            mp_batch_free(b);
            return st;
        }
        std::memcpy(dst[l], tmp.data(), sizeof(mp_aligned_pair) * std::min(counts[l], caps[l]));
    }
    if (logProb) *logProb = lp;
    mp_batch_free(b);
    return MP_OK;
}
