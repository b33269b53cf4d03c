// mp_wide.hpp -- bands wider than 64 cells on some anti-diagonal (anchor-free
// calls, large band half-widths), gfx950.
//
// One job per wavefront; cell i = lane + 64c for c < C (up to 2048 cells).
// A workgroup holds as many wavefronts as its LDS allows (each running jobs of
// its own, synchronised within the wave only), sharing one copy of the
// emission table.  The (X, Y, M) outputs of the last three anti-diagonals live
// in the wave's LDS slice and are read by band coordinate; one anti-diagonal
// per step.  Scaling is lazy: the
// exponent chosen at a grid-cell start comes from the previous anti-diagonal's
// maximum, and the two-back match input is corrected by the same exponent on
// the next step.  The forward sweep of a window runs to tb, then the backward
// sweep runs down with posterior records (backward-side order: anti-diagonals
// descending, cells ascending).  Same windows, modes and outputs as the
// narrow kernels (pairwiseAligner.c:687-808, 813-865, 1151-1164), window jobs
// included: a restarted job's forward starts at fstart from a flat vector and
// leaves its chain value (F . endV at the previous window's tb).
#pragma once

#include "mp_device.hpp"

namespace mpk {

struct WideCtx {
    const KParams &kp;
    float *xb;  // [3 slots][3 states][W]
    int W;
    int64_t bandOff;
    __device__ float *slot(int d, int st) const { return xb + ((size_t)(((d % 3) + 3) % 3) * 3 + st) * W; }
    __device__ void band(int d, int &L, int &R) const {
        const int b = kp.band[bandOff + d];
        L = 2 * band_xL(b) - d;
        R = L + 2 * band_wm1_wide(b);
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

// LDS of a workgroup: the emission table, then per wave three anti-diagonals
// of (X, Y, M) outputs for C columns.
constexpr int kWideLdsLimit = 160 * 1024;
__host__ __device__ constexpr int wide_wave_lds_bytes(int c) { return 9 * 64 * c * 4; }
constexpr int kWideMaxGroupWaves = 8;  // __launch_bounds__(512): the kernel keeps its registers unspilled
__host__ __device__ constexpr int wide_group_waves(int c) {
    return (kWideLdsLimit - kTabReach) / wide_wave_lds_bytes(c) > kWideMaxGroupWaves
               ? kWideMaxGroupWaves
               : ((kWideLdsLimit - kTabReach) / wide_wave_lds_bytes(c) < 1
                      ? 1
                      : (kWideLdsLimit - kTabReach) / wide_wave_lds_bytes(c));
}
__host__ __device__ constexpr int wide_lds_bytes(int c, int waves) { return kTabReach + waves * wide_wave_lds_bytes(c); }
static_assert(wide_lds_bytes(kWideMaxC, wide_group_waves(kWideMaxC)) <= kWideLdsLimit, "wide kernel LDS");

// Orders one wave's LDS writes before its following reads (no workgroup barrier:
// the other waves of the workgroup run other jobs).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(64 * kWideMaxGroupWaves) void pairhmm_wide_kernel(KParams kp) {
    extern __shared__ float ldsw[];
    float *tabL = ldsw;
    for (int i = threadIdx.x; i < kTabFloats; i += blockDim.x) tabL[i] = kp.tables[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int gw = blockIdx.x * (blockDim.x >> 6) + wv;  // this wave's ring slice
    if (gw >= kp.wideWaves) return;
    const int C = kp.wideC, W = 64 * C;
    float *xbuf = ldsw + kTabReach / 4 + (size_t)wv * 9 * W;
    // forward values: a row of the band's cells per anti-diagonal, rows placed one after the
    // other and wrapping to the start of the wave's ring; rowOf[d % R] = the row's first cell
    float4 *ring = reinterpret_cast<float4 *>(kp.ringWide) + (size_t)gw * kp.ringWideCells;
    int32_t *rowOf = kp.ringWideIdx + (size_t)gw * kp.ringWideLen;
    const int R = kp.ringWideLen;
    const int64_t ringCap = kp.ringWideCells;
    const uint64_t below = (1ull << lane) - 1ull;
    const bool expMode = kp.mode == MP_MODE_EXPECTATIONS, fwdOnly = kp.mode == MP_MODE_FORWARD_ONLY;
    const bool postMode = kp.mode <= MP_MODE_MATCH_ONLY;
    const bool withIndels = kp.withIndels != 0;
    const int stopOff = expMode ? 0 : 1;
    const float thr = kp.threshold;
    const float T[9] = {kp.T[0], kp.T[1], kp.T[2], kp.T[3], kp.T[4], kp.T[5], kp.T[6], kp.T[7], kp.T[8]};

    for (;;) {
        int jn = 0;
        if (lane == 0) jn = atomicAdd(&kp.queue[3], 1);
        jn = __shfl(jn, 0, 64);
        const int pos = kp.wideBegin + jn;
        if (pos >= kp.wideEnd) break;
        const int job = kp.jobMap ? kp.jobMap[pos] : pos;
        const DevJob jb = kp.jobs[job];
        const int D = jb.lX + jb.lY;
        const bool rgL = (jb.flags >> 1) & 1, rgR = (jb.flags >> 2) & 1;
        WideCtx cx{kp, xbuf, W, jb.bandOff};
        const int64_t outBase = jb.outOff + jb.cap[0];
        const int cap = jb.cap[1];
        int cnt = 0;
        double likeAcc = 0.0;
        double eacc[9];
        for (int i = 0; i < 9; i++) eacc[i] = 0.0;
        int S = 0;          // forward exponent
        float lastMax = 0.75f;
        int pendM = 0;      // exponent correction of the two-back match input
        int64_t nextRow = 0;     // ring cell of the next forward row
        int dStart = jb.fstart;  // first forward anti-diagonal of this window
        const bool restart = (jb.flags & kJobRestart) != 0;
        // a restarted window job's chain anti-diagonal: the previous window's tb
        const int capD = restart ? kp.win[jb.winOff].lo + kp.tbd + 1 : -1;
        for (int i = lane; i < 9 * W; i += 64) xbuf[i] = 0.f;
        wave_sync();
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
                int64_t row = nextRow;
                if (!fwdOnly) {
                    const int wd = (R0 - L0) / 2 + 1;
                    if (row + wd > ringCap) row = 0;
                    nextRow = row + wd;
                    rowOf[d % R] = (int32_t)row;  // every lane stores the same offset
                }
                float mx = 0.f, uc = 0.f;
                for (int c = 0; c < C; c++) {
                    const int i = lane + 64 * c, xmy = L0 + 2 * i;
                    const bool act = xmy <= R0;
                    const int x = (d + xmy) >> 1, y = (d - xmy) >> 1;
                    float p0 = 0.f, p1 = 0.f, p2 = 0.f;
                    if (act) {
                        const Emis e = emissions(tabL, kp.sym[jb.symX + x - 1], kp.sym[jb.symY + y - 1]);
                        if (d == 0) {
                            p0 = rgL ? kp.startV[1][0] : kp.startV[0][0];
                            p1 = rgL ? kp.startV[1][1] : kp.startV[0][1];
                            p2 = rgL ? kp.startV[1][2] : kp.startV[0][2];
                        } else if (restart && d == jb.fstart) {
                            p0 = p1 = p2 = 1.f;  // flat restart vector
                        } else {
                            const int sX = xmy - 1, sY = xmy + 1;
                            const float inX = (sX >= L1 && sX <= R1) ? cx.slot(d - 1, 1)[(sX - L1) >> 1] : 0.f;
                            const float inY = (sY >= L1 && sY <= R1) ? cx.slot(d - 1, 2)[(sY - L1) >> 1] : 0.f;
                            const float inM = (xmy >= L2 && xmy <= R2) ? cx.slot(d - 2, 0)[(xmy - L2) >> 1] : 0.f;
                            p0 = e.eM * inM * sclM;
                            p1 = e.eX * inX * scl;
                            p2 = e.eY * inY * scl;
                        }
                    }
                    mx = fmaxf(mx, fmaxf(p0, fmaxf(p1, p2)));
                    cx.slot(d, 0)[i] = fmaf(T[6], p2, fmaf(T[3], p1, T[0] * p0));
                    cx.slot(d, 1)[i] = fmaf(T[7], p2, fmaf(T[4], p1, T[1] * p0));
                    cx.slot(d, 2)[i] = fmaf(T[8], p2, fmaf(T[5], p1, T[2] * p0));
                    if (!fwdOnly && act) ring[row + i] = make_float4(p0, p1, p2, __int_as_float(S + e2));
                    if (d == capD) uc += p0 * kp.endV[0][0] + p1 * kp.endV[0][1] + p2 * kp.endV[0][2];
                    if (d == wn.tb) {
                        const float e0 = rgR && atEnd ? kp.endV[1][0] : kp.endV[0][0];
                        const float e1 = rgR && atEnd ? kp.endV[1][1] : kp.endV[0][1];
                        const float e2v = rgR && atEnd ? kp.endV[1][2] : kp.endV[0][2];
                        lastMax = p0 * e0 + p1 * e1 + p2 * e2v + (c == 0 ? 0.f : lastMax);  // partial total
                    }
                }
                if (d == capD) {
                    const float u = wave_sum(uc);
                    if (lane == 0) kp.chain[job] = make_int2(__float_as_int(u), S + e2);
                }
                S += e2;
                pendM = e2;
                if (d != wn.tb) lastMax = mx;
                wave_sync();
            }
            const float t = wave_sum(lastMax);
            const int SFtb = S;
            const float totalInv = t > 0.f ? 1.f / t : 0.f;
            const double lt = log((double)t) + (double)SFtb * 0.69314718055994530942;
            if (win + 1 == jb.nWin && lane == 0) kp.logTotal[job] = make_int2(__float_as_int(t), SFtb);
            likeAcc += (double)(wn.from - wn.lo) * lt;
            if (fwdOnly) break;
            // ---- backward from tb down to lo + stopOff ----
            for (int i = lane; i < 9 * W; i += 64) xbuf[i] = 0.f;
            wave_sync();
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
                const int64_t row = rowOf[d % R];
                float mx = 0.f;
                float ep[9];
                for (int q = 0; q < 9; q++) ep[q] = 0.f;
                for (int c = 0; c < C; c++) {
                    const int i = lane + 64 * c, xmy = L0 + 2 * i;
                    const bool act = xmy <= R0;
                    const int x = (d + xmy) >> 1, y = (d - xmy) >> 1;
                    float q0 = 0.f, q1 = 0.f, q2 = 0.f, eX = 0.f, eY = 0.f, eM = 0.f, inX = 0.f, inY = 0.f, inM = 0.f;
                    if (act) {
                        const Emis e = emissions(tabL, kp.sym[jb.symX + x - 1], kp.sym[jb.symY + y - 1]);
                        eX = e.eX;
                        eY = e.eY;
                        eM = e.eM;
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
                            q0 = fmaf(T[2], inY, fmaf(T[1], inX, T[0] * inM)) * scl;
                            q1 = fmaf(T[5], inY, fmaf(T[4], inX, T[3] * inM)) * scl;
                            q2 = fmaf(T[8], inY, fmaf(T[7], inX, T[6] * inM)) * scl;
                        }
                    }
                    mx = fmaxf(mx, fmaxf(q0, fmaxf(q1, q2)));
                    const float4 f = act ? ring[row + i] : make_float4(0.f, 0.f, 0.f, 0.f);
                    if (expMode && act) {
                        const float sc = ldexpf(totalInv, __float_as_int(f.w) + SB - SFtb);
                        const bool okXY = d + 1 > wn.lo && d + 1 <= wn.from, okM = d + 2 > wn.lo && d + 2 <= wn.from;
                        const float tX = okXY ? inX : 0.f, tY = okXY ? inY : 0.f, tM = okM ? inM : 0.f;
                        const float a0 = f.x * sc, a1 = f.y * sc, a2 = f.z * sc;
                        ep[0] += a0 * T[0] * tM; ep[1] += a0 * T[1] * tX; ep[2] += a0 * T[2] * tY;
                        ep[3] += a1 * T[3] * tM; ep[4] += a1 * T[4] * tX; ep[5] += a1 * T[5] * tY;
                        ep[6] += a2 * T[6] * tM; ep[7] += a2 * T[7] * tX; ep[8] += a2 * T[8] * tY;
                        // expected emission counts (state posteriors by emission-table entry)
                        if (d > wn.lo && d <= wn.from) {
                            const float scq = ldexpf(totalInv, __float_as_int(f.w) + SB + e2 - SFtb);
                            const uint32_t sx = kp.sym[jb.symX + x - 1], sy = kp.sym[jb.symY + y - 1];
                            const uint32_t sxy = sx + sy;
                            const double pM = (double)(f.x * q0 * scq);
                            atomicAdd(&kp.emitCount[kTabE + (((sxy >> 16) & 0xffu) >> 2)], pM);
                            atomicAdd(&kp.emitCount[kTabRL + (min(sxy & 0xffffu, kRLLastByte) >> 2)], pM);
                            atomicAdd(&kp.emitCount[(sx >> 24) >> 2], (double)(f.y * q1 * scq));
                            atomicAdd(&kp.emitCount[(sy >> 24) >> 2], (double)(f.z * q2 * scq));
                        }
                    }
                    if (postMode && d <= wn.from) {
                        const float sc = ldexpf(totalInv, __float_as_int(f.w) + SB + e2 - SFtb);
                        const float pM = f.x * q0 * sc, pX = f.y * q1 * sc, pY = f.z * q2 * sc;
                        const bool emit = act && (pM >= thr || (withIndels && (pX >= thr || pY >= thr)));
                        const uint64_t bm = __ballot(emit);
                        if (emit) {
                            const int idx = cnt + __popcll(bm & below);
                            if (idx < cap)
                                kp.out[outBase + idx] = make_int4(__float_as_int(pM), __float_as_int(pX),
                                                                  __float_as_int(pY), (int)rec_code(d, i));
                        }
                        cnt += __popcll(bm);
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
                wave_sync();
            }
            if (lane == 0) {
                kp.winEnd[2 * (jb.winOff + win)] = 0;
                kp.winEnd[2 * (jb.winOff + win) + 1] = cnt;
            }
            // ---- resume the forward sweep after tb: rebuild its exchange state from the ring ----
            if (win + 1 < jb.nWin) {
                const int tb = wn.tb;
                int L0, R0, L1, R1;
                cx.band(tb, L0, R0);
                cx.band(tb - 1, L1, R1);
                const int64_t row0 = rowOf[tb % R], row1 = rowOf[(tb - 1) % R];
                for (int c = 0; c < C; c++) {
                    const int i = lane + 64 * c;
                    const bool a0 = L0 + 2 * i <= R0, a1 = L1 + 2 * i <= R1;
                    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
                    const float4 f = a0 ? ring[row0 + i] : z;
                    const float4 g = a1 ? ring[row1 + i] : z;
                    const float sg = ldexpf(1.f, __float_as_int(g.w) - __float_as_int(f.w));
                    cx.slot(tb, 0)[i] = a0 ? T[0] * f.x + T[3] * f.y + T[6] * f.z : 0.f;
                    cx.slot(tb, 1)[i] = a0 ? T[1] * f.x + T[4] * f.y + T[7] * f.z : 0.f;
                    cx.slot(tb, 2)[i] = a0 ? T[2] * f.x + T[5] * f.y + T[8] * f.z : 0.f;
                    cx.slot(tb - 1, 0)[i] = a1 ? (T[0] * g.x + T[3] * g.y + T[6] * g.z) * sg : 0.f;
                }
                wave_sync();
                dStart = tb + 1;
                pendM = 0;
                lastMax = 0.75f;
            }
        }
        if (lane == 0) {
            kp.counts[2 * job] = 0;
            kp.counts[2 * job + 1] = cnt;
        }
        if (expMode) {
            for (int q = 0; q < 9; q++) {
                const double v = wave_sum_d(eacc[q]);
                if (lane == 0) kp.expect[(size_t)job * 10 + q] = v;
            }
            if (lane == 0) kp.expect[(size_t)job * 10 + 9] = likeAcc;
        }
        wave_sync();
    }
}

}  // namespace mpk
