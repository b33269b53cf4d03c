// mp_wide5.hpp -- the five-state machine (M, short gap X, short gap Y, long
// gap X, long gap Y), every job, gfx950.  No counterpart in the reference.
//
// Same sweep structure as the 3-state wide kernel: one job per wavefront,
// cell i = lane + 64c (c < C), one anti-diagonal per step, the per-target
// "send" values of the last three anti-diagonals in LDS, read by band
// coordinate; forward to tb, backward down with posterior records
// (anti-diagonals descending, cells ascending); same windows and outputs.
//
// Scaling is per cell: each cell keeps its five values normalised to a
// maximum in [0.5, 1) with an int exponent of its own.  A cell aligns its
// neighbours' inputs to the largest neighbour exponent before the update.  A
// long gap extends with probability ~0.997 at the gap emission, so along one
// anti-diagonal the long-gap values at the band edges and the match values
// near the alignment path differ by more than the FP32 range once the band is
// a few hundred cells wide; one exponent per anti-diagonal (the 3-state
// kernels' scheme) would flush the path cells to zero.
//
// A gap state of either length emits like the 3-state gap of its side; the
// gap-X (gap-Y) posterior of a cell is the sum over its two gap-X (gap-Y)
// states.
#pragma once

#include "mp_device.hpp"

namespace mpk {

constexpr int kNS5 = 5;
constexpr int kWide5MaxC = 20;  // up to 64 * 20 = 1280 cells per anti-diagonal
// LDS: emission table, then send values [3 slots][5 targets][W] and the
// cells' exponents [3 slots][W].
constexpr int wide5_lds_bytes(int c) { return kTabReach + 3 * (kNS5 + 1) * 64 * c * 4; }
constexpr int kWide5LdsBytes = wide5_lds_bytes(kWide5MaxC);

// state kind: 0 match (input from (x-1, y-1)), 1 gap X (from (x-1, y)), 2 gap Y (from (x, y-1))
__host__ __device__ constexpr int kind5(int s) { return s == 0 ? 0 : ((s & 1) ? 1 : 2); }

struct Wide5Ctx {
    float *val;  // [3][5][W]
    int *ex;     // [3][W]
    int W;
    __device__ static int slotOf(int d) { return ((d % 3) + 3) % 3; }
    __device__ float *v(int d, int s) const { return val + (size_t)(slotOf(d) * kNS5 + s) * W; }
    __device__ int *e(int d) const { return ex + (size_t)slotOf(d) * W; }
};

__device__ __forceinline__ int wave_max_i(int v) {
    for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float wave_sum5(float v) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// (acc, accE) += (v, ve): a running sum of values with exponents of their own
__device__ __forceinline__ void acc_add(float &acc, int &accE, float v, int ve) {
    if (v == 0.f) return;
    if (acc == 0.f) {
        acc = v;
        accE = ve;
    } else if (ve > accE) {
        acc = ldexpf(acc, accE - ve) + v;
        accE = ve;
    } else {
        acc += ldexpf(v, ve - accE);
    }
}

// Exponent of a cell without mass (all five values 0).  It loses every
// maximum against a cell with mass, and differences with it stay far from
// int overflow while ldexpf of them is 0.
constexpr int kNoMass = -(1 << 29);

// Scale the five values (of exponent E) to a maximum in [0.5, 1); returns the
// new exponent, kNoMass when all values are 0.
__device__ __forceinline__ int renorm5(float *v, int E) {
    float m = v[0];
#pragma unroll
    for (int s = 1; s < kNS5; s++) m = fmaxf(m, v[s]);
    if (m == 0.f) return kNoMass;
    const int k = __builtin_amdgcn_frexp_expf(m);
    const float sc = ldexpf(1.f, -k);
#pragma unroll
    for (int s = 0; s < kNS5; s++) v[s] *= sc;
    return E + k;
}

__global__ __launch_bounds__(64) void pairhmm_wide5_kernel(KParams kp) {
    extern __shared__ float ldsw5[];
    float *tabL = ldsw5;
    const int lane = threadIdx.x;
    for (int i = lane; i < kTabFloats; i += 64) tabL[i] = kp.tables[i];
    __syncthreads();
    const int C = kp.wideC, W = 64 * C;
    Wide5Ctx cx{ldsw5 + kTabReach / 4, reinterpret_cast<int *>(ldsw5 + kTabReach / 4 + 3 * kNS5 * W), W};
    float4 *ring = reinterpret_cast<float4 *>(kp.ringWide) + (size_t)blockIdx.x * kp.ringWideLen * W * 2;
    const int R = kp.ringWideLen;
    auto ringAt = [&](int d, int k, int i) -> float4 & { return ring[((size_t)(d % R) * 2 + k) * W + i]; };
    const uint64_t below = (1ull << lane) - 1ull;
    const bool fwdOnly = kp.mode == MP_MODE_FORWARD_ONLY;
    const bool postMode = kp.mode <= MP_MODE_MATCH_ONLY;
    const bool withIndels = kp.withIndels != 0;
    const float thr = kp.threshold;
    float T[kNS5 * kNS5];
#pragma unroll
    for (int q = 0; q < kNS5 * kNS5; q++) T[q] = kp.T5[q];

    for (;;) {
        int jn = 0;
        if (lane == 0) jn = atomicAdd(&kp.queue[3], 1);
        jn = __shfl(jn, 0, 64);
        const int job = kp.wideBegin + jn;
        if (job >= kp.wideEnd) break;
        const DevJob jb = kp.jobs[job];
        const int D = jb.lX + jb.lY;
        const bool rgL = (jb.flags >> 1) & 1, rgR = (jb.flags >> 2) & 1;
        auto band = [&](int d, int &L, int &Rr) {
            const int b = kp.band[jb.bandOff + d];
            L = 2 * band_xL(b) - d;
            Rr = L + 2 * band_wm1_wide(b);
        };
        const int64_t outBase = jb.outOff + jb.cap[0];
        const int cap = jb.cap[1];
        int cnt = 0;
        int dStart = 0;
        for (int i = lane; i < 3 * (kNS5 + 1) * W; i += 64) cx.val[i] = 0.f;
        __syncthreads();
        for (int win = 0; win < jb.nWin; win++) {
            const DevWin wn = kp.win[jb.winOff + win];
            const bool atEnd = wn.tb == D;
            const int rEnd = rgR && atEnd ? 1 : 0;
            float acc = 0.f;  // this lane's part of the total at tb
            int accE = 0;
            // ---- forward from dStart to tb ----
            for (int d = dStart; d <= wn.tb; d++) {
                int L0, R0, L1 = 1, R1 = -1, L2 = 1, R2 = -1;
                band(d, L0, R0);
                if (d >= 1) band(d - 1, L1, R1);
                if (d >= 2) band(d - 2, L2, R2);
                for (int c = 0; c < C; c++) {
                    const int i = lane + 64 * c, xmy = L0 + 2 * i;
                    const bool act = xmy <= R0;
                    const int x = (d + xmy) >> 1, y = (d - xmy) >> 1;
                    float p[kNS5] = {0.f, 0.f, 0.f, 0.f, 0.f};
                    int E = kNoMass;
                    if (act) {
                        const Emis em = emissions(tabL, kp.sym[jb.symX + x - 1], kp.sym[jb.symY + y - 1]);
                        if (d == 0) {
#pragma unroll
                            for (int s = 0; s < kNS5; s++) p[s] = s == 0 ? (rgL ? 0.f : 1.f) : (rgL ? 1.f : 0.f);
                            E = renorm5(p, 0);
                        } else {
                            const int sX = xmy - 1, sY = xmy + 1;
                            const bool okX = sX >= L1 && sX <= R1, okY = sY >= L1 && sY <= R1;
                            const bool okM = xmy >= L2 && xmy <= R2;
                            const int jX = (sX - L1) >> 1, jY = (sY - L1) >> 1, jM = (xmy - L2) >> 1;
                            const int eXn = okX ? cx.e(d - 1)[jX] : kNoMass, eYn = okY ? cx.e(d - 1)[jY] : kNoMass;
                            const int eMn = okM ? cx.e(d - 2)[jM] : kNoMass;
                            const int Ein = max(eXn, max(eYn, eMn));
                            const float fX = okX ? em.eX * ldexpf(1.f, eXn - Ein) : 0.f;
                            const float fY = okY ? em.eY * ldexpf(1.f, eYn - Ein) : 0.f;
                            const float fM = okM ? em.eM * ldexpf(1.f, eMn - Ein) : 0.f;
                            p[0] = okM ? fM * cx.v(d - 2, 0)[jM] : 0.f;
                            p[1] = okX ? fX * cx.v(d - 1, 1)[jX] : 0.f;
                            p[3] = okX ? fX * cx.v(d - 1, 3)[jX] : 0.f;
                            p[2] = okY ? fY * cx.v(d - 1, 2)[jY] : 0.f;
                            p[4] = okY ? fY * cx.v(d - 1, 4)[jY] : 0.f;
                            E = renorm5(p, Ein);
                        }
                    }
#pragma unroll
                    for (int t = 0; t < kNS5; t++) {
                        float a = T[t] * p[0];
#pragma unroll
                        for (int s = 1; s < kNS5; s++) a = fmaf(T[s * kNS5 + t], p[s], a);
                        cx.v(d, t)[i] = a;
                    }
                    cx.e(d)[i] = E;
                    if (!fwdOnly) {
                        ringAt(d, 0, i) = make_float4(p[0], p[1], p[2], p[3]);
                        ringAt(d, 1, i) = make_float4(p[4], __int_as_float(E), 0.f, 0.f);
                    }
                    if (d == wn.tb && act) {
                        float tv = p[0] * kp.endV5[rEnd][0];
#pragma unroll
                        for (int s = 1; s < kNS5; s++) tv = fmaf(p[s], kp.endV5[rEnd][s], tv);
                        acc_add(acc, accE, tv, E);
                    }
                }
                __syncthreads();
            }
            // total at tb: t * 2^SFtb
            const int SFtb = wave_max_i(acc > 0.f ? accE : kNoMass);
            const float t = wave_sum5(acc > 0.f ? ldexpf(acc, accE - SFtb) : 0.f);
            const float totalInv = t > 0.f ? 1.f / t : 0.f;
            if (atEnd && lane == 0) kp.logTotal[job] = make_int2(__float_as_int(t), t > 0.f ? SFtb : 0);
            if (fwdOnly) break;
            // ---- backward from tb down to lo + 1 ----
            for (int d = wn.tb; d >= wn.lo + 1; d--) {
                int L0, R0, L1 = 1, R1 = -1, L2 = 1, R2 = -1;
                band(d, L0, R0);
                if (d + 1 <= wn.tb) band(d + 1, L1, R1);
                if (d + 2 <= wn.tb) band(d + 2, L2, R2);
                for (int c = 0; c < C; c++) {
                    const int i = lane + 64 * c, xmy = L0 + 2 * i;
                    const bool act = xmy <= R0;
                    const int x = (d + xmy) >> 1, y = (d - xmy) >> 1;
                    float q[kNS5] = {0.f, 0.f, 0.f, 0.f, 0.f};
                    float eX = 0.f, eY = 0.f, eM = 0.f;
                    int E = kNoMass;
                    if (act) {
                        const Emis em = emissions(tabL, kp.sym[jb.symX + x - 1], kp.sym[jb.symY + y - 1]);
                        eX = em.eX;
                        eY = em.eY;
                        eM = em.eM;
                        if (d == wn.tb) {
#pragma unroll
                            for (int s = 0; s < kNS5; s++) q[s] = kp.endV5[rEnd][s];
                            E = renorm5(q, 0);
                        } else {
                            const int sX = xmy + 1, sY = xmy - 1;
                            const bool okX = sX >= L1 && sX <= R1, okY = sY >= L1 && sY <= R1;
                            const bool okM = xmy >= L2 && xmy <= R2;
                            const int jX = (sX - L1) >> 1, jY = (sY - L1) >> 1, jM = (xmy - L2) >> 1;
                            const int eXn = okX ? cx.e(d + 1)[jX] : kNoMass, eYn = okY ? cx.e(d + 1)[jY] : kNoMass;
                            const int eMn = okM ? cx.e(d + 2)[jM] : kNoMass;
                            const int Ein = max(eXn, max(eYn, eMn));
                            const float fX = okX ? ldexpf(1.f, eXn - Ein) : 0.f;
                            const float fY = okY ? ldexpf(1.f, eYn - Ein) : 0.f;
                            const float fM = okM ? ldexpf(1.f, eMn - Ein) : 0.f;
                            float in[kNS5];
                            in[0] = okM ? fM * cx.v(d + 2, 0)[jM] : 0.f;
                            in[1] = okX ? fX * cx.v(d + 1, 1)[jX] : 0.f;
                            in[3] = okX ? fX * cx.v(d + 1, 3)[jX] : 0.f;
                            in[2] = okY ? fY * cx.v(d + 1, 2)[jY] : 0.f;
                            in[4] = okY ? fY * cx.v(d + 1, 4)[jY] : 0.f;
#pragma unroll
                            for (int s = 0; s < kNS5; s++) {
                                float a = T[s * kNS5] * in[0];
#pragma unroll
                                for (int tt = 1; tt < kNS5; tt++) a = fmaf(T[s * kNS5 + tt], in[tt], a);
                                q[s] = a;
                            }
                            E = renorm5(q, Ein);
                        }
                    }
                    if (postMode && d <= wn.from) {
                        float pM = 0.f, pX = 0.f, pY = 0.f;
                        if (act) {
                            const float4 f0 = ringAt(d, 0, i), f1 = ringAt(d, 1, i);
                            const float sc = ldexpf(totalInv, __float_as_int(f1.y) + E - SFtb);
                            pM = f0.x * q[0] * sc;
                            pX = fmaf(f0.w, q[3], f0.y * q[1]) * sc;
                            pY = fmaf(f1.x, q[4], f0.z * q[2]) * sc;
                        }
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
#pragma unroll
                    for (int s = 0; s < kNS5; s++) {
                        const int k = kind5(s);
                        cx.v(d, s)[i] = (k == 0 ? eM : (k == 1 ? eX : eY)) * q[s];
                    }
                    cx.e(d)[i] = E;
                }
                __syncthreads();
            }
            if (lane == 0) {
                kp.winEnd[2 * (jb.winOff + win)] = 0;
                kp.winEnd[2 * (jb.winOff + win) + 1] = cnt;
            }
            // ---- resume the forward sweep after tb: its send values at tb and
            // the match send values at tb - 1, rebuilt from the ring ----
            if (win + 1 < jb.nWin) {
                const int tb = wn.tb;
                int L0, R0, L1, R1;
                band(tb, L0, R0);
                band(tb - 1, L1, R1);
                for (int c = 0; c < C; c++) {
                    const int i = lane + 64 * c;
                    const bool a0 = L0 + 2 * i <= R0, a1 = L1 + 2 * i <= R1;
                    const float4 f0 = ringAt(tb, 0, i), f1 = ringAt(tb, 1, i);
                    const float4 g0 = ringAt(tb - 1, 0, i), g1 = ringAt(tb - 1, 1, i);
                    const float f[kNS5] = {f0.x, f0.y, f0.z, f0.w, f1.x};
                    const float g[kNS5] = {g0.x, g0.y, g0.z, g0.w, g1.x};
#pragma unroll
                    for (int tt = 0; tt < kNS5; tt++) {
                        float a = T[tt] * f[0];
#pragma unroll
                        for (int s = 1; s < kNS5; s++) a = fmaf(T[s * kNS5 + tt], f[s], a);
                        cx.v(tb, tt)[i] = a0 ? a : 0.f;
                    }
                    cx.e(tb)[i] = a0 ? __float_as_int(f1.y) : kNoMass;
                    float a = T[0] * g[0];
#pragma unroll
                    for (int s = 1; s < kNS5; s++) a = fmaf(T[s * kNS5], g[s], a);
                    cx.v(tb - 1, 0)[i] = a1 ? a : 0.f;
                    cx.e(tb - 1)[i] = a1 ? __float_as_int(g1.y) : kNoMass;
                }
                __syncthreads();
                dStart = tb + 1;
            }
        }
        if (lane == 0) {
            kp.counts[2 * job] = 0;
            kp.counts[2 * job + 1] = cnt;
        }
        __syncthreads();
    }
}

}  // namespace mpk
