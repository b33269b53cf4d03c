// mp_wide5.hpp -- the five-state machine (M, short gap X, short gap Y, long
// gap X, long gap Y), every job, in FP32 or FP64, gfx950.  No counterpart in
// the reference.
//
// Same sweep structure as the 3-state wide kernel, with one job per
// workgroup of up to 8 wavefronts: cell i = lane + 64c (c < C), wavefront w
// takes the columns c = w, w + NW, ...; one anti-diagonal per step (one
// workgroup barrier), neighbours read from LDS by band coordinate; forward to
// tb, backward down with posterior records; same windows and outputs.  The
// records of one anti-diagonal are contiguous in a window's record stream (a
// shared LDS counter hands out slots), cells in any order inside it: the list
// assembly ranks them by cell.
//
// Scaling is per cell: each cell keeps its five values normalised to a
// maximum in [0.5, 1) with an int exponent of its own, and aligns its
// neighbours' inputs to the largest neighbour exponent before the update.  A
// long gap extends with probability ~0.997 at the gap emission, so along one
// anti-diagonal the long-gap values at the band edges and the match values
// near the alignment path differ by more than the FP32 range once the band is
// a few hundred cells wide; one exponent per anti-diagonal (the 3-state
// kernels' scheme) would flush the path cells to zero.
//
// LDS: the emission table (only in-band cells form emission offsets, and the
// table's clamps keep them inside it), then per cell the "send" values of the
// last anti-diagonals -- the match target of three ([3][W]: the match input
// comes from two back), the four gap targets of two ([2][4][W]) -- and the
// cell exponents of three ([3][W]).
//
// A gap state of either length emits like the 3-state gap of its side; the
// gap-X (gap-Y) posterior of a cell is the sum over its two gap-X (gap-Y)
// states.
#pragma once

#include "mp_device.hpp"

namespace mpk {

constexpr int kNS5 = 5;
constexpr int kWide5TabBytes = kTabFloats * 4;  // 41904, a multiple of 16
constexpr int kWide5MaxWaves = 8;
constexpr int kWide5CtlBytes = kWide5MaxWaves * 16 + 16;  // per-wave reduction slots, job and record counters
template <class Real>
constexpr int wide5_lds_bytes(int c) {
    return kWide5TabBytes + 64 * c * (11 * (int)sizeof(Real) + 3 * 4) + kWide5CtlBytes;
}
template <class Real>
constexpr int wide5_max_c() {  // columns of 64 cells that fit 160 KB of LDS
    return (160 * 1024 - kWide5TabBytes - kWide5CtlBytes) / (64 * (11 * (int)sizeof(Real) + 3 * 4));
}
// wavefronts per workgroup for a batch whose widest band needs c columns
__host__ __device__ constexpr int wide5_waves(int c) { return c < kWide5MaxWaves ? c : kWide5MaxWaves; }
constexpr int kWide5MaxC32 = wide5_max_c<float>() < 32 ? wide5_max_c<float>() : 32;  // 2048 cells
constexpr int kWide5MaxC64 = wide5_max_c<double>();                                  // 19: 1216 cells

// state kind: 0 match (input from (x-1, y-1)), 1 gap X (from (x-1, y)), 2 gap Y (from (x, y-1))
__host__ __device__ constexpr int kind5(int s) { return s == 0 ? 0 : ((s & 1) ? 1 : 2); }

// Exponent of a cell without mass (all five values 0).  It loses every
// maximum against a cell with mass, differences with it stay far from int
// overflow, and ldexp of them is 0.
constexpr int kNoMass = -(1 << 29);

__device__ __forceinline__ int frexp_exp(float v) { return __builtin_amdgcn_frexp_expf(v); }
__device__ __forceinline__ int frexp_exp(double v) { return __builtin_amdgcn_frexp_exp(v); }
__device__ __forceinline__ float ldexp_r(float v, int k) { return ldexpf(v, k); }
__device__ __forceinline__ double ldexp_r(double v, int k) { return ldexp(v, k); }
__device__ __forceinline__ float fma_r(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double fma_r(double a, double b, double c) { return fma(a, b, c); }
// an int exponent kept in a ring row of the value type (its bit pattern, not its value)
__device__ __forceinline__ float exp_slot(int e, float) { return __int_as_float(e); }
__device__ __forceinline__ double exp_slot(int e, double) { return __longlong_as_double((long long)e); }
__device__ __forceinline__ int slot_exp(float v) { return __float_as_int(v); }
__device__ __forceinline__ int slot_exp(double v) { return (int)__double_as_longlong(v); }

__device__ __forceinline__ int wave_max_i(int v) {
    for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}
template <class Real>
__device__ __forceinline__ Real wave_sum_r(Real v) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// (acc, accE) += (v, ve): a running sum of values with exponents of their own
template <class Real>
__device__ __forceinline__ void acc_add(Real &acc, int &accE, Real v, int ve) {
    if (v == Real(0)) return;
    if (acc == Real(0)) {
        acc = v;
        accE = ve;
    } else if (ve > accE) {
        acc = ldexp_r(acc, accE - ve) + v;
        accE = ve;
    } else {
        acc += ldexp_r(v, ve - accE);
    }
}

// Scale the five values (of exponent E) to a maximum in [0.5, 1); returns the
// new exponent, kNoMass when all values are 0.
template <class Real>
__device__ __forceinline__ int renorm5(Real *v, int E) {
    Real m = v[0];
#pragma unroll
    for (int s = 1; s < kNS5; s++) m = v[s] > m ? v[s] : m;
    if (m == Real(0)) return kNoMass;
    const int k = frexp_exp(m);
    const Real sc = ldexp_r(Real(1), -k);
#pragma unroll
    for (int s = 0; s < kNS5; s++) v[s] *= sc;
    return E + k;
}

template <class Real>
struct Wide5Lds {
    Real *sendM;  // [3][W]    match target
    Real *sendG;  // [2][4][W] targets short X, short Y, long X, long Y
    int *ex;      // [3][W]
    int W;
    __device__ static int m3(int d) { return ((d % 3) + 3) % 3; }
    __device__ Real *m(int d) const { return sendM + (size_t)m3(d) * W; }
    __device__ Real *g(int d, int s) const { return sendG + (size_t)((d & 1) * 4 + s - 1) * W; }  // s = 1..4
    __device__ int *e(int d) const { return ex + (size_t)m3(d) * W; }
};

template <class Real>
__global__ __launch_bounds__(64 * kWide5MaxWaves) void pairhmm_wide5_kernel(KParams kp) {
    extern __shared__ float ldsw5[];
    float *tabL = ldsw5;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, NW = blockDim.x >> 6;
    for (int i = tid; i < kTabFloats; i += blockDim.x) tabL[i] = kp.tables[i];
    const int C = kp.wideC, W = 64 * C;
    Real *arr = reinterpret_cast<Real *>(reinterpret_cast<char *>(ldsw5) + kWide5TabBytes);
    const Wide5Lds<Real> lx{arr, arr + 3 * W, reinterpret_cast<int *>(arr + 11 * W), W};
    // control: per-wave totals (value, exponent), the job index, the record counter
    Real *redT = reinterpret_cast<Real *>(lx.ex + 3 * W);
    int *redE = reinterpret_cast<int *>(redT + kWide5MaxWaves);
    int *ctl = redE + kWide5MaxWaves;
    // ring: per anti-diagonal six rows of W entries (five values, the exponent)
    Real *ring = reinterpret_cast<Real *>(kp.ringWide) + (size_t)blockIdx.x * kp.ringWideLen * W * 6;
    const int R = kp.ringWideLen;
    auto ringRow = [&](int d, int s) { return ring + ((size_t)(d % R) * 6 + s) * W; };
    const uint64_t below = (1ull << lane) - 1ull;
    const bool fwdOnly = kp.mode == MP_MODE_FORWARD_ONLY;
    const bool postMode = kp.mode <= MP_MODE_MATCH_ONLY;
    const bool withIndels = kp.withIndels != 0;
    const float thr = kp.threshold;
    Real T[kNS5 * kNS5];
#pragma unroll
    for (int q = 0; q < kNS5 * kNS5; q++) T[q] = (Real)kp.T5[q];
    __syncthreads();

    for (;;) {
        if (tid == 0) {
            ctl[0] = atomicAdd(&kp.queue[3], 1);
            ctl[1] = 0;
        }
        __syncthreads();
        const int pos = kp.wideBegin + ctl[0];
        if (pos >= kp.wideEnd) break;
        const int job = kp.jobMap ? kp.jobMap[pos] : pos;
        const DevJob jb = kp.jobs[job];
        const int D = jb.lX + jb.lY;
        const int rStart = (jb.flags >> 1) & 1, rgR = (jb.flags >> 2) & 1;
        auto band = [&](int d, int &L, int &Rr) {
            const int b = kp.band[jb.bandOff + d];
            L = 2 * band_xL(b) - d;
            Rr = L + 2 * band_wm1_wide(b);
        };
        const int64_t outBase = jb.outOff + jb.cap[0];
        const int cap = jb.cap[1];
        int dStart = 0;
        for (int i = tid; i < 11 * W; i += blockDim.x) lx.sendM[i] = Real(0);
        for (int i = tid; i < 3 * W; i += blockDim.x) lx.ex[i] = kNoMass;
        __syncthreads();
        for (int win = 0; win < jb.nWin; win++) {
            const DevWin wn = kp.win[jb.winOff + win];
            const bool atEnd = wn.tb == D;
            const int rEnd = rgR && atEnd ? 1 : 0;
            Real acc = Real(0);  // this lane's part of the total at tb
            int accE = 0;
            // ---- forward from dStart to tb ----
            for (int d = dStart; d <= wn.tb; d++) {
                int L0, R0, L1 = 1, R1 = -1, L2 = 1, R2 = -1;
                band(d, L0, R0);
                if (d >= 1) band(d - 1, L1, R1);
                if (d >= 2) band(d - 2, L2, R2);
                for (int c = wave; c < C; c += NW) {
                    const int i = lane + 64 * c, xmy = L0 + 2 * i;
                    const bool act = xmy <= R0;
                    const int x = (d + xmy) >> 1, y = (d - xmy) >> 1;
                    Real p[kNS5] = {Real(0), Real(0), Real(0), Real(0), Real(0)};
                    int E = kNoMass;
                    if (act) {
                        if (d == 0) {
#pragma unroll
                            for (int s = 0; s < kNS5; s++) p[s] = (Real)kp.startV5[rStart][s];
                            E = renorm5(p, 0);
                        } else {
                            const Emis em = emissions(tabL, kp.sym[jb.symX + x - 1], kp.sym[jb.symY + y - 1]);
                            const int sX = xmy - 1, sY = xmy + 1;
                            const bool okX = sX >= L1 && sX <= R1, okY = sY >= L1 && sY <= R1;
                            const bool okM = xmy >= L2 && xmy <= R2;
                            const int jX = (sX - L1) >> 1, jY = (sY - L1) >> 1, jM = (xmy - L2) >> 1;
                            const int eXn = okX ? lx.e(d - 1)[jX] : kNoMass, eYn = okY ? lx.e(d - 1)[jY] : kNoMass;
                            const int eMn = okM ? lx.e(d - 2)[jM] : kNoMass;
                            const int Ein = max(eXn, max(eYn, eMn));
                            const Real fX = okX ? (Real)em.eX * ldexp_r(Real(1), eXn - Ein) : Real(0);
                            const Real fY = okY ? (Real)em.eY * ldexp_r(Real(1), eYn - Ein) : Real(0);
                            const Real fM = okM ? (Real)em.eM * ldexp_r(Real(1), eMn - Ein) : Real(0);
                            p[0] = okM ? fM * lx.m(d - 2)[jM] : Real(0);
                            p[1] = okX ? fX * lx.g(d - 1, 1)[jX] : Real(0);
                            p[3] = okX ? fX * lx.g(d - 1, 3)[jX] : Real(0);
                            p[2] = okY ? fY * lx.g(d - 1, 2)[jY] : Real(0);
                            p[4] = okY ? fY * lx.g(d - 1, 4)[jY] : Real(0);
                            E = renorm5(p, Ein);
                        }
                    }
#pragma unroll
                    for (int t = 0; t < kNS5; t++) {
                        Real a = T[t] * p[0];
#pragma unroll
                        for (int s = 1; s < kNS5; s++) a = fma_r(T[s * kNS5 + t], p[s], a);
                        if (t == 0) lx.m(d)[i] = a;
                        else lx.g(d, t)[i] = a;
                    }
                    lx.e(d)[i] = E;
                    if (!fwdOnly) {
#pragma unroll
                        for (int s = 0; s < kNS5; s++) ringRow(d, s)[i] = p[s];
                        ringRow(d, 5)[i] = exp_slot(E, Real(0));
                    }
                    if (d == wn.tb && act) {
                        Real tv = p[0] * (Real)kp.endV5[rEnd][0];
#pragma unroll
                        for (int s = 1; s < kNS5; s++) tv = fma_r(p[s], (Real)kp.endV5[rEnd][s], tv);
                        acc_add(acc, accE, tv, E);
                    }
                }
                __syncthreads();
            }
            // total at tb: t * 2^SFtb, over the wave, then over the waves in wave order
            {
                const int e = wave_max_i(acc > Real(0) ? accE : kNoMass);
                const Real tw = wave_sum_r(acc > Real(0) ? ldexp_r(acc, accE - e) : Real(0));
                if (lane == 0) {
                    redT[wave] = tw;
                    redE[wave] = e;
                }
            }
            __syncthreads();
            int SFtb = kNoMass;
            for (int w = 0; w < NW; w++)
                if (redT[w] > Real(0)) SFtb = max(SFtb, redE[w]);
            Real t = Real(0);
            for (int w = 0; w < NW; w++)
                if (redT[w] > Real(0)) t += ldexp_r(redT[w], redE[w] - SFtb);
            const Real totalInv = t > Real(0) ? Real(1) / t : Real(0);
            if (atEnd && tid == 0) {
                // (t, S) with t in [0.5, 1) as FP32: the host's log t + S ln 2 loses < 1e-7 to the rounding of t
                const int k = t > Real(0) ? frexp_exp(t) : 0;
                kp.logTotal[job] = make_int2(__float_as_int((float)ldexp_r(t, -k)), t > Real(0) ? SFtb + k : 0);
            }
            if (fwdOnly) break;
            // ---- backward from tb down to lo + 1 ----
            for (int d = wn.tb; d >= wn.lo + 1; d--) {
                int L0, R0, L1 = 1, R1 = -1, L2 = 1, R2 = -1;
                band(d, L0, R0);
                if (d + 1 <= wn.tb) band(d + 1, L1, R1);
                if (d + 2 <= wn.tb) band(d + 2, L2, R2);
                for (int c = wave; c < C; c += NW) {
                    const int i = lane + 64 * c, xmy = L0 + 2 * i;
                    const bool act = xmy <= R0;
                    const int x = (d + xmy) >> 1, y = (d - xmy) >> 1;
                    Real q[kNS5] = {Real(0), Real(0), Real(0), Real(0), Real(0)};
                    Real eX = Real(0), eY = Real(0), eM = Real(0);
                    int E = kNoMass;
                    if (act) {
                        const Emis em = emissions(tabL, kp.sym[jb.symX + x - 1], kp.sym[jb.symY + y - 1]);
                        eX = (Real)em.eX;
                        eY = (Real)em.eY;
                        eM = (Real)em.eM;
                        if (d == wn.tb) {
#pragma unroll
                            for (int s = 0; s < kNS5; s++) q[s] = (Real)kp.endV5[rEnd][s];
                            E = renorm5(q, 0);
                        } else {
                            const int sX = xmy + 1, sY = xmy - 1;
                            const bool okX = sX >= L1 && sX <= R1, okY = sY >= L1 && sY <= R1;
                            const bool okM = xmy >= L2 && xmy <= R2;
                            const int jX = (sX - L1) >> 1, jY = (sY - L1) >> 1, jM = (xmy - L2) >> 1;
                            const int eXn = okX ? lx.e(d + 1)[jX] : kNoMass, eYn = okY ? lx.e(d + 1)[jY] : kNoMass;
                            const int eMn = okM ? lx.e(d + 2)[jM] : kNoMass;
                            const int Ein = max(eXn, max(eYn, eMn));
                            const Real fX = okX ? ldexp_r(Real(1), eXn - Ein) : Real(0);
                            const Real fY = okY ? ldexp_r(Real(1), eYn - Ein) : Real(0);
                            const Real fM = okM ? ldexp_r(Real(1), eMn - Ein) : Real(0);
                            Real in[kNS5];
                            in[0] = okM ? fM * lx.m(d + 2)[jM] : Real(0);
                            in[1] = okX ? fX * lx.g(d + 1, 1)[jX] : Real(0);
                            in[3] = okX ? fX * lx.g(d + 1, 3)[jX] : Real(0);
                            in[2] = okY ? fY * lx.g(d + 1, 2)[jY] : Real(0);
                            in[4] = okY ? fY * lx.g(d + 1, 4)[jY] : Real(0);
#pragma unroll
                            for (int s = 0; s < kNS5; s++) {
                                Real a = T[s * kNS5] * in[0];
#pragma unroll
                                for (int tt = 1; tt < kNS5; tt++) a = fma_r(T[s * kNS5 + tt], in[tt], a);
                                q[s] = a;
                            }
                            E = renorm5(q, Ein);
                        }
                    }
                    if (postMode && d <= wn.from) {
                        float pM = 0.f, pX = 0.f, pY = 0.f;
                        if (act) {
                            Real f[kNS5];
#pragma unroll
                            for (int s = 0; s < kNS5; s++) f[s] = ringRow(d, s)[i];
                            const Real sc = ldexp_r(totalInv, slot_exp(ringRow(d, 5)[i]) + E - SFtb);
                            pM = (float)(f[0] * q[0] * sc);
                            pX = (float)(fma_r(f[3], q[3], f[1] * q[1]) * sc);
                            pY = (float)(fma_r(f[4], q[4], f[2] * q[2]) * sc);
                        }
                        const bool emit = act && (pM >= thr || (withIndels && (pX >= thr || pY >= thr)));
                        const uint64_t bm = __ballot(emit);
                        int base = 0;
                        if (lane == 0 && bm) base = atomicAdd(&ctl[1], __popcll(bm));
                        base = __shfl(base, 0, 64);
                        if (emit) {
                            const int idx = base + __popcll(bm & below);
                            if (idx < cap)
                                kp.out[outBase + idx] = make_int4(__float_as_int(pM), __float_as_int(pX),
                                                                  __float_as_int(pY), (int)rec_code(d, i));
                        }
                    }
                    lx.m(d)[i] = eM * q[0];
                    lx.g(d, 1)[i] = eX * q[1];
                    lx.g(d, 3)[i] = eX * q[3];
                    lx.g(d, 2)[i] = eY * q[2];
                    lx.g(d, 4)[i] = eY * q[4];
                    lx.e(d)[i] = E;
                }
                __syncthreads();
            }
            if (tid == 0) {
                kp.winEnd[2 * (jb.winOff + win)] = 0;
                kp.winEnd[2 * (jb.winOff + win) + 1] = ctl[1];
            }
            // ---- resume the forward sweep after tb: its send values at tb and
            // the match send values at tb - 1, rebuilt from the ring ----
            if (win + 1 < jb.nWin) {
                const int tb = wn.tb;
                int L0, R0, L1, R1;
                band(tb, L0, R0);
                band(tb - 1, L1, R1);
                for (int c = wave; c < C; c += NW) {
                    const int i = lane + 64 * c;
                    const bool a0 = L0 + 2 * i <= R0, a1 = L1 + 2 * i <= R1;
                    Real f[kNS5], g[kNS5];
#pragma unroll
                    for (int s = 0; s < kNS5; s++) {
                        f[s] = ringRow(tb, s)[i];
                        g[s] = ringRow(tb - 1, s)[i];
                    }
#pragma unroll
                    for (int tt = 0; tt < kNS5; tt++) {
                        Real a = T[tt] * f[0];
#pragma unroll
                        for (int s = 1; s < kNS5; s++) a = fma_r(T[s * kNS5 + tt], f[s], a);
                        if (tt == 0) lx.m(tb)[i] = a0 ? a : Real(0);
                        else lx.g(tb, tt)[i] = a0 ? a : Real(0);
                    }
                    lx.e(tb)[i] = a0 ? slot_exp(ringRow(tb, 5)[i]) : kNoMass;
                    Real a = T[0] * g[0];
#pragma unroll
                    for (int s = 1; s < kNS5; s++) a = fma_r(T[s * kNS5], g[s], a);
                    lx.m(tb - 1)[i] = a1 ? a : Real(0);
                    lx.e(tb - 1)[i] = a1 ? slot_exp(ringRow(tb - 1, 5)[i]) : kNoMass;
                }
                __syncthreads();
                dStart = tb + 1;
            }
        }
        if (tid == 0) {
            kp.counts[2 * job] = 0;
            kp.counts[2 * job + 1] = ctl[1];
        }
        __syncthreads();
    }
}

}  // namespace mpk
