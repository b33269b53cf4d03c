// mp_ref.hpp -- the pair-HMM in the reference's own arithmetic on the GPU:
// FP64 log space, the reference's polynomial logAdd (or exact log-sum-exp),
// the reference's operation order, windows and total-probability cadence.
//
// Reference algorithm: getPosteriorProbsWithBanding (pairwiseAligner.c:
// 687-808) with diagonalCalculationForward/Backward (528-565),
// diagonalCalculationTotalProbability (567-584), the posterior extraction
// (586-664), stateMachine3_cellCalculate (stateMachine.c:559-581) and the
// nucleotide / run-length emissions (stateMachine.c:360-380, 702-736;
// polisher.c:1756-1775).
//
// What must match the reference bit for bit is the order of the floating-point
// operations, not their parallel schedule: every cell of an anti-diagonal is
// independent, so the cells are spread over the workgroup's threads, and each
// cell performs exactly the reference's sequence of logAdds.  The forward cell
// pulls from its three predecessors in stateMachine3_cellCalculate's order.
// The reference's backward pushes each cell's values into its predecessors;
// the target cell of those pushes receives them in the order the reference
// walks the cells -- the match push of anti-diagonal e+2 first, then (on
// anti-diagonal e+1, cells by ascending x - y) the gap-Y push of (x, y+1),
// then the gap-X push of (x+1, y) -- and this kernel pulls them in that same
// order.  Totals fold the cells' dot products serially in the reference's cell
// order (one thread), every 10th posterior anti-diagonal of a window.
//
// One job (one sub-problem after large-gap splitting) per workgroup; a
// persistent grid takes jobs from a queue.  Per workgroup: the forward values
// of every anti-diagonal of the job, the backward values of three
// anti-diagonals, the cell offsets of the anti-diagonals, and two dot-product
// rows.  Tuples are written per job and list in the reference's append order
// (windows first to last, anti-diagonals descending, cells by ascending
// x - y); the list pass reverses each job's lists, as convertAlignedPairs does
// (pairwiseAligner.c:1068-1086).
#pragma once

#include "mp_device.hpp"
#include "mp_lists.hpp"

namespace mpk {

struct RefJob {
    int64_t symX, symY;   // first symbol of each sequence in the raw uint16 pool
    int64_t lrOff;        // the job's (xmyL, xmyR) per anti-diagonal
    int64_t outOff[3];    // the job's tuple regions (raw order), one per list
    int32_t cap[3];       // tuples each region holds
    int32_t lX, lY;
    int32_t flags;        // bit0 strand, bit1 ragged left, bit2 ragged right
    int32_t xAdd, yAdd;   // output coordinate shifts (split offset + reference crop - 1, split offset - 1)
};

struct RefModel {
    double trans[9];      // natural log: matchContinue, matchFromGapX, matchFromGapY, gapOpenX, gapOpenY,
                          // gapExtendX, gapExtendY, gapSwitchToX, gapSwitchToY
    double match[2][16];  // [strand][x base * 4 + y base], natural log
    double gapX[2][4], gapY[2][4];
    int32_t rle;
};

struct RefParams {
    const RefJob *jobs;
    int32_t njobs;
    const uint16_t *sym;
    const int2 *lr;
    const double *repeat;  // [row][underlying][observed] log10, rows 0-3, 51 x 51 (run-length models)
    RefModel model;
    double threshold;
    int32_t withIndels;
    int32_t poly;          // 1: the reference's polynomial logAdd, 0: exact log-sum-exp
    int64_t minDiags, tbd, expansion;
    double *F;             // per workgroup fStride doubles
    int64_t fStride;
    double *B;             // per workgroup 3 * bStride doubles
    int64_t bStride;       // 3 * widest anti-diagonal
    int64_t *offs;         // per workgroup oStride
    int64_t oStride;       // longest job's anti-diagonals + 2
    double *dots;          // per workgroup 2 * wStride
    int64_t wStride;       // widest anti-diagonal
    ListTuple *out;        // raw tuples
    int32_t *counts;       // per job: tuples of each list (may exceed cap: the host grows and runs again)
    double *logTotal;      // per job: log-likelihood of its last window
    int32_t *queue;
    int32_t *error;
};

constexpr double kLogZero = -__builtin_inf();

// The reference's lookup (pairwiseAligner.c:279-299): float-suffixed literals
// evaluated in double; the casts reproduce that rounding.
#define MPK_FC(v) ((double)(float)(v))
__device__ __forceinline__ double ref_poly_log1pexp(double x) {
    if (x <= MPK_FC(1.00))
        return ((MPK_FC(-0.009350833524763) * x + MPK_FC(0.130659527668286)) * x + MPK_FC(0.498799810682272)) * x +
               MPK_FC(0.693203116424741);
    if (x <= MPK_FC(2.50))
        return ((MPK_FC(-0.014532321752540) * x + MPK_FC(0.139942324101744)) * x + MPK_FC(0.495635523139337)) * x +
               MPK_FC(0.692140569840976);
    if (x <= MPK_FC(4.50))
        return ((MPK_FC(-0.004605031767994) * x + MPK_FC(0.063427417320019)) * x + MPK_FC(0.695956496475118)) * x +
               MPK_FC(0.514272634594009);
    return ((MPK_FC(-0.000458661602210) * x + MPK_FC(0.009695946122598)) * x + MPK_FC(0.930734667215156)) * x +
           MPK_FC(0.168037164329057);
}
#undef MPK_FC

template <bool POLY>
__device__ __forceinline__ double ref_logadd(double x, double y) {
    if constexpr (POLY) {
        if (x < y) return (x == kLogZero || y - x >= 7.5) ? y : ref_poly_log1pexp(y - x) + x;
        return (y == kLogZero || x - y >= 7.5) ? x : ref_poly_log1pexp(x - y) + y;
    } else {
        if (x < y) return x == kLogZero ? y : y + log1p(exp(x - y));
        return y == kLogZero ? x : x + log1p(exp(y - x));
    }
}

// exp of a log posterior.  Weights are floor(p * 1e7), so the last bit of p
// decides a weight whenever p is within rounding of 1 - k * 1e-7, which happens
// at p ~ 1 (|x| below ~1e-7) on every well-aligned cell.  There the host
// library's FP64 exp is correctly rounded, and so is 1 + (x + x^2 / 2) for
// |x| < 2^-24 (the cubic term is below 2^-72); the device exp is within one
// unit of the last place elsewhere, where a weight flips only if p * 1e7 lies
// within ~1e-9 of an integer.
__device__ __forceinline__ double ref_posterior_exp(double x) {
    return fabs(x) < 0x1p-24 ? 1.0 + (x + x * x * 0.5) : exp(x);
}

struct RefEmit {
    const RefModel *m;
    const double *repeat;
    int strand;
    __device__ __forceinline__ double gapX(uint16_t s) const {
        const int b = m->rle ? (s & 255) : s;
        return b >= 4 ? -1.386294361 : m->gapX[strand][b];
    }
    __device__ __forceinline__ double gapY(uint16_t s) const {
        const int b = m->rle ? (s & 255) : s;
        return b >= 4 ? -1.386294361 : m->gapY[strand][b];
    }
    __device__ __forceinline__ double match(uint16_t sx, uint16_t sy) const {
        if (!m->rle) return (sx >= 4 || sy >= 4) ? -2.772588722 : m->match[strand][sx * 4 + sy];
        const int xb = sx & 255, yb = sy & 255;
        const double e = (xb >= 4 || yb >= 4) ? -2.772588722 : m->match[strand][xb * 4 + yb];
        // (N reads the A row; run lengths beyond the table's 50 read its last entry)
        const int b = xb >= 4 ? 0 : xb;
        const int row = strand ? b : 3 - b;
        const double lp = repeat[(int64_t)row * 51 * 51 + (int64_t)min(sx >> 8, 50) * 51 + min(sy >> 8, 50)];
        return e + 2.3025 * lp;
    }
};

// NT threads per workgroup, one job at a time: 256, or 512 for batches with bands
// wider than 512 cells (the stitch rectangles), whose anti-diagonals take most of
// the time (1024 would cap the kernel at 128 registers and spill).
template <bool POLY, int NT>
__global__ __launch_bounds__(NT) void pairhmm_ref_kernel(RefParams rp) {
    __shared__ int sJob;
    __shared__ double sTotal;
    __shared__ int sScan[3][NT / 64];
    __shared__ int sCount[3];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const double T_MM = rp.model.trans[0], T_XM = rp.model.trans[1], T_YM = rp.model.trans[2];
    const double T_MX = rp.model.trans[3], T_MY = rp.model.trans[4], T_XX = rp.model.trans[5];
    const double T_YY = rp.model.trans[6], T_YX = rp.model.trans[7], T_XY = rp.model.trans[8];
    double *F = rp.F + (size_t)blockIdx.x * rp.fStride;
    double *Bb = rp.B + (size_t)blockIdx.x * 3 * rp.bStride;
    int64_t *offs = rp.offs + (size_t)blockIdx.x * rp.oStride;
    double *dotA = rp.dots + (size_t)blockIdx.x * 2 * rp.wStride;
    double *dotB = dotA + rp.wStride;
    for (;;) {
        if (tid == 0) sJob = atomicAdd(rp.queue, 1);
        __syncthreads();
        const int job = sJob;
        __syncthreads();
        if (job >= rp.njobs) break;
        const RefJob J = rp.jobs[job];
        const int64_t D = (int64_t)J.lX + J.lY;
        const int2 *lr = rp.lr + J.lrOff;
        const uint16_t *sX = rp.sym + J.symX, *sY = rp.sym + J.symY;
        const RefEmit em{&rp.model, rp.repeat, J.flags & 1};
        auto width = [&](int64_t d) { return (int64_t)((lr[d].y - lr[d].x) / 2 + 1); };
        auto xsym = [&](int64_t x) -> uint16_t { return x > 0 ? sX[x - 1] : (uint16_t)4; };
        auto ysym = [&](int64_t y) -> uint16_t { return y > 0 ? sY[y - 1] : (uint16_t)4; };
        int cnt[3] = {0, 0, 0};
        if (tid < 3) sCount[tid] = 0;
        // cell offsets of the anti-diagonals (F storage)
        {
            int64_t carry = 0;
            __shared__ int64_t sh64[NT / 64];
            for (int64_t d0 = 0; d0 <= D; d0 += NT) {
                const int64_t d = d0 + tid;
                const int64_t v = d <= D ? width(d) : 0;
                int64_t tot;
                const int64_t pre = block_excl_scan<int64_t, NT>(v, sh64, &tot);
                if (d <= D) offs[d] = carry + pre;
                carry += tot;
            }
            if (tid == 0) offs[D + 1] = carry;
        }
        __syncthreads();
        auto Fc = [&](int64_t d, int64_t xmy) -> double * {  // forward cell or null (outside the band)
            if (d < 0 || xmy < lr[d].x || xmy > lr[d].y) return nullptr;
            return F + 3 * (offs[d] + (xmy - lr[d].x) / 2);
        };
        auto Bc = [&](int64_t d, int64_t xmy) -> double * {
            return Bb + (d % 3) * rp.bStride + 3 * ((xmy - lr[d].x) / 2);
        };
        // forward start (anti-diagonal 0: the cell (0, 0))
        if (tid == 0) {
            const bool rg = (J.flags >> 1) & 1;
            F[0] = rg ? kLogZero : 0.0;
            F[1] = rg ? 0.0 : kLogZero;
            F[2] = rg ? 0.0 : kLogZero;
        }
        __syncthreads();
        int64_t tracedBackTo = 0;
        for (int64_t xay = 1; xay <= D; xay++) {
            // ---- forward anti-diagonal xay (cellCalculate with doTransitionForward) ----
            const int64_t w = width(xay), L0 = lr[xay].x;
            for (int64_t i = tid; i < w; i += NT) {
                const int64_t xmy = L0 + 2 * i, x = (xay + xmy) / 2, y = (xay - xmy) / 2;
                double *cur = F + 3 * (offs[xay] + i);
                double cM = kLogZero, cX = kLogZero, cY = kLogZero;
                const double *lo = Fc(xay - 1, xmy - 1), *mid = xay >= 2 ? Fc(xay - 2, xmy) : nullptr,
                             *up = Fc(xay - 1, xmy + 1);
                const uint16_t cx = xsym(x), cy = ysym(y);
                if (lo) {
                    const double e = em.gapX(cx);
                    cX = ref_logadd<POLY>(cX, lo[0] + (e + T_MX));
                    cX = ref_logadd<POLY>(cX, lo[1] + (e + T_XX));
                    cX = ref_logadd<POLY>(cX, lo[2] + (e + T_YX));
                }
                if (mid) {
                    const double e = em.match(cx, cy);
                    cM = ref_logadd<POLY>(cM, mid[0] + (e + T_MM));
                    cM = ref_logadd<POLY>(cM, mid[1] + (e + T_XM));
                    cM = ref_logadd<POLY>(cM, mid[2] + (e + T_YM));
                }
                if (up) {
                    const double e = em.gapY(cy);
                    cY = ref_logadd<POLY>(cY, up[0] + (e + T_MY));
                    cY = ref_logadd<POLY>(cY, up[2] + (e + T_YY));
                    cY = ref_logadd<POLY>(cY, up[1] + (e + T_XY));
                }
                cur[0] = cM;
                cur[1] = cX;
                cur[2] = cY;
            }
            __syncthreads();
            const bool atEnd = xay == D;
            const bool tbPoint = xay >= tracedBackTo + rp.minDiags && w <= rp.expansion * 2 + 1;
            if (!(atEnd || tbPoint)) continue;
            // ---- traceback window ending at xay ----
            {
                const bool rg = atEnd && ((J.flags >> 2) & 1);
                const double e0 = rg ? (rp.model.trans[3] + rp.model.trans[4]) / 2.0 : T_MM;
                const double e1 = rg ? T_XX : T_XM, e2 = rg ? T_YY : T_YM;
                for (int64_t i = tid; i < w; i += NT) {
                    double *b = Bc(xay, L0 + 2 * i);
                    b[0] = e0;
                    b[1] = e1;
                    b[2] = e2;
                }
            }
            const int64_t from = xay - (atEnd ? 0 : rp.tbd + 1);
            double total = kLogZero;
            int64_t count = 0;
            __syncthreads();
            for (int64_t d2 = xay; d2 > tracedBackTo; d2--) {
                const int64_t w2 = width(d2), L2 = lr[d2].x;
                if (d2 < xay) {
                    // backward values of d2: the pushes of d2 + 2 (match) and of d2 + 1 (gap Y of
                    // (x, y+1), then gap X of (x+1, y)), in the reference's order
                    for (int64_t i = tid; i < w2; i += NT) {
                        const int64_t xmy = L2 + 2 * i, x = (d2 + xmy) / 2, y = (d2 - xmy) / 2;
                        double bM = kLogZero, bX = kLogZero, bY = kLogZero;
                        if (d2 + 2 <= xay && xmy >= lr[d2 + 2].x && xmy <= lr[d2 + 2].y) {
                            const double *p = Bc(d2 + 2, xmy);
                            const double e = em.match(xsym(x + 1), ysym(y + 1));
                            bM = ref_logadd<POLY>(bM, p[0] + (e + T_MM));
                            bX = ref_logadd<POLY>(bX, p[0] + (e + T_XM));
                            bY = ref_logadd<POLY>(bY, p[0] + (e + T_YM));
                        }
                        if (xmy - 1 >= lr[d2 + 1].x && xmy - 1 <= lr[d2 + 1].y) {
                            const double *u = Bc(d2 + 1, xmy - 1);
                            const double e = em.gapY(ysym(y + 1));
                            bM = ref_logadd<POLY>(bM, u[2] + (e + T_MY));
                            bY = ref_logadd<POLY>(bY, u[2] + (e + T_YY));
                            bX = ref_logadd<POLY>(bX, u[2] + (e + T_XY));
                        }
                        if (xmy + 1 >= lr[d2 + 1].x && xmy + 1 <= lr[d2 + 1].y) {
                            const double *l = Bc(d2 + 1, xmy + 1);
                            const double e = em.gapX(xsym(x + 1));
                            bM = ref_logadd<POLY>(bM, l[1] + (e + T_MX));
                            bX = ref_logadd<POLY>(bX, l[1] + (e + T_XX));
                            bY = ref_logadd<POLY>(bY, l[1] + (e + T_YX));
                        }
                        double *b = Bc(d2, xmy);
                        b[0] = bM;
                        b[1] = bX;
                        b[2] = bY;
                    }
                    __syncthreads();
                }
                if (d2 > from) continue;
                if (count++ % 10 == 0) {
                    // diagonalCalculationTotalProbability: the cut at d2 plus the matches through it
                    for (int64_t i = tid; i < w2; i += NT) {
                        const double *f = F + 3 * (offs[d2] + i), *b = Bc(d2, L2 + 2 * i);
                        double t = f[0] + b[0];
                        t = ref_logadd<POLY>(t, f[1] + b[1]);
                        t = ref_logadd<POLY>(t, f[2] + b[2]);
                        dotA[i] = t;
                    }
                    const bool through = d2 + 1 <= xay && d2 + 1 <= D;
                    const int64_t w1 = through ? width(d2 + 1) : 0, L1 = through ? lr[d2 + 1].x : 0;
                    for (int64_t i = tid; i < w1; i += NT) {
                        const int64_t xmy = L1 + 2 * i, x = (d2 + 1 + xmy) / 2, y = (d2 + 1 - xmy) / 2;
                        const double *mid = Fc(d2 - 1, xmy);
                        double mt = kLogZero;
                        if (mid) {
                            const double e = em.match(xsym(x), ysym(y));
                            mt = ref_logadd<POLY>(mt, mid[0] + (e + T_MM));
                            mt = ref_logadd<POLY>(mt, mid[1] + (e + T_XM));
                            mt = ref_logadd<POLY>(mt, mid[2] + (e + T_YM));
                        }
                        const double *b = Bc(d2 + 1, xmy);
                        double t = mt + b[0];
                        t = ref_logadd<POLY>(t, kLogZero + b[1]);
                        t = ref_logadd<POLY>(t, kLogZero + b[2]);
                        dotB[i] = t;
                    }
                    __syncthreads();
                    if (tid == 0) {
                        double t = kLogZero;
                        for (int64_t i = 0; i < w2; i++) t = ref_logadd<POLY>(t, dotA[i]);
                        if (through) {
                            double acc = kLogZero;
                            for (int64_t i = 0; i < w1; i++) acc = ref_logadd<POLY>(acc, dotB[i]);
                            t = ref_logadd<POLY>(t, acc);
                        }
                        sTotal = t;
                        if (count == 1 && atEnd) rp.logTotal[job] = t;
                    }
                    __syncthreads();
                    total = sTotal;
                }
                // ---- posteriors of d2 (addPosteriorProb), cells by ascending x - y ----
                for (int64_t i0 = 0; i0 < w2; i0 += NT) {
                    const int64_t i = i0 + tid;
                    int64_t wgt[3] = {0, 0, 0};
                    bool on[3] = {false, false, false};
                    int64_t x = 0, y = 0;
                    if (i < w2) {
                        const int64_t xmy = L2 + 2 * i;
                        x = (d2 + xmy) / 2;
                        y = (d2 - xmy) / 2;
                        const double *f = F + 3 * (offs[d2] + i), *b = Bc(d2, xmy);
                        const bool ok[3] = {x > 0 && y > 0, rp.withIndels && x > 0, rp.withIndels && y > 0};
                        for (int s = 0; s < 3; s++) {
                            if (!ok[s]) continue;
                            double pr = ref_posterior_exp((f[s] + b[s]) - total);
                            if (pr >= rp.threshold) {
                                if (pr > 1.0) pr = 1.0;
                                wgt[s] = (int64_t)floor(pr * 10000000.0);
                                on[s] = true;
                            }
                        }
                    }
                    for (int l = 0; l < (rp.withIndels ? 3 : 1); l++) {  // match-only batches: one list
                        const uint64_t m = __ballot(on[l]);
                        const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                        if (lane == 0) sScan[l][wv] = __popcll(m);
                        __syncthreads();
                        int before = 0, tot = 0;
                        for (int k = 0; k < NT / 64; k++) {
                            if (k < wv) before += sScan[l][k];
                            tot += sScan[l][k];
                        }
                        const int idx = cnt[l] + before + below;
                        if (on[l] && idx < J.cap[l])
                            rp.out[J.outOff[l] + idx] = ListTuple{wgt[l], (int32_t)(x + J.xAdd), (int32_t)(y + J.yAdd)};
                        cnt[l] += tot;
                        __syncthreads();
                    }
                }
            }
            tracedBackTo = from;
            __syncthreads();
        }
        if (tid == 0) {
            rp.counts[3 * job] = cnt[0];
            rp.counts[3 * job + 1] = cnt[1];
            rp.counts[3 * job + 2] = cnt[2];
            if (D == 0) rp.logTotal[job] = 0.0;
        }
        __syncthreads();
    }
}

// Per slot: the tuples of each list, exclusive scan -> each job's first
// tuple of each list; a job whose tuples exceed its region sets *overflow.
__global__ __launch_bounds__(1024) void ref_slot_scan(const int32_t *slotJob, int32_t nslots, const int32_t *counts,
                                                      const RefJob *jobs, const int32_t *error, int64_t *slotBase3,
                                                      int32_t *overflow) {
    __shared__ int64_t sh[3][16];
    int64_t carry[3] = {0, 0, 0};
    bool ovf = false;
    const bool bad = *error != 0;
    for (int32_t s0 = 0; s0 < nslots; s0 += 1024) {
        const int32_t s = s0 + (int32_t)threadIdx.x;
        int64_t v[3] = {0, 0, 0};
        if (s < nslots && !bad && slotJob[s] >= 0) {
            const int32_t q = slotJob[s];
            for (int l = 0; l < 3; l++) {
                const int32_t c = counts[3 * q + l];
                ovf |= c > jobs[q].cap[l] || c < 0;
                v[l] = min(max(c, 0), jobs[q].cap[l]);
            }
        }
        for (int l = 0; l < 3; l++) {
            int64_t tot;
            const int64_t pre = block_excl_scan<int64_t, 1024>(v[l], sh[l], &tot);
            if (s < nslots) slotBase3[3 * (int64_t)s + l] = carry[l] + pre;
            carry[l] += tot;
        }
    }
    if (threadIdx.x == 0)
        for (int l = 0; l < 3; l++) slotBase3[3 * (int64_t)nslots + l] = carry[l];
    if (__any(ovf) && (threadIdx.x & 63) == 0) atomicOr(overflow, 1);
}

// Each job's raw tuples, reversed, at its place in the three lists (one
// workgroup per job); a list position at or beyond the list's capacity is
// dropped (the host grows that list and assembles again).
__global__ __launch_bounds__(256) void ref_lists_copy(const int32_t *slotJob, int32_t nslots, const int32_t *counts,
                                                      const RefJob *jobs, const ListTuple *raw,
                                                      const int64_t *slotBase3, const int32_t *skip, int4 cap, int pk,
                                                      ListTuple *out0, ListTuple *out1, ListTuple *out2) {
    if (*skip) return;
    ListTuple *out[3] = {out0, out1, out2};
    const int capL[3] = {cap.x, cap.y, cap.z};
    for (int32_t s = blockIdx.x; s < nslots; s += gridDim.x) {
        const int32_t q = slotJob[s];
        if (q < 0) continue;
        const RefJob J = jobs[q];
        for (int l = 0; l < 3; l++) {
            const int32_t n = counts[3 * q + l];
            const int64_t base = slotBase3[3 * (int64_t)s + l];
            for (int32_t i = threadIdx.x; i < n; i += blockDim.x)
                if (base + i < capL[l]) {
                    const ListTuple r = raw[J.outOff[l] + (n - 1 - i)];
                    put_tuple(out[l], base + i, pk, (int)r.weight, r.x, r.y);
                }
        }
    }
}

__global__ __launch_bounds__(256) void ref_pair_starts(const int32_t *pairSlot, int64_t np,
                                                       const int64_t *slotBase3, const int32_t *skip,
                                                       int64_t *starts, int4 *totals) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= np; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = *skip ? 0 : pairSlot[i];
        for (int l = 0; l < 3; l++) starts[l * (np + 1) + i] = *skip ? 0 : slotBase3[3 * s + l];
        if (i == np)
            *totals = *skip ? make_int4(0, 0, 0, 0)
                            : make_int4((int)slotBase3[3 * s], (int)slotBase3[3 * s + 1], (int)slotBase3[3 * s + 2], 0);
    }
}

// FP64 log-likelihood per pair: the sum over its sub-problems of the last
// window's total.
__global__ __launch_bounds__(256) void ref_pair_log_likelihood(const int32_t *pairSlot, int64_t np,
                                                               const int32_t *slotJob, const double *logTotal,
                                                               const int32_t *error, double *out) {
    if (*error) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < np; i += (int64_t)gridDim.x * blockDim.x) {
        double ll = 0.0;
        for (int32_t s = pairSlot[i]; s < pairSlot[i + 1]; s++)
            if (slotJob[s] >= 0) ll += logTotal[slotJob[s]];
        out[i] = ll;
    }
}

}  // namespace mpk
