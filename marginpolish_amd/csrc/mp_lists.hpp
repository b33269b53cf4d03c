// Posterior lists and log-likelihoods assembled on the device, inside the
// batch's run (every size is read from device memory, so the whole pass is
// captured in the run's second hipGraph and needs no host round trip).
//
// The pair-HMM kernels leave per job two record streams (forward side,
// backward side), each a run of anti-diagonal blocks in sweep order with the
// cells of a block in lane order.  The reference's lists
// (getPosteriorProbsWithBanding, pairwiseAligner.c:744-782, with the
// reversal of convertAlignedPairs at 1072-1074) visit the windows last to
// first and, inside a window, the anti-diagonals ascending with the cells of
// one anti-diagonal by x descending; a cell joins the match list at
// p_M >= threshold with x, y > 0, the gap-X list at p_X >= threshold with
// x > 0 and the gap-Y list at p_Y >= threshold with y > 0 (586-664), weight
// floor(min(p, 1) * 1e7).
//
// Jobs sit in pair-major "slots": a pair's sub-problems in split order, the
// window jobs of one sub-problem last window first (so the slot order is the
// emission order).
// 1. `slot_record_scan`: record count per slot, exclusive scan -> each job's
//    first emission index; flags an overflowing record region.
// 2. `order_records`: one workgroup per job side places every record at its
//    emission index and decodes it: coordinates, three weights, list flags.
// 3. `list_flag_sums` / `list_block_offsets` / `list_scan_emit`: a scan of the
//    three list flags in emission order that writes the (weight, x, y) tuples.
// 4. `pair_list_starts`: per-pair list offsets.
// 5. `pair_log_likelihood` / `batch_log_likelihood`: FP64 log-likelihood per
//    pair (window jobs chained through their traceback anti-diagonals) and
//    their sum, in a fixed order.
#pragma once

#include "mp_device.hpp"

namespace mpk {

struct JobLists {
    int64_t outF, outB;     // record regions: forward side, backward side
    int64_t bandOff;        // the job's band entries
    int32_t capF, capB;     // record capacities of the two sides
    int32_t winOff, nWin;
    int32_t xAdd, yAdd;     // output coordinate shifts (split offset + reference crop - 1, split offset - 1)
    int32_t slot;           // the job's slot
    int32_t pairSlot;       // first slot of the job's pair
};

struct ListTuple {          // the C ABI's mp_aligned_pair
    int64_t weight;
    int32_t x, y;
};

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;                              // records per thread
constexpr int kScanTile = kScanThreads * kScanItems;       // records per workgroup

// Exclusive scan over a workgroup of 64-lane waves; returns this thread's
// prefix and writes the workgroup total to *total (every thread).
template <class T, int NT>
__device__ __forceinline__ T block_excl_scan(T v, T *sh, T *total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    T x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const T y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[w] = x;
    __syncthreads();
    T waveOff = 0, tot = 0;
    for (int i = 0; i < NT / 64; i++) {
        if (i < w) waveOff += sh[i];
        tot += sh[i];
    }
    __syncthreads();
    *total = tot;
    return waveOff + x - v;
}

// Record count per slot and its exclusive scan (one workgroup).  A slot whose
// job wrote more records than its region holds sets *overflow (the host then
// sizes the regions exactly and runs the batch again).
__global__ __launch_bounds__(1024) void slot_record_scan(const int32_t *slotJob, int32_t nslots,
                                                         const int32_t *counts, const JobLists *jl,
                                                         const int32_t *error, int64_t *slotBase,
                                                         int32_t *overflow) {
    __shared__ int64_t sh[16];
    int64_t carry = 0;
    bool ovf = false;
    const bool bad = *error != 0;
    for (int32_t s0 = 0; s0 < nslots; s0 += 1024) {
        const int32_t s = s0 + (int32_t)threadIdx.x;
        int64_t v = 0;
        if (s < nslots && !bad) {
            const int32_t q = slotJob[s];
            if (q >= 0) {
                const int32_t cF = counts[2 * q], cB = counts[2 * q + 1];
                const JobLists J = jl[q];
                ovf |= cF > J.capF || cB > J.capB || cF < 0 || cB < 0;
                v = (int64_t)min(max(cF, 0), J.capF) + min(max(cB, 0), J.capB);
            }
        }
        int64_t tot;
        const int64_t pre = block_excl_scan<int64_t, 1024>(v, sh, &tot);
        if (s < nslots) slotBase[s] = carry + pre;
        carry += tot;
    }
    if (threadIdx.x == 0) slotBase[nslots] = carry;
    if (__any(ovf) && (threadIdx.x & 63) == 0) atomicOr(overflow, 1);
}

// Records of one job side are ordered in tiles of kOrderTile: the tile's
// record codes, with kOrderHalo records either side, are staged in LDS, so the
// scans for the run of records on the same anti-diagonal (and the cell rank in
// it) read LDS; a run reaching beyond the halo (bands wider than kOrderHalo
// cells) reads the rest from global memory.
constexpr int kOrderTile = 256;
constexpr int kOrderHalo = 64;

__global__ __launch_bounds__(kOrderTile) void order_records(const int4 *rec, const JobLists *jl,
                                                            const int32_t *counts, const int64_t *slotBase,
                                                            const int32_t *winEnd, const int *band,
                                                            const int32_t *skip, const int32_t *error, float thr,
                                                            int withIndels, int nseg, int2 *xy, int4 *wf,
                                                            uint8_t *flg) {
    __shared__ uint32_t code[kOrderTile + 2 * kOrderHalo];
    if (*skip || *error) return;  // (grid-uniform)
    for (int seg = blockIdx.x; seg < nseg; seg += gridDim.x) {
        const JobLists J = jl[seg >> 1];
        const int side = seg & 1;  // 0 forward, 1 backward
        const int nF = counts[2 * (seg >> 1)], nB = counts[2 * (seg >> 1) + 1];
        const int n = side ? nB : nF;
        const int4 *r = rec + (side ? J.outB : J.outF);
        const int32_t *we = winEnd + 2 * J.winOff;
        const int64_t base = slotBase[J.slot];
        for (int t0 = 0; t0 < n; t0 += kOrderTile) {
            __syncthreads();  // the previous tile's LDS reads are done
            for (int i = threadIdx.x; i < kOrderTile + 2 * kOrderHalo; i += kOrderTile) {
                const int gi = t0 - kOrderHalo + i;
                code[i] = gi >= 0 && gi < n ? (uint32_t)r[gi].w : 0xffffffffu;
            }
            __syncthreads();
            const int t = t0 + (int)threadIdx.x;
            if (t >= n) continue;
            auto codeAt = [&](int i) -> uint32_t {
                const int li = i - t0 + kOrderHalo;
                return li >= 0 && li < kOrderTile + 2 * kOrderHalo ? code[li] : (uint32_t)r[i].w;
            };
            int w = 0;
            while (w + 1 < J.nWin && t >= we[2 * w + side]) w++;
            const int fBeg = w ? we[2 * (w - 1)] : 0, fEnd = we[2 * w];
            const int bBeg = w ? we[2 * (w - 1) + 1] : 0, bEnd = we[2 * w + 1];
            const int beg = side ? bBeg : fBeg, end = min(side ? bEnd : fEnd, n);
            // windows after this one come first
            const int64_t off = (int64_t)(nF - fEnd) + (nB - bEnd);
            const int4 me = r[t];
            const uint32_t d = (uint32_t)me.w >> 11, cell = (uint32_t)me.w & 2047u;
            int u = t, v = t + 1;
            while (u > beg && (codeAt(u - 1) >> 11) == d) u--;
            while (v < end && (codeAt(v) >> 11) == d) v++;
            int rank = 0;
            for (int i = u; i < v; i++) rank += (codeAt(i) & 2047u) > cell;
            // backward blocks are stored anti-diagonals descending: walk them from the end
            const int64_t pos = side ? (int64_t)(bEnd - v) + rank : (int64_t)(bEnd - bBeg) + (u - fBeg) + rank;
            const int64_t g = base + off + pos;
            const int x = band_xL(band[J.bandOff + d]) + (int)cell, y = (int)d - x;
            const float pM = __int_as_float(me.x), pX = __int_as_float(me.y), pY = __int_as_float(me.z);
            const int flags = ((pM >= thr) & (x > 0) & (y > 0)) | (((withIndels != 0) & (pX >= thr) & (x > 0)) << 1) |
                              (((withIndels != 0) & (pY >= thr) & (y > 0)) << 2);
            xy[g] = make_int2(x + J.xAdd, y + J.yAdd);
            wf[g] = make_int4((int)floorf(fminf(pM, 1.f) * 1e7f), (int)floorf(fminf(pX, 1.f) * 1e7f),
                              (int)floorf(fminf(pY, 1.f) * 1e7f), flags);
            flg[g] = (uint8_t)flags;
        }
    }
}

// The three list scans read the records in emission order in chunks of 64
// (one wavefront, coalesced); a chunk's list prefixes come from ballots.
// Tile t covers records [t * kScanTile, (t + 1) * kScanTile): chunk (i, w) =
// records tile0 + i * 256 + w * 64 + lane, i = 0..7, wave w = 0..3.
__device__ __forceinline__ uint64_t ballot_bits(bool b) { return __ballot(b); }
__device__ __forceinline__ int prefix_in_wave(uint64_t m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Per tile: the number of tuples of each list.
__global__ __launch_bounds__(kScanThreads) void list_flag_sums(const uint8_t *flg, const int64_t *nPtr,
                                                               const int32_t *skip, int4 *tileSums) {
    __shared__ int sh[3][kScanThreads / 64];
    const int64_t n = *skip ? 0 : *nPtr;
    const int64_t tile0 = (int64_t)blockIdx.x * kScanTile;
    if (tile0 >= n) return;  // workgroup-uniform
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int c0 = 0, c1 = 0, c2 = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        const int64_t g = tile0 + i * kScanThreads + threadIdx.x;
        const int f = g < n ? flg[g] : 0;
        c0 += __popcll(ballot_bits(f & 1));
        c1 += __popcll(ballot_bits(f & 2));
        c2 += __popcll(ballot_bits(f & 4));
    }
    if (lane == 0) {
        sh[0][w] = c0;
        sh[1][w] = c1;
        sh[2][w] = c2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int4 s = make_int4(0, 0, 0, 0);
        for (int k = 0; k < kScanThreads / 64; k++) {
            s.x += sh[0][k];
            s.y += sh[1][k];
            s.z += sh[2][k];
        }
        tileSums[blockIdx.x] = s;
    }
}

// Exclusive scan of the tile sums of the ceil(n / kScanTile) tiles in place
// (one workgroup, four tiles per thread per pass); the list lengths go to
// *totals.
__global__ __launch_bounds__(1024) void list_block_offsets(int4 *tileSums, const int64_t *nPtr,
                                                           const int32_t *skip, int4 *totals) {
    __shared__ int sh[3][16];
    const int64_t n = *skip ? 0 : *nPtr;
    const int64_t nTiles = (n + kScanTile - 1) / kScanTile;
    int carry[3] = {0, 0, 0};
    for (int64_t b0 = 0; b0 < nTiles; b0 += 4 * 1024) {
        const int64_t b = b0 + 4 * (int64_t)threadIdx.x;
        int4 v[4];
        int mine[3] = {0, 0, 0};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            v[k] = b + k < nTiles ? tileSums[b + k] : make_int4(0, 0, 0, 0);
            mine[0] += v[k].x;
            mine[1] += v[k].y;
            mine[2] += v[k].z;
        }
        int pre[3], tot[3];
        for (int l = 0; l < 3; l++) pre[l] = carry[l] + block_excl_scan<int, 1024>(mine[l], sh[l], &tot[l]);
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (b + k < nTiles) {
                tileSums[b + k] = make_int4(pre[0], pre[1], pre[2], 0);
                pre[0] += v[k].x;
                pre[1] += v[k].y;
                pre[2] += v[k].z;
            }
        for (int l = 0; l < 3; l++) carry[l] += tot[l];
    }
    if (threadIdx.x == 0) *totals = make_int4(carry[0], carry[1], carry[2], 0);
}

// Every tuple written at its list position (positions at or beyond a list's
// capacity are dropped: the host then grows that list and assembles again).
__global__ __launch_bounds__(kScanThreads) void list_scan_emit(const int2 *xy, const int4 *wf, const int64_t *nPtr,
                                                               const int32_t *skip, const int4 *tileOff,
                                                               int4 cap, ListTuple *out0, ListTuple *out1,
                                                               ListTuple *out2) {
    __shared__ int cnt[3][kScanItems * kScanThreads / 64];
    const int64_t n = *skip ? 0 : *nPtr;
    const int64_t tile0 = (int64_t)blockIdx.x * kScanTile;
    if (tile0 >= n) return;  // workgroup-uniform
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    constexpr int NC = kScanItems * kScanThreads / 64;  // chunks per tile
    int f[kScanItems];
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        const int64_t g = tile0 + i * kScanThreads + threadIdx.x;
        f[i] = g < n ? wf[g].w : 0;
        const int c = i * (kScanThreads / 64) + w;
        const int a = __popcll(ballot_bits(f[i] & 1)), b = __popcll(ballot_bits(f[i] & 2)),
                  d = __popcll(ballot_bits(f[i] & 4));
        if (lane == 0) {
            cnt[0][c] = a;
            cnt[1][c] = b;
            cnt[2][c] = d;
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {  // exclusive prefix over the tile's chunks, per list
        int run = 0;
        for (int c = 0; c < NC; c++) {
            const int v = cnt[threadIdx.x][c];
            cnt[threadIdx.x][c] = run;
            run += v;
        }
    }
    __syncthreads();
    const int4 o = tileOff[blockIdx.x];
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        const int c = i * (kScanThreads / 64) + w;
        const uint64_t m0 = ballot_bits(f[i] & 1), m1 = ballot_bits(f[i] & 2), m2 = ballot_bits(f[i] & 4);
        if (!f[i]) continue;
        const int64_t g = tile0 + i * kScanThreads + threadIdx.x;
        const int4 wv = wf[g];
        const int2 c2 = xy[g];
        const int p0 = o.x + cnt[0][c] + prefix_in_wave(m0);
        const int p1 = o.y + cnt[1][c] + prefix_in_wave(m1);
        const int p2 = o.z + cnt[2][c] + prefix_in_wave(m2);
        if ((f[i] & 1) && p0 < cap.x) out0[p0] = ListTuple{wv.x, c2.x, c2.y};
        if ((f[i] & 2) && p1 < cap.y) out1[p1] = ListTuple{wv.y, c2.x, c2.y};
        if ((f[i] & 4) && p2 < cap.z) out2[p2] = ListTuple{wv.z, c2.x, c2.y};
    }
}

// starts[l * (np + 1) + i] = index in list l of pair i's first tuple; entry np
// = list length.  One wavefront per pair: the tile's offset plus the list
// flags of the tile's records before the pair's first record.
__global__ __launch_bounds__(256) void pair_list_starts(const int32_t *pairSlot, int64_t np, const int64_t *slotBase,
                                                        int32_t nslots, const uint8_t *flg, const int4 *tileOff,
                                                        const int4 *totals, const int32_t *skip, int64_t *starts) {
    const int lane = threadIdx.x & 63;
    const int64_t n = *skip ? 0 : slotBase[nslots];
    for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; i <= np;
         i += ((int64_t)gridDim.x * blockDim.x) >> 6) {
        const int64_t g = *skip ? 0 : slotBase[pairSlot[i]];
        int4 p;
        if (g >= n) {
            p = *skip ? make_int4(0, 0, 0, 0) : *totals;
        } else {
            const int64_t t = g / kScanTile, t0 = t * kScanTile;
            int c0 = 0, c1 = 0, c2 = 0;
            for (int64_t k = t0 + lane; k < g; k += 64) {
                const int f = flg[k];
                c0 += f & 1;
                c1 += (f >> 1) & 1;
                c2 += (f >> 2) & 1;
            }
            for (int o = 32; o > 0; o >>= 1) {
                c0 += __shfl_xor(c0, o, 64);
                c1 += __shfl_xor(c1, o, 64);
                c2 += __shfl_xor(c2, o, 64);
            }
            const int4 b = tileOff[t];
            p = make_int4(b.x + c0, b.y + c1, b.z + c2, 0);
        }
        if (lane == 0) {
            starts[i] = p.x;
            starts[(np + 1) + i] = p.y;
            starts[2 * (np + 1) + i] = p.z;
        }
    }
}

// FP64 log-likelihood of every pair: the sum over its sub-problems of the
// last window's total.  A sub-problem run as window jobs (its slots hold the
// windows last first, the first slot carries their number K) chains the
// scales: C_0 = 0, C_k = C_{k-1} + log t_{k-1} - log u_k for a restarted
// window k (u_k = its forward . endV at tb_{k-1}, t_{k-1} = window k-1's
// total, both over the paths that end on tb_{k-1}), C_k = 0 when window k's
// forward starts at the matrix start; log P = log t_{K-1} + C_{K-1}.
__device__ __forceinline__ double log_t(int2 v) {
    return log((double)__int_as_float(v.x)) + (double)v.y * 0.69314718055994530942;
}
__global__ __launch_bounds__(256) void pair_log_likelihood(const int32_t *pairSlot, int64_t np,
                                                           const int32_t *slotJob, const int32_t *slotK,
                                                           const DevJob *jobs, const int2 *logTotal,
                                                           const int2 *chain, const int32_t *error, double *out) {
    if (*error) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < np; i += (int64_t)gridDim.x * blockDim.x) {
        double ll = 0.0;
        for (int32_t s = pairSlot[i]; s < pairSlot[i + 1];) {
            const int32_t K = max(slotK[s], 1);
            if (slotJob[s] < 0) {
                s += K;
                continue;
            }
            double C = 0.0, lt = 0.0;
            for (int32_t k = 0; k < K; k++) {
                const int32_t q = slotJob[s + K - 1 - k];
                const double ltk = log_t(logTotal[q]);
                if (k > 0) C = (jobs[q].flags & kJobRestart) ? C + lt - log_t(chain[q]) : 0.0;
                lt = ltk;
            }
            ll += lt + C;
            s += K;
        }
        out[i] = ll;
    }
}

// Sum of the pairs' log-likelihoods in a fixed order (one workgroup).
__global__ __launch_bounds__(1024) void batch_log_likelihood(const double *pairLL, int64_t np,
                                                             const int32_t *error, double *out) {
    __shared__ double sh[1024];
    double v = 0.0;
    if (!*error)
        for (int64_t i = threadIdx.x; i < np; i += 1024) v += pairLL[i];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

}  // namespace mpk
