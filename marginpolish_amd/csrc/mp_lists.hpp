// Posterior lists assembled on the device.
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
// 1. `order_records`: one workgroup per job side places every record at its
//    emission index (pairs in order, a pair's jobs in split order, then the
//    order above) and decodes it: coordinates, three weights, list flags.
// 2. three exclusive scans of the list flags give each tuple's list index.
// 3. `emit_lists` writes the (weight, x, y) tuples; `pair_list_starts` the
//    per-pair list offsets.
#pragma once

#include "mp_device.hpp"

namespace mpk {

struct JobLists {
    int64_t outF, outB;   // record regions: forward side, backward side
    int64_t base;         // emission index of the job's first record
    int64_t bandOff;      // the job's band entries
    int32_t nF, nB;       // records written by each side
    int32_t winOff, nWin;
    int32_t xAdd, yAdd;   // output coordinate shifts (split offset + reference crop - 1, split offset - 1)
};

struct ListTuple {        // the C ABI's mp_aligned_pair
    int64_t weight;
    int32_t x, y;
};

__global__ __launch_bounds__(256) void order_records(const int4 *rec, const JobLists *jl, const int32_t *winEnd,
                                                     const int *band, float thr, int withIndels, int nseg,
                                                     int2 *xy, int4 *wf) {
    for (int seg = blockIdx.x; seg < nseg; seg += gridDim.x) {
        const JobLists J = jl[seg >> 1];
        const int side = seg & 1;  // 0 forward, 1 backward
        const int n = side ? J.nB : J.nF;
        const int4 *r = rec + (side ? J.outB : J.outF);
        const int32_t *we = winEnd + 2 * J.winOff;
        for (int t = threadIdx.x; t < n; t += blockDim.x) {
            int w = 0;
            while (w + 1 < J.nWin && t >= we[2 * w + side]) w++;
            const int fBeg = w ? we[2 * (w - 1)] : 0, fEnd = we[2 * w];
            const int bBeg = w ? we[2 * (w - 1) + 1] : 0, bEnd = we[2 * w + 1];
            const int beg = side ? bBeg : fBeg, end = side ? bEnd : fEnd;
            // windows after this one come first
            const int64_t off = (int64_t)(J.nF - fEnd) + (J.nB - bEnd);
            const int4 me = r[t];
            const uint32_t d = (uint32_t)me.w >> 11, cell = (uint32_t)me.w & 2047u;
            int u = t, v = t + 1;
            while (u > beg && ((uint32_t)r[u - 1].w >> 11) == d) u--;
            while (v < end && ((uint32_t)r[v].w >> 11) == d) v++;
            int rank = 0;
            for (int i = u; i < v; i++) rank += ((uint32_t)r[i].w & 2047u) > cell;
            // backward blocks are stored anti-diagonals descending: walk them from the end
            const int64_t pos = side ? (int64_t)(bEnd - v) + rank : (int64_t)(bEnd - bBeg) + (u - fBeg) + rank;
            const int64_t g = J.base + off + pos;
            const int x = band_xL(band[J.bandOff + d]) + (int)cell, y = (int)d - x;
            const float pM = __int_as_float(me.x), pX = __int_as_float(me.y), pY = __int_as_float(me.z);
            const int flags = ((pM >= thr) & (x > 0) & (y > 0)) | (((withIndels != 0) & (pX >= thr) & (x > 0)) << 1) |
                              (((withIndels != 0) & (pY >= thr) & (y > 0)) << 2);
            xy[g] = make_int2(x + J.xAdd, y + J.yAdd);
            wf[g] = make_int4((int)floorf(fminf(pM, 1.f) * 1e7f), (int)floorf(fminf(pX, 1.f) * 1e7f),
                              (int)floorf(fminf(pY, 1.f) * 1e7f), flags);
        }
    }
}

struct ListBit {
    int l;
    __host__ __device__ int operator()(const int4 &f) const { return (f.w >> l) & 1; }
};

__global__ __launch_bounds__(256) void emit_lists(const int2 *xy, const int4 *wf, const int32_t *pos0,
                                                  const int32_t *pos1, const int32_t *pos2, int64_t n,
                                                  ListTuple *out0, ListTuple *out1, ListTuple *out2) {
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x) {
        const int4 f = wf[g];
        const int2 c = xy[g];
        if (f.w & 1) out0[pos0[g]] = ListTuple{f.x, c.x, c.y};
        if (f.w & 2) out1[pos1[g]] = ListTuple{f.y, c.x, c.y};
        if (f.w & 4) out2[pos2[g]] = ListTuple{f.z, c.x, c.y};
    }
}

// starts[l * (np + 1) + i] = index in list l of pair i's first tuple; entry np = list length
__global__ __launch_bounds__(256) void pair_list_starts(const int64_t *pairStart, int64_t np, int64_t n,
                                                        const int4 *wf, const int32_t *pos0, const int32_t *pos1,
                                                        const int32_t *pos2, int64_t *starts) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= np; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = pairStart[i];
        const int32_t *pos[3] = {pos0, pos1, pos2};
        for (int l = 0; l < 3; l++)
            starts[l * (np + 1) + i] =
                g < n ? pos[l][g] : (n > 0 ? (int64_t)pos[l][n - 1] + ((wf[n - 1].w >> l) & 1) : 0);
    }
}

}  // namespace mpk
