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
// 1. `slot_record_scan`: record count per slot (the batch's record total);
//    flags an overflowing record region.
// 2. `tile_counts` + `side_tile_scan` (and `window_counts` for jobs of several
//    windows): the number of records of each list per tile and per segment
//    (a window's records on one side).
// 3. `slot_list_scan`: the counts scanned in slot order -> each slot's and
//    each segment's first position in the three lists.
// 4. `emit_tiles` / `emit_windows`: every record's (weight, x, y) tuples
//    written at their list positions; `ref_pair_starts` (mp_ref.hpp): per-pair
//    list offsets.
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

// A list entry on the device: with pk > 0 one 64-bit word, weight (below 2^24)
// in bits 0-23, x + 1 in the next pk bits, y + 1 above (coordinates start at
// -1; the batch chose pk from its longest reference and read so that both
// fit); with pk == 0 a ListTuple.
__device__ __forceinline__ void put_tuple(ListTuple *out, int64_t pos, int pk, int w, int x, int y) {
    if (pk)
        reinterpret_cast<uint64_t *>(out)[pos] = (uint64_t)(uint32_t)w | ((uint64_t)(uint32_t)(x + 1) << 24) |
                                                 ((uint64_t)(uint32_t)(y + 1) << (24 + pk));
    else
        *reinterpret_cast<int4 *>(out + pos) = make_int4(w, w >> 31, x, y);
}

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

// Records of one job side form one segment per window (the side's records of
// that window, [winEnd of the previous window, winEnd of this one)).  The
// reference's order inside a window puts the backward segment first, then the
// forward one; a forward segment (anti-diagonals ascending) is emitted in
// storage order with each anti-diagonal's run of cells reversed by x, a
// backward segment (anti-diagonals descending) fully reversed, again with each
// run by x descending.  So a record's place in list l inside its segment is
// the number of list-l records before its run (forward: in storage order
// before the run; backward: after it) plus those of its run with a larger
// cell.
//
// Sides of one window (every side in window-job batches) are cut into tiles
// of kBigTile records, one workgroup each: `tile_counts` counts each tile's list
// records, `side_tile_scan` scans them along the side (one thread per side)
// and `emit_tiles` places every record's tuples.  Sides of several windows
// (exact-chain batches) are counted and emitted per segment by one workgroup
// per side walking its tiles (`window_counts`, `emit_windows`).
// `slot_list_scan` turns the segment counts into every slot's and segment's
// first list positions.
constexpr int kTile = 256;        // records per tile of emit_windows
constexpr int kHalo = 64;
constexpr int kPer = 4;           // records per thread of tile_counts / emit_tiles
constexpr int kBigTile = kTile * kPer;

// Bit l set: the record joins list l (pairwiseAligner.c:586-664 with the
// x, y > 0 conditions of the three lists).
__device__ __forceinline__ int record_flags(int4 r, const int *bandJ, float thr, int withIndels) {
    const uint32_t c = (uint32_t)r.w;
    const int d = (int)(c >> 11);
    const int x = band_xL(bandJ[d]) + (int)(c & 2047u), y = d - x;
    const float pM = __int_as_float(r.x), pX = __int_as_float(r.y), pY = __int_as_float(r.z);
    return ((pM >= thr) & (x > 0) & (y > 0)) | (((withIndels != 0) & (pX >= thr) & (x > 0)) << 1) |
           (((withIndels != 0) & (pY >= thr) & (y > 0)) << 2);
}

__device__ __forceinline__ int record_flags_x(int4 r, const int *bandJ, float thr, int withIndels, int *xOut) {
    const uint32_t c = (uint32_t)r.w;
    const int d = (int)(c >> 11);
    const int x = band_xL(bandJ[d]) + (int)(c & 2047u), y = d - x;
    const float pM = __int_as_float(r.x), pX = __int_as_float(r.y), pY = __int_as_float(r.z);
    *xOut = x;
    return ((pM >= thr) & (x > 0) & (y > 0)) | (((withIndels != 0) & (pX >= thr) & (x > 0)) << 1) |
           (((withIndels != 0) & (pY >= thr) & (y > 0)) << 2);
}

// The segment [beg, end) of window w on one side of a job.
__device__ __forceinline__ void segment_range(const int32_t *we, int w, int side, int n, int *beg, int *end) {
    *end = min(we[2 * w + side], n);
    *beg = min(w ? we[2 * (w - 1) + side] : 0, *end);
}

// Lanes below this one with their bit set in m.
__device__ __forceinline__ int prefix_in_wave(uint64_t m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Three list counts of at most 1023 in 10-bit fields.
__device__ __forceinline__ int pack3(int f) { return (f & 1) | (((f >> 1) & 1) << 10) | ((f >> 2) << 20); }
__device__ __forceinline__ int field3(int v, int l) { return (v >> (10 * l)) & 1023; }

// Tuples of record t (flags myF) at their list positions.  pt[l] = list-l
// records of the segment in storage order before t; T = the segment's list
// counts; base = the segment's first list positions; codeAt / flagsAt read a
// record's code and flags inside [beg, end).
template <class CodeAt, class FlagsAt>
__device__ __forceinline__ void emit_record(int t, int4 me, int x, int myF, int beg, int end, int side, const int pt[3],
                                            int4 T, const int64_t base[3], const JobLists &J, const int *bandJ,
                                            CodeAt codeAt, FlagsAt flagsAt, int4 cap, int pk, ListTuple *out0,
                                            ListTuple *out1, ListTuple *out2) {
    const uint32_t d = (uint32_t)me.w >> 11, cell = (uint32_t)me.w & 2047u;
    // over the run of this anti-diagonal: list records before / after this one in storage,
    // and those with a larger cell (16-bit fields; a run is at most 2048 records)
    uint64_t before = 0, after = 0, larger = 0;
    for (int i = t - 1; i >= beg; i--) {
        const uint32_t c = codeAt(i);
        if ((c >> 11) != d) break;
        const int f = flagsAt(i);
        const uint64_t fp = (uint64_t)(f & 1) | ((uint64_t)((f >> 1) & 1) << 16) | ((uint64_t)(f >> 2) << 32);
        before += fp;
        if ((c & 2047u) > cell) larger += fp;
    }
    for (int i = t + 1; i < end; i++) {
        const uint32_t c = codeAt(i);
        if ((c >> 11) != d) break;
        const int f = flagsAt(i);
        const uint64_t fp = (uint64_t)(f & 1) | ((uint64_t)((f >> 1) & 1) << 16) | ((uint64_t)(f >> 2) << 32);
        after += fp;
        if ((c & 2047u) > cell) larger += fp;
    }
    const int y = (int)d - x;
    const int segTot[3] = {T.x, T.y, T.z};
    const float p[3] = {__int_as_float(me.x), __int_as_float(me.y), __int_as_float(me.z)};
    ListTuple *out[3] = {out0, out1, out2};
    const int capL[3] = {cap.x, cap.y, cap.z};
#pragma unroll
    for (int l = 0; l < 3; l++) {
        if (!((myF >> l) & 1)) continue;
        const int bf = (int)((before >> (16 * l)) & 0xffff), af = (int)((after >> (16 * l)) & 0xffff);
        const int lg = (int)((larger >> (16 * l)) & 0xffff);
        // forward: list records before the run, then the run by cell descending;
        // backward: list records after the run, then the run by cell descending
        const int inSeg = side ? segTot[l] - (pt[l] + 1 + af) + lg : pt[l] - bf + lg;
        const int64_t pos = base[l] + inSeg;
        if (pos < capL[l]) put_tuple(out[l], pos, pk, (int)floorf(fminf(p[l], 1.f) * 1e7f), x + J.xAdd, y + J.yAdd);
    }
}

// Tile k of the one-window sides: tileMeta[k] = (side 2 job + s, first record,
// the side's winEnd / segment index 2 window + s, 0).  A tile's kBigTile
// records are kPer rounds of kTile, record t0 + i kTile + threadIdx.x in round
// i (coalesced), so chunk 4 i + wave holds records in storage order.
// List counts of every tile of the one-window sides (one workgroup per tile).
__global__ __launch_bounds__(kTile) void tile_counts(const int4 *rec, const JobLists *jl, const int32_t *counts,
                                                     const int32_t *winEnd, const int *band, const int4 *tileMeta,
                                                     int64_t ntiles, const int32_t *skip, const int32_t *error,
                                                     float thr, int withIndels, int4 *tileCnt) {
    __shared__ int sh[3][kTile / 64];
    if (*skip || *error) return;  // (grid-uniform)
    for (int64_t k = blockIdx.x; k < ntiles; k += gridDim.x) {
        const int4 m = tileMeta[k];
        const JobLists J = jl[m.x >> 1];
        const int end = min(winEnd[m.z], counts[m.x]);
        if (m.y >= end) continue;  // (workgroup-uniform) past the side's records
        const int4 *r = rec + ((m.x & 1) ? J.outB : J.outF);
        const int *bandJ = band + J.bandOff;
        int4 v[kPer];
#pragma unroll
        for (int i = 0; i < kPer; i++) {
            const int j = i * kTile + (int)threadIdx.x;
            v[i] = m.y + j < end ? r[m.y + j] : make_int4(0, 0, 0, -1);
        }
        int c[3] = {0, 0, 0};
#pragma unroll
        for (int i = 0; i < kPer; i++) {
            const int t = m.y + i * kTile + (int)threadIdx.x;
            const int f = t < end ? record_flags(v[i], bandJ, thr, withIndels) : 0;
            c[0] += __popcll(__ballot(f & 1));
            c[1] += __popcll(__ballot(f & 2));
            c[2] += __popcll(__ballot(f & 4));
        }
        __syncthreads();  // the previous tile's sums are read
        if ((threadIdx.x & 63) == 0)
            for (int l = 0; l < 3; l++) sh[l][threadIdx.x >> 6] = c[l];
        __syncthreads();
        if (threadIdx.x == 0) {
            int4 tot = make_int4(0, 0, 0, 0);
            for (int w = 0; w < kTile / 64; w++) {
                tot.x += sh[0][w];
                tot.y += sh[1][w];
                tot.z += sh[2][w];
            }
            tileCnt[k] = tot;
        }
    }
}

// Per one-window side (one thread each): exclusive list prefixes of its tiles
// -> tilePre[k], and the side's list counts -> segCnt.
__global__ __launch_bounds__(256) void side_tile_scan(const JobLists *jl, const int32_t *counts,
                                                      const int32_t *winEnd, const int32_t *segTile, int32_t nseg,
                                                      const int4 *tileCnt, const int32_t *skip,
                                                      const int32_t *error, int4 *tilePre, int4 *segCnt) {
    if (*skip || *error) return;
    for (int32_t seg = blockIdx.x * blockDim.x + threadIdx.x; seg < nseg; seg += gridDim.x * blockDim.x) {
        const JobLists J = jl[seg >> 1];
        if (J.nWin != 1) continue;
        const int side = seg & 1;
        const int end = min(winEnd[2 * J.winOff + side], counts[seg]);
        const int64_t k0 = segTile[seg];
        int4 pre = make_int4(0, 0, 0, 0);
        for (int j = 0; j * kBigTile < end; j++) {
            tilePre[k0 + j] = pre;
            const int4 c = tileCnt[k0 + j];
            pre.x += c.x;
            pre.y += c.y;
            pre.z += c.z;
        }
        segCnt[2 * J.winOff + side] = pre;
    }
}

// Per segment of the sides of several windows (one workgroup per side): the
// number of records of each list.
__global__ __launch_bounds__(kTile) void window_counts(const int4 *rec, const JobLists *jl, const int32_t *counts,
                                                       const int32_t *winEnd, const int *band, const int32_t *skip,
                                                       const int32_t *error, float thr, int withIndels, int nseg,
                                                       int4 *segCnt) {
    __shared__ int sh[kTile / 64];
    if (*skip || *error) return;  // (grid-uniform)
    for (int seg = blockIdx.x; seg < nseg; seg += gridDim.x) {
        const JobLists J = jl[seg >> 1];
        if (J.nWin == 1) continue;  // (workgroup-uniform)
        const int side = seg & 1;  // 0 forward, 1 backward
        const int n = counts[seg];
        const int4 *r = rec + (side ? J.outB : J.outF);
        const int32_t *we = winEnd + 2 * J.winOff;
        const int *bandJ = band + J.bandOff;
        for (int w = 0; w < J.nWin; w++) {
            int beg, end;
            segment_range(we, w, side, n, &beg, &end);
            int4 v = make_int4(0, 0, 0, 0);
            for (int t0 = beg; t0 < end; t0 += kTile) {
                const int t = t0 + (int)threadIdx.x;
                int c = t < end ? pack3(record_flags(r[t], bandJ, thr, withIndels)) : 0;
                for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
                __syncthreads();
                if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = c;
                __syncthreads();
                for (int k = 0; k < kTile / 64; k++) {
                    v.x += field3(sh[k], 0);
                    v.y += field3(sh[k], 1);
                    v.z += field3(sh[k], 2);
                }
            }
            if (threadIdx.x == 0) segCnt[2 * (J.winOff + w) + side] = v;
        }
    }
}

// Per slot the number of tuples of each list and their exclusive scan in slot
// order (one workgroup) -> slotBase3[3 s + l]; every segment's first list
// position -> segBase[3 (2 window + side) + l], the windows of a job last
// first and, in a window, the backward segment before the forward one.
__global__ __launch_bounds__(1024) void slot_list_scan(const int32_t *slotJob, int32_t nslots, const JobLists *jl,
                                                       const int4 *segCnt, const int32_t *skip, const int32_t *error,
                                                       int64_t *slotBase3, int64_t *segBase) {
    constexpr int PER = 4;  // slots per thread and pass
    __shared__ int64_t sh[3][16];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int64_t carry[3] = {0, 0, 0};
    const bool bad = *skip || *error;
    for (int32_t s0 = 0; s0 < nslots; s0 += 1024 * PER) {
        int32_t q[PER], winOff[PER], nWin[PER];
        int64_t v[PER][3];
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int32_t s = s0 + PER * (int32_t)threadIdx.x + k;
            q[k] = s < nslots && !bad ? slotJob[s] : -1;
        }
#pragma unroll
        for (int k = 0; k < PER; k++) {
            winOff[k] = q[k] >= 0 ? jl[q[k]].winOff : 0;
            nWin[k] = q[k] >= 0 ? jl[q[k]].nWin : 0;
        }
        int64_t mine[3] = {0, 0, 0};
#pragma unroll
        for (int k = 0; k < PER; k++) {
            v[k][0] = v[k][1] = v[k][2] = 0;
            for (int32_t w = 0; w < nWin[k]; w++) {
                const int4 f = segCnt[2 * (winOff[k] + w)], b = segCnt[2 * (winOff[k] + w) + 1];
                v[k][0] += f.x + b.x;
                v[k][1] += f.y + b.y;
                v[k][2] += f.z + b.z;
            }
            for (int l = 0; l < 3; l++) mine[l] += v[k][l];
        }
        // exclusive scan of the three per-thread sums over the workgroup
        int64_t x[3] = {mine[0], mine[1], mine[2]};
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
            for (int l = 0; l < 3; l++) {
                const int64_t y = __shfl_up(x[l], o, 64);
                if (lane >= o) x[l] += y;
            }
        __syncthreads();  // the previous pass's sums are read
        if (lane == 63)
            for (int l = 0; l < 3; l++) sh[l][wv] = x[l];
        __syncthreads();
        int64_t pre[3];
        for (int l = 0; l < 3; l++) {
            int64_t off = 0, tot = 0;
            for (int i = 0; i < 16; i++) {
                if (i < wv) off += sh[l][i];
                tot += sh[l][i];
            }
            pre[l] = carry[l] + off + x[l] - mine[l];
            carry[l] += tot;
        }
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int32_t s = s0 + PER * (int32_t)threadIdx.x + k;
            if (s < nslots)
                for (int l = 0; l < 3; l++) slotBase3[3 * (int64_t)s + l] = pre[l];
            int64_t p[3] = {pre[0], pre[1], pre[2]};
            for (int32_t w = nWin[k] - 1; w >= 0; w--) {
                const int4 f = segCnt[2 * (winOff[k] + w)], b = segCnt[2 * (winOff[k] + w) + 1];
                int64_t *bw = segBase + 3 * (2 * (int64_t)(winOff[k] + w) + 1), *fw = bw - 3;
                bw[0] = p[0];
                bw[1] = p[1];
                bw[2] = p[2];
                p[0] += b.x;
                p[1] += b.y;
                p[2] += b.z;
                fw[0] = p[0];
                fw[1] = p[1];
                fw[2] = p[2];
                p[0] += f.x;
                p[1] += f.y;
                p[2] += f.z;
            }
            for (int l = 0; l < 3; l++) pre[l] += v[k][l];
        }
    }
    if (threadIdx.x == 0)
        for (int l = 0; l < 3; l++) slotBase3[3 * (int64_t)nslots + l] = carry[l];
}

// Every record of the one-window sides placed in the lists (one workgroup per
// tile, records in rounds as in tile_counts).  The tile's record codes, with
// kHalo records either side, and the exclusive prefix of their list flags
// (21-bit fields) are staged in LDS.  A record's anti-diagonal run holds its
// cells in lane order: ascending, or for rows of x-aligned lanes an ascending
// run rotated once (one descent at w, every cell after w below every cell
// before it).  Then the run's cells larger than the record's are those after
// it up to w, plus for a record after w the whole part before w, so a walk
// over the codes for the run's ends and w and four prefix differences place
// the record.  Runs beyond the staged records or with more than one descent
// take the general count (emit_record).  Positions at or beyond a list's
// capacity are dropped (the host then grows that list and assembles again).
__global__ __launch_bounds__(kTile) void emit_tiles(const int4 *rec, const JobLists *jl, const int32_t *counts,
                                                    const int32_t *winEnd, const int *band, const int4 *tileMeta,
                                                    int64_t ntiles, const int4 *tilePre, const int4 *segCnt,
                                                    const int64_t *segBase, const int32_t *skip, const int32_t *error,
                                                    float thr, int withIndels, int4 cap, int pk, ListTuple *out0,
                                                    ListTuple *out1, ListTuple *out2) {
    constexpr int NS = kBigTile + 2 * kHalo;
    constexpr int NC = NS / 64;  // chunks of staged records: left halo, kPer rounds x waves, right halo
    static_assert(kHalo == 64, "one halo chunk per side");
    __shared__ uint32_t code[NS];
    __shared__ uint64_t ss[NS + 1];  // exclusive prefix of the staged list flags, 21-bit fields
    __shared__ uint64_t cnt[NC];     // per chunk: list counts, then their exclusive prefix
    if (*skip || *error) return;     // (grid-uniform)
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    auto fld = [](uint64_t v, int l) -> int { return (int)((v >> (21 * l)) & 0x1fffff); };
    for (int64_t k = blockIdx.x; k < ntiles; k += gridDim.x) {
        const int4 m = tileMeta[k];
        const int t0 = m.y, side = m.x & 1;
        const int end = min(winEnd[m.z], counts[m.x]);
        const JobLists J = jl[m.x >> 1];
        const int4 P = tilePre[k], T = segCnt[m.z];
        const int64_t base[3] = {segBase[3 * (int64_t)m.z], segBase[3 * (int64_t)m.z + 1],
                                 segBase[3 * (int64_t)m.z + 2]};
        if (t0 >= end) continue;  // (workgroup-uniform)
        const int4 *r = rec + (side ? J.outB : J.outF);
        const int *bandJ = band + J.bandOff;
        // waves 0 and 1 stage the left and right halo chunks
        int4 me[kPer];
        int f[kPer], xs[kPer];
#pragma unroll
        for (int i = 0; i < kPer; i++) {
            const int t = t0 + i * kTile + (int)threadIdx.x;
            me[i] = t < end ? r[t] : make_int4(0, 0, 0, -1);
        }
        const int hg = wv == 0 ? t0 - kHalo + lane : t0 + kBigTile + lane;
        const bool hv = wv < 2 && hg >= 0 && hg < end;
        const int4 hr = hv ? r[hg] : make_int4(0, 0, 0, -1);
        uint64_t bal[kPer][3];
#pragma unroll
        for (int i = 0; i < kPer; i++) {
            const int t = t0 + i * kTile + (int)threadIdx.x;
            f[i] = t < end ? record_flags_x(me[i], bandJ, thr, withIndels, &xs[i]) : 0;
            for (int l = 0; l < 3; l++) bal[i][l] = __ballot((f[i] >> l) & 1);
        }
        const int hf = hv ? record_flags(hr, bandJ, thr, withIndels) : 0;
        uint64_t hb[3];
        for (int l = 0; l < 3; l++) hb[l] = __ballot((hf >> l) & 1);
        __syncthreads();  // the previous tile's LDS reads are done
#pragma unroll
        for (int i = 0; i < kPer; i++) {
            code[kHalo + i * kTile + threadIdx.x] = (uint32_t)me[i].w;
            if (lane == 0)
                cnt[1 + i * (kTile / 64) + wv] = (uint64_t)__popcll(bal[i][0]) |
                                                 ((uint64_t)__popcll(bal[i][1]) << 21) |
                                                 ((uint64_t)__popcll(bal[i][2]) << 42);
        }
        if (wv < 2) {
            code[wv == 0 ? lane : kHalo + kBigTile + lane] = (uint32_t)hr.w;
            if (lane == 0)
                cnt[wv == 0 ? 0 : NC - 1] = (uint64_t)__popcll(hb[0]) | ((uint64_t)__popcll(hb[1]) << 21) |
                                            ((uint64_t)__popcll(hb[2]) << 42);
        }
        __syncthreads();
        if (threadIdx.x == 0) {  // exclusive prefix over the chunks
            uint64_t run = 0;
            for (int c = 0; c < NC; c++) {
                const uint64_t v = cnt[c];
                cnt[c] = run;
                run += v;
            }
            ss[NS] = run;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kPer; i++) {
            const int c = 1 + i * (kTile / 64) + wv;
            ss[64 * c + lane] = cnt[c] + (uint64_t)prefix_in_wave(bal[i][0]) +
                                ((uint64_t)prefix_in_wave(bal[i][1]) << 21) +
                                ((uint64_t)prefix_in_wave(bal[i][2]) << 42);
        }
        if (wv < 2) {
            const int c = wv == 0 ? 0 : NC - 1;
            ss[64 * c + lane] = cnt[c] + (uint64_t)prefix_in_wave(hb[0]) + ((uint64_t)prefix_in_wave(hb[1]) << 21) +
                                ((uint64_t)prefix_in_wave(hb[2]) << 42);
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kPer; i++) {
            if (!f[i]) continue;
            const int t = t0 + i * kTile + (int)threadIdx.x, lt = kHalo + i * kTile + (int)threadIdx.x;
            const uint32_t cm = (uint32_t)me[i].w, dd = cm >> 11;
            // list records of the tile before this one (the chunk prefixes count the left halo first)
            const uint64_t cp = cnt[1 + i * (kTile / 64) + wv] - cnt[1];
            const int pt[3] = {P.x + fld(cp, 0) + prefix_in_wave(bal[i][0]), P.y + fld(cp, 1) + prefix_in_wave(bal[i][1]),
                               P.z + fld(cp, 2) + prefix_in_wave(bal[i][2])};
            // run start lu, end lv (staged indexes) and its descents
            int lu = lt, wB = -1, nDesc = 0;
            uint32_t cur = cm;
            while (lu > 0) {
                const uint32_t c = code[lu - 1];
                if ((c >> 11) != dd) break;
                if ((cur & 2047u) < (c & 2047u)) {
                    nDesc++;
                    if (wB < 0) wB = lu;
                }
                cur = c;
                lu--;
            }
            const uint32_t firstCell = cur & 2047u;
            int lv = lt + 1, wA = -1;
            cur = cm;
            while (lv < NS) {
                const uint32_t c = code[lv];
                if ((c >> 11) != dd) break;
                if ((c & 2047u) < (cur & 2047u)) {
                    nDesc++;
                    if (wA < 0) wA = lv;
                }
                cur = c;
                lv++;
            }
            // (one descent: a rotation only if the run's last cell is below its first)
            const bool general = nDesc > 1 || (nDesc == 1 && (cur & 2047u) >= firstCell) ||
                                 (lu == 0 && t0 - kHalo > 0) || (lv == NS && t0 - kHalo + NS < end);
            if (general) {
                auto codeAt = [&](int i2) -> uint32_t {
                    const int li = i2 - t0 + kHalo;
                    return li >= 0 && li < NS ? code[li] : (uint32_t)r[i2].w;
                };
                auto flagsAt = [&](int i2) -> int { return record_flags(r[i2], bandJ, thr, withIndels); };
                emit_record(t, me[i], xs[i], f[i], 0, end, side, pt, T, base, J, bandJ, codeAt, flagsAt, cap, pk, out0,
                            out1, out2);
                continue;
            }
            const uint64_t sU = ss[lu], sT = ss[lt], sT1 = ss[lt + 1], sV = ss[lv];
            const bool inB = wB >= 0;
            const uint64_t sW = ss[inB ? wB : (wA >= 0 ? wA : lv)];
            // larger cells of the run: after t up to w, plus (after w) everything before w
            const uint64_t larger = inB ? (sW - sU) + (sV - sT1) : sW - sT1;
            const int y0 = (int)dd - xs[i];
            const float p[3] = {__int_as_float(me[i].x), __int_as_float(me[i].y), __int_as_float(me[i].z)};
            ListTuple *out[3] = {out0, out1, out2};
            const int capL[3] = {cap.x, cap.y, cap.z};
            const int segTot[3] = {T.x, T.y, T.z};
#pragma unroll
            for (int l = 0; l < 3; l++) {
                if (!((f[i] >> l) & 1)) continue;
                // forward: list records before the run, then the larger cells of the run;
                // backward: list records after the run, then the larger cells of the run
                const int inSeg = side ? segTot[l] - (pt[l] + fld(sV - sT, l)) + fld(larger, l)
                                       : pt[l] - fld(sT - sU, l) + fld(larger, l);
                const int64_t pos = base[l] + inSeg;
                if (pos < capL[l])
                    put_tuple(out[l], pos, pk, (int)floorf(fminf(p[l], 1.f) * 1e7f), xs[i] + J.xAdd, y0 + J.yAdd);
            }
        }
    }
}

// Every record of the sides of several windows placed in the lists (one
// workgroup per side, its segments in tiles of kTile records staged as in
// emit_tiles).
__global__ __launch_bounds__(kTile) void emit_windows(const int4 *rec, const JobLists *jl, const int32_t *counts,
                                                      const int32_t *winEnd, const int *band, const int4 *segCnt,
                                                      const int64_t *segBase, const int32_t *skip,
                                                      const int32_t *error, float thr, int withIndels, int nseg,
                                                      int4 cap, int pk, ListTuple *out0, ListTuple *out1,
                                                      ListTuple *out2) {
    constexpr int NS = kTile + 2 * kHalo;
    __shared__ uint32_t code[NS];
    __shared__ uint8_t fl[NS];
    __shared__ int scanSh[kTile / 64];
    if (*skip || *error) return;  // (grid-uniform)
    for (int seg = blockIdx.x; seg < nseg; seg += gridDim.x) {
        const JobLists J = jl[seg >> 1];
        if (J.nWin == 1) continue;  // (workgroup-uniform)
        const int side = seg & 1;
        const int n = counts[seg];
        const int4 *r = rec + (side ? J.outB : J.outF);
        const int32_t *we = winEnd + 2 * J.winOff;
        const int *bandJ = band + J.bandOff;
        for (int w = 0; w < J.nWin; w++) {
            int beg, end;
            segment_range(we, w, side, n, &beg, &end);
            const int64_t sb = 3 * (2 * (int64_t)(J.winOff + w) + side);
            const int64_t base[3] = {segBase[sb], segBase[sb + 1], segBase[sb + 2]};
            const int4 T = segCnt[2 * (J.winOff + w) + side];
            int carry[3] = {0, 0, 0};
            for (int t0 = beg; t0 < end; t0 += kTile) {
                __syncthreads();  // the previous tile's LDS reads are done
                for (int i = threadIdx.x; i < NS; i += kTile) {
                    const int gi = t0 - kHalo + i;
                    if (gi >= beg && gi < end) {
                        const int4 v = r[gi];
                        code[i] = (uint32_t)v.w;
                        fl[i] = (uint8_t)record_flags(v, bandJ, thr, withIndels);
                    } else {
                        code[i] = 0xffffffffu;
                        fl[i] = 0;
                    }
                }
                __syncthreads();
                const int t = t0 + (int)threadIdx.x;
                const int myF = t < end ? fl[kHalo + threadIdx.x] : 0;
                int tileTot;
                const int ex = block_excl_scan<int, kTile>(pack3(myF), scanSh, &tileTot);
                if (t < end && myF) {
                    const int pt[3] = {carry[0] + field3(ex, 0), carry[1] + field3(ex, 1), carry[2] + field3(ex, 2)};
                    auto codeAt = [&](int i) -> uint32_t {
                        const int li = i - t0 + kHalo;
                        return li >= 0 && li < NS ? code[li] : (uint32_t)r[i].w;
                    };
                    auto flagsAt = [&](int i) -> int {
                        const int li = i - t0 + kHalo;
                        return li >= 0 && li < NS ? (int)fl[li] : record_flags(r[i], bandJ, thr, withIndels);
                    };
                    const int4 me = r[t];
                    const int x = band_xL(bandJ[(uint32_t)me.w >> 11]) + (int)((uint32_t)me.w & 2047u);
                    emit_record(t, me, x, myF, beg, end, side, pt, T, base, J, bandJ, codeAt, flagsAt, cap, pk, out0,
                                out1, out2);
                }
                for (int l = 0; l < 3; l++) carry[l] += field3(tileTot, l);
            }
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
