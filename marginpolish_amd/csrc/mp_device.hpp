// mp_device.hpp -- device-side layout shared by the pair-HMM kernels
// (gfx950): job/window descriptors, the kernel parameter block, the symbol
// encoding that turns an emission into four LDS reads, band entry packing,
// and wavefront helpers (DPP shifts and row reductions).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "mp_pairhmm.h"

namespace mpk {

// ---------------------------------------------------------------------------
// Emission table (floats, copied to LDS).  Both strands share the run-length
// table: the strand only changes which of its four base rows a reference
// base selects (polisher.c:1756-1765: row = strand ? base : 3 - base, N -> A).
//   [0] = 0, GX[strand][5] at 1, GY[strand][5] at 11, E[strand][5][5] at 21,
//   [71] = 0, RL[4][51][51] at 72
// ---------------------------------------------------------------------------
constexpr int kTabGX = 1, kTabGY = 11, kTabE = 21, kTabZeroE = 71, kTabRL = 72;
constexpr int kRL = 51;                             // run lengths 0..50
constexpr int kTabFloats = kTabRL + 4 * kRL * kRL;  // 10476
// Every byte offset an encoded symbol pair can form stays below this bound,
// so the LDS allocation is padded to it (no read can leave the allocation).
constexpr int kTabReach = kTabRL * 4 + 65536;
constexpr uint32_t kRLLastByte = (4 * kRL * kRL - 1) * 4;

// Symbol encoding (uint32 per sequence position, built on the host per pair
// and strand).  Reference side (x):
//   bits  0-15  byte offset of the run-length row: (row * 51 + rl_x) * 51 * 4
//   bits 16-23  byte offset of the match row:       (strand * 25 + base_x * 5) * 4
//   bits 24-31  byte offset of the gap-X entry in the table
// Read side (y):
//   bits  0-15  rl_y * 4
//   bits 16-23  base_y * 4
//   bits 24-31  byte offset of the gap-Y entry in the table
// so x + y carries the run-length and match-table offsets in one add.  The
// zero symbol, read for both x and y by cells outside the band, selects the
// table's two zero entries: all three emissions are 0, so such cells send 0.
__host__ __device__ inline uint32_t enc_x(uint32_t sym, int strand) {
    const uint32_t base = (sym & 255u) < 4u ? (sym & 255u) : 4u;
    const uint32_t rl = (sym >> 8) < 50u ? (sym >> 8) : 50u;
    const uint32_t row = (base & 3u) ^ (strand ? 0u : 3u);
    return ((row * kRL + rl) * kRL * 4u) | (((uint32_t)strand * 25u + base * 5u) * 4u << 16) |
           (((uint32_t)kTabGX + (uint32_t)strand * 5u + base) * 4u << 24);
}
__host__ __device__ inline uint32_t enc_y(uint32_t sym, int strand) {
    const uint32_t base = (sym & 255u) < 4u ? (sym & 255u) : 4u;
    const uint32_t rl = (sym >> 8) < 50u ? (sym >> 8) : 50u;
    return (rl * 4u) | ((base * 4u) << 16) | ((((uint32_t)kTabGY + (uint32_t)strand * 5u + base) * 4u) << 24);
}
constexpr uint32_t kZeroSym = (uint32_t)((kTabZeroE - kTabE) * 4 / 2) << 16;

// A batch's symbol pool goes to the device as 16-bit raw entries, encoded there:
// base min(sym & 255, 4) (bits 0-2), run length min(sym >> 8, 50) (bits 3-8),
// side (bit 9: 0 reference x, 1 read y), strand (bit 10); kRawZero marks the
// zero-emission entry read by cells outside the band.
constexpr uint16_t kRawZero = 0xffff;
__host__ __device__ inline uint16_t raw_sym(uint32_t sym, int y, int strand) {
    const uint32_t base = (sym & 255u) < 4u ? (sym & 255u) : 4u;
    const uint32_t rl = (sym >> 8) < 50u ? (sym >> 8) : 50u;
    return (uint16_t)(base | (rl << 3) | ((uint32_t)y << 9) | ((uint32_t)strand << 10));
}
__host__ __device__ inline uint32_t enc_raw(uint16_t v) {
    if (v == kRawZero) return kZeroSym;
    const uint32_t s = (v & 7u) | (((uint32_t)v >> 3 & 63u) << 8);
    const int strand = (v >> 10) & 1;
    return ((v >> 9) & 1) ? enc_y(s, strand) : enc_x(s, strand);
}

struct Emis {
    float eX, eY, eM;
};

__device__ __forceinline__ Emis emissions(const float *tab, uint32_t sx, uint32_t sy) {
    const char *t = reinterpret_cast<const char *>(tab);
    const uint32_t s = sx + sy;
    // The clamp keeps cells outside the band (which may read any pool entry)
    // inside the table; in-band cells never reach it.
    const float rl = *reinterpret_cast<const float *>(t + kTabRL * 4 + min(s & 0xffffu, kRLLastByte));
    const float e = *reinterpret_cast<const float *>(t + kTabE * 4 + ((s >> 16) & 0xffu));
    const float ex = *reinterpret_cast<const float *>(t + (sx >> 24));
    const float ey = *reinterpret_cast<const float *>(t + (sy >> 24));
    return Emis{ex, ey, e * rl};
}

// ---------------------------------------------------------------------------
// Band entries, one int32 per anti-diagonal d: bits 31-11 xL = x of the
// band's first cell (signed; xmyL = 2 xL - d), bits 10-0 width - 1 (wide
// jobs).  Narrow jobs (width <= 64) use bits 7-0 for width - 1, bit 8 = "this
// 8-diagonal cell needs the general neighbour exchange going forward",
// bit 9 = the same going backward.  Cell k of diagonal d is (xL + k, d - xL - k).
// ---------------------------------------------------------------------------
constexpr int kBandShift = 11;
constexpr int kSlowF = 1 << 8, kSlowB = 1 << 9;
constexpr int kMaxDiagonals = 1 << 20;  // per job (record coordinates use 20 bits)
constexpr int kEmptyBand = INT32_MIN;   // entry of a missing anti-diagonal: no cell matches it
__host__ __device__ inline int band_xL(int b) { return b >> kBandShift; }
__host__ __device__ inline int band_wm1_narrow(int b) { return b & 255; }
__host__ __device__ inline int band_wm1_wide(int b) { return b & 2047; }

// Posterior record: (pM, pX, pY) as float bits and (d << 11) | cell index.
__host__ __device__ inline uint32_t rec_code(int d, int cell) { return ((uint32_t)d << 11) | (uint32_t)cell; }

constexpr int kNumClasses = 3;  // rows of 16, 32, 64 lanes
constexpr int kK = 8;           // anti-diagonals per macro step (grid cell)
constexpr int kSymPad = 128;    // padding entries around the symbol pool
constexpr int kWideMaxC = 32;   // wide kernel: up to 64 * 32 cells per anti-diagonal
constexpr int kMaxPairs = 8;    // forward/backward wavefront pairs per workgroup

struct DevJob {
    int64_t symX, symY, bandOff, outOff;
    int32_t lX, lY, winOff, nWin;
    int32_t cap[2];  // record capacity: forward side, backward side
    int32_t flags;   // bit0 strand, bit1 ragged left, bit2 ragged right, bit3 restarted forward (window job)
    int32_t fstart;  // anti-diagonal the forward sweep starts at (a multiple of kK; 0 = the matrix start)
};
constexpr int kJobRestart = 8;

// Window jobs: one traceback window of a sub-problem with several windows
// runs as a job of its own.  Its forward sweep restarts kWarmup anti-diagonals
// before the window's first posterior anti-diagonal from a flat vector (every
// state of every band cell 1); a window's posteriors F.B / total do not depend
// on the scale of F, and the direction of F forgets its start within a few
// dozen anti-diagonals of an anchored band (measured: relative direction error
// below 1e-12 after 128 anti-diagonals on 100 kb pairs), so 512 leaves a wide
// margin.  The pair's log-likelihood is chained through the traceback
// anti-diagonals: window k's forward dotted with the end vector at tb_{k-1}
// (chain) against window k-1's total.
constexpr int kWarmup = 512;

// One traceback window (pairwiseAligner.c:715-783): forward to tb, backward
// from tb to lo+1, posteriors on (lo, from]; cx = crossing diagonal of the
// concurrent forward/backward sweeps.
struct DevWin {
    int32_t tb, lo, from, cx;
};

struct KParams {
    const DevJob *jobs;
    const uint32_t *sym;
    const int *band;
    const DevWin *win;
    int4 *out;           // posterior records
    int32_t *counts;     // per job: records (forward side, backward side)
    int32_t *winEnd;     // per window: cumulative records at window end (forward side, backward side)
    int2 *logTotal;      // per job: total likelihood as (FP32 t, exponent S): log = ln t + S ln 2
    int2 *chain;         // per restarted window job: its forward . endV at tb of the previous window (t, S)
    float *ring;         // forward/backward values kept for the other sweep
    int32_t *ringExp;    // per 8-diagonal cell exponents of those values
    const float *tables;
    int32_t *queue;      // kNumClasses + 1 counters (the last for wide jobs)
    const int32_t *jobMap;  // claimed positions -> job indices (a partial re-run), or null: position = job
    double *expect;      // per job: 9 expected transition counts + likelihood
    double *emitCount;   // per batch: expected emission counts by emission-table entry (kTabFloats)
    int32_t *error;      // set when a wavefront gives up waiting for its partner (never in a correct run)
    float *ringWide;
    int32_t ringWideLen, wideC, wideBegin, wideEnd;
    int32_t wideWaves;   // ring slices of the 3-state wide kernel (one per wave)
    // 3-state wide kernel: per wave a ring of ringWideCells cells (a row per anti-diagonal holding
    // its band's cells only) and ringWideLen row offsets
    int64_t ringWideCells;
    int32_t *ringWideIdx;
    int32_t jobBegin[kNumClasses], jobEnd[kNumClasses];
    int32_t ringLen;     // anti-diagonals per ring (multiple of kK)
    int32_t tbd;         // traceBackDiagonals: a window's lo is tb of the previous window - (tbd + 1)
    int32_t withIndels;
    int32_t mode;
    float threshold;
    float T[9];          // probability transitions, [from * 3 + to], M = 0, X = 1, Y = 2
    float startV[2][3];  // [ragged][state]
    float endV[2][3];
    // pairhmm_wide5_kernel: the five-state machine (M, short X, short Y, long X,
    // long Y), or a 3-state machine with the long gaps unreachable (FP64 batches)
    int32_t nStates;     // 5 when the five-state kernel runs the batch, else 3
    double T5[25];       // [from * 5 + to]
    double startV5[2][5];
    double endV5[2][5];
};

// ---------------------------------------------------------------------------
// Buffer resources.  A buffer store whose offset (raw) or index (structured)
// lies beyond the resource's extent is dropped by the address unit: a masked
// store without an EXEC branch, so the steps around it stay one scheduling
// region.  Word 3 of a gfx9 resource: 32-bit data format, no swizzle.
// ---------------------------------------------------------------------------
constexpr int kBufWord3 = 0x00020000;
constexpr int kCachePolicyNT = 2;           // gfx94x/gfx950 cache-policy bit "nt"
constexpr uint32_t kBufOff = 0xfffffff0u;  // a raw offset beyond every extent used here

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// (non-temporal: ring values are read back by the partner sweep much later)
__device__ __forceinline__ void buf_store4(__amdgpu_buffer_rsrc_t r, uint32_t off, float a, float b, float c,
                                           float d) {
    const u32x4 v = {__float_as_uint(a), __float_as_uint(b), __float_as_uint(c), __float_as_uint(d)};
    __builtin_amdgcn_raw_buffer_store_b128(v, r, (int)off, 0, kCachePolicyNT);
}

// ---------------------------------------------------------------------------
// Wavefront helpers
// ---------------------------------------------------------------------------
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

template <int CTRL, bool BOUND = true>
__device__ __forceinline__ float dpp(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, BOUND));
}
template <int CTRL>
__device__ __forceinline__ int dppi(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}
constexpr int kRowShr1 = 0x111, kRowShl1 = 0x101, kWaveShr1 = 0x138, kWaveShl1 = 0x130;
constexpr int kRowRor1 = 0x121, kRowRor15 = 0x12F;  // lane i <- lane i-1 / i+1 (mod 16)
constexpr int kQuadSwap1 = 0xB1, kQuadSwap2 = 0x4E, kRowHalfMirror = 0x141, kRowMirror = 0x140;
constexpr int kRowBcast0 = 0x150, kRowBcast15 = 0x15F;

// Lane i <- lane i-1 / i+1 inside rows of SEG lanes; 0 where the source lane
// is outside the row.
template <int SEG>
__device__ __forceinline__ float row_shr(float v, int k) {
    if constexpr (SEG == 16) {
        return dpp<kRowShr1>(v);
    } else if constexpr (SEG == 64) {
        return dpp<kWaveShr1>(v);
    } else {
        const float r = dpp<kWaveShr1>(v);
        return k == 0 ? 0.f : r;
    }
}
template <int SEG>
__device__ __forceinline__ float row_shl(float v, int k) {
    if constexpr (SEG == 16) {
        return dpp<kRowShl1>(v);
    } else if constexpr (SEG == 64) {
        return dpp<kWaveShl1>(v);
    } else {
        const float r = dpp<kWaveShl1>(v);
        return k == SEG - 1 ? 0.f : r;
    }
}

// Reductions over a row of SEG lanes; every lane of the row receives the
// same, bitwise identical result.
template <int SEG>
__device__ __forceinline__ float rmax(float v) {
    v = fmaxf(v, dpp<kQuadSwap1, false>(v));
    v = fmaxf(v, dpp<kQuadSwap2, false>(v));
    v = fmaxf(v, dpp<kRowHalfMirror, false>(v));
    v = fmaxf(v, dpp<kRowMirror, false>(v));
    if constexpr (SEG >= 32) v = fmaxf(v, __shfl_xor(v, 16, 64));
    if constexpr (SEG >= 64) v = fmaxf(v, __shfl_xor(v, 32, 64));
    return v;
}
// Row maximum of non-negative floats through their bit patterns (same order).
template <int SEG>
__device__ __forceinline__ float rmaxpos(float f) {
    uint32_t v = __float_as_uint(f);
    v = max(v, (uint32_t)dppi<kQuadSwap1>((int)v));
    v = max(v, (uint32_t)dppi<kQuadSwap2>((int)v));
    v = max(v, (uint32_t)dppi<kRowHalfMirror>((int)v));
    v = max(v, (uint32_t)dppi<kRowMirror>((int)v));
    if constexpr (SEG >= 32) v = max(v, (uint32_t)__shfl_xor((int)v, 16, 64));
    if constexpr (SEG >= 64) v = max(v, (uint32_t)__shfl_xor((int)v, 32, 64));
    return __uint_as_float(v);
}
template <int SEG>
__device__ __forceinline__ float rsum(float v) {
    v += dpp<kQuadSwap1, false>(v);
    v += dpp<kQuadSwap2, false>(v);
    v += dpp<kRowHalfMirror, false>(v);
    v += dpp<kRowMirror, false>(v);
    if constexpr (SEG >= 32) v += __shfl_xor(v, 16, 64);
    if constexpr (SEG >= 64) v += __shfl_xor(v, 32, 64);
    return v;
}
// Value of lane `src` of the row (row-uniform broadcast).
template <int SEG>
__device__ __forceinline__ int rbcast(int v, int rowBase, int src) {
    if constexpr (SEG == 16) {
        return src == 0 ? dppi<kRowBcast0>(v) : dppi<kRowBcast15>(v);
    } else {
        return __shfl(v, rowBase + src, 64);
    }
}

__device__ __forceinline__ float bperm(float v, int srcLane) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(srcLane << 2, __float_as_int(v)));
}

// Neighbour input from another step's register v: this cell reads cell
// `src` of the source anti-diagonal (cells 0..srcWm1); 0 outside it.
__device__ __forceinline__ float gather(float v, int src, int srcWm1, bool act, int rowBase, int k) {
    const bool ok = act && src >= 0 && src <= srcWm1;
    const float r = bperm(v, rowBase + (ok ? src : k));
    return ok ? r : 0.f;
}

// Consecutive idle iterations (every row of the wavefront waiting, each with
// an s_sleep) after which a wavefront reports an error and leaves: about two
// seconds, far beyond any legitimate wait for the partner wavefront.
constexpr int kSpinLimit = 1 << 24;

// s_waitcnt vmcnt(0): every vector memory operation of this wave has completed.
__device__ __forceinline__ void wait_vm() { __builtin_amdgcn_s_waitcnt(0x0F70); }

}  // namespace mpk
