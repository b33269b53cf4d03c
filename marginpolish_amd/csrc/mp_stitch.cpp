// mp_stitch.cpp -- chunk stitching: removeOverlap for many boundaries as one
// match-only batch, and mergeContigChunks on top of it.
#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "mp_internal.hpp"
#include "mp_stitch.h"

namespace {

inline uint16_t base_code(char c) {  // symbolString_construct: convertNucleotideCharToSymbol (stateMachine.c:25-42)
    switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 4;
    }
}

// One overlap alignment: x = the prefix string's window, y = the suffix string's window.
struct Overlap {
    const char *x;
    int64_t lx, off;  // window length and its start in the prefix string
    const char *y;
    int64_t ly;
    int64_t prefixLen;
};

// The chosen tuple: x relative to the window (the prefix keeps x + window
// start characters), y, and its weight; weight 0 = no tuple.
struct Cut {
    int64_t x, cropStart, weight;
};

// removeOverlap's alignment and choice (polisher.c:1954-2003) for every job, one batch.
mp_status align_overlaps(mp_engine *e, const mp_band_params *p, const std::vector<Overlap> &jobs,
                         std::vector<Cut> &cuts) {
    cuts.assign(jobs.size(), Cut{0, 0, 0});
    std::vector<size_t> live;
    for (size_t k = 0; k < jobs.size(); k++)
        if (jobs[k].lx > 0 && jobs[k].ly > 0) live.push_back(k);  // an empty side aligns nothing
    if (live.empty()) return MP_OK;
    std::vector<int64_t> symOff(live.size() + 1, 0);
    for (size_t q = 0; q < live.size(); q++) symOff[q + 1] = symOff[q] + jobs[live[q]].lx + jobs[live[q]].ly;
    std::vector<uint16_t> sym((size_t)symOff.back());
    std::vector<mp_pair> pairs(live.size());
    for (size_t q = 0; q < live.size(); q++) {
        const Overlap &o = jobs[live[q]];
        uint16_t *sx = sym.data() + symOff[q], *sy = sx + o.lx;
        for (int64_t i = 0; i < o.lx; i++) sx[i] = base_code(o.x[i]);
        for (int64_t i = 0; i < o.ly; i++) sy[i] = base_code(o.y[i]);
        mp_pair &pr = pairs[q];
        pr = mp_pair{};
        pr.sX = sx;
        pr.lX = o.lx;
        pr.sY = sy;
        pr.lY = o.ly;
        pr.strand = 1;
        pr.raggedLeft = 1;
        pr.raggedRight = 1;
    }
    mp_model *m = nullptr;  // stateMachine3_constructNucleotide(threeState), as removeOverlap builds it
    mp_status st = mp_model_default_nucleotide(&m);
    if (st) return st;
    mp_batch *b = nullptr;
    // the reference's arithmetic: the cut is an integer decision on near-tied weights
    st = mp_batch_create(e, m, pairs.data(), (int64_t)pairs.size(), p, MP_MODE_MATCH_ONLY,
                         MP_BATCH_NO_GRAPH | MP_BATCH_REFERENCE_ARITHMETIC, &b);
    mp_model_free(m);
    if (st) return st;
    std::vector<int64_t> counts(3 * pairs.size());
    std::vector<mp_aligned_pair> list;
    if (!(st = mp_batch_run(b)) && !(st = mp_batch_fetch_counts(b, counts.data(), nullptr))) {
        int64_t tot = 0;
        for (size_t q = 0; q < pairs.size(); q++) tot += counts[3 * q];
        list.resize((size_t)std::max<int64_t>(tot, 1));
        st = mp_batch_fetch_list(b, MP_LIST_MATCH, list.data());
    }
    mp_batch_free(b);
    if (st) return st;
    int64_t at = 0;
    for (size_t q = 0; q < pairs.size(); q++) {
        const mp_aligned_pair *best = nullptr;  // first tuple of maximal weight, in list order
        for (int64_t t = at; t < at + counts[3 * q]; t++)
            if (!best || list[t].weight > best->weight) best = &list[t];
        at += counts[3 * q];
        if (best) cuts[live[q]] = Cut{best->x, best->y, best->weight};
    }
    return MP_OK;
}

// The windows removeOverlap aligns (polisher.c:1941-1953).
Overlap window(const char *prefix, int64_t prefixLen, const char *suffix, int64_t suffixLen, int64_t approx) {
    const int64_t i = std::max<int64_t>(prefixLen - approx, 0);
    const int64_t j = std::min(approx, suffixLen);
    return Overlap{prefix + i, prefixLen - i, i, suffix, j, prefixLen};
}

bool bad_string(const char *s, int64_t n) { return n < 0 || (n > 0 && !s); }

}  // namespace

extern "C" mp_status mp_remove_overlaps(mp_engine *e, const mp_band_params *p, int64_t n, const char *const *prefix,
                                        const int64_t *prefixLen, const char *const *suffix,
                                        const int64_t *suffixLen, const int64_t *approxOverlap,
                                        int64_t *prefixCropEnd, int64_t *suffixCropStart, int64_t *overlapWeight) {
    if (!e || !p || n < 0 || (n > 0 && (!prefix || !prefixLen || !suffix || !suffixLen || !approxOverlap ||
                                        !prefixCropEnd || !suffixCropStart))) {
        mp::set_error("mp_remove_overlaps: null or negative argument");
        return MP_EINVAL;
    }
    try {
        std::vector<Overlap> jobs((size_t)n);
        for (int64_t k = 0; k < n; k++) {
            if (bad_string(prefix[k], prefixLen[k]) || bad_string(suffix[k], suffixLen[k]) || approxOverlap[k] < 0) {
                mp::set_error("mp_remove_overlaps: bad string or overlap at pair " + std::to_string(k));
                return MP_EINVAL;
            }
            jobs[k] = window(prefix[k], prefixLen[k], suffix[k], suffixLen[k], approxOverlap[k]);
        }
        std::vector<Cut> cuts;
        mp_status st = align_overlaps(e, p, jobs, cuts);
        if (st) return st;
        for (int64_t k = 0; k < n; k++) {
            const bool found = cuts[k].weight > 0;  // tuples are kept at p >= threshold > 0
            prefixCropEnd[k] = found ? cuts[k].x + jobs[k].off : prefixLen[k];
            suffixCropStart[k] = found ? cuts[k].cropStart : 0;
            if (overlapWeight) overlapWeight[k] = cuts[k].weight;
        }
    } catch (const std::bad_alloc &) {
        mp::set_error("mp_remove_overlaps: out of host memory");
        return MP_ENOMEM;
    }
    return MP_OK;
}

extern "C" mp_status mp_merge_chunks(mp_engine *e, const mp_band_params *p, int64_t nChunks,
                                     const char *const *chunks, const int64_t *chunkLen, int64_t overlap,
                                     const char *missingSpacer, char *out, int64_t outCap, int64_t *outLen) {
    if (!e || !p || !outLen || nChunks < 0 || overlap < 0 || (nChunks > 0 && (!chunks || !chunkLen))) {
        mp::set_error("mp_merge_chunks: null or negative argument");
        return MP_EINVAL;
    }
    for (int64_t c = 0; c < nChunks; c++)
        if (bad_string(chunks[c], chunkLen[c])) {
            mp::set_error("mp_merge_chunks: bad chunk " + std::to_string(c));
            return MP_EINVAL;
        }
    static const char kNs[] = "NNNNNNNNNN";
    const char *spacer = missingSpacer ? missingSpacer : "";
    const int64_t spacerLen = (int64_t)std::strlen(spacer);
    // a polished string as the merge holds it: a view [s, s + n)
    struct View {
        const char *s;
        int64_t n;
    };
    try {
        // 1. every boundary in one batch, each against the predecessor it has
        //    when no earlier crop reaches into its overlap window
        std::vector<Overlap> jobs;
        for (int64_t c = 1; c < nChunks; c++) {
            View prev{chunks[c - 1], chunkLen[c - 1]};
            if (c - 1 > 0 && chunkLen[c - 1] == 0) prev = View{spacer, spacerLen};  // an empty chunk became the spacer
            jobs.push_back(window(prev.s, prev.n, chunks[c], chunkLen[c], overlap));
        }
        std::vector<Cut> cuts;
        mp_status st = align_overlaps(e, p, jobs, cuts);
        if (st) return st;
        // 2. the reference's sequential merge; a boundary whose actual predecessor
        //    differs inside the window is aligned again with the actual string
        std::vector<View> pieces;
        if (nChunks > 0) pieces.push_back(View{chunks[0], chunkLen[0]});
        for (int64_t c = 1; c < nChunks; c++) {
            View &prev = pieces.back();
            const Overlap want = window(prev.s, prev.n, chunks[c], chunkLen[c], overlap);
            Cut cut = cuts[c - 1];
            const Overlap &had = jobs[c - 1];
            if (want.x != had.x || want.lx != had.lx) {
                std::vector<Cut> one;
                if ((st = align_overlaps(e, p, std::vector<Overlap>{want}, one))) return st;
                cut = one[0];
            }
            View cur{chunks[c], chunkLen[c]};
            if (cut.weight > 0) {
                prev.n = cut.x + want.off;  // crop the predecessor's suffix, the chunk's prefix
                cur = View{chunks[c] + cut.cropStart, chunkLen[c] - cut.cropStart};
            } else if (chunkLen[c] == 0) {
                cur = View{spacer, spacerLen};
            } else if (overlap != 0) {
                pieces.push_back(View{kNs, 10});
            }
            pieces.push_back(cur);
        }
        int64_t total = 0;
        for (const View &v : pieces) total += v.n;
        *outLen = total;
        if (!out || outCap < total) {
            mp::set_error("mp_merge_chunks: output buffer too small (need " + std::to_string(total) + ")");
            return MP_EINVAL;
        }
        int64_t at = 0;
        for (const View &v : pieces) {
            if (v.n) std::memcpy(out + at, v.s, (size_t)v.n);
            at += v.n;
        }
    } catch (const std::bad_alloc &) {
        mp::set_error("mp_merge_chunks: out of host memory");
        return MP_ENOMEM;
    }
    return MP_OK;
}
