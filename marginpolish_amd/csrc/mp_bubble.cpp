// mp_bubble.cpp -- per-read x per-allele forward likelihoods of many
// bubbles as one forward-only batch (bubbleGraph.c:895-928).
//
// Every (allele, distinct read) pair of every bubble becomes one pair of a
// MP_MODE_FORWARD_ONLY batch (static band of diagonalExpansion, no anchors,
// no ragged ends: computeForwardProbability, pairwiseAligner.c:813-865), so
// the whole bubble graph's likelihoods take one persistent launch instead of
// one call per pair.
#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "mp_bubble.h"
#include "mp_internal.hpp"

namespace {

inline uint16_t base_code(char c) {  // convertNucleotideCharToSymbol (stateMachine.c:25-42)
    switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 4;
    }
}

// rleString_constructSymbolString (stateMachine.c:133-146) into `out`.
void symbols(const mp_rle_string &s, bool rl, uint64_t maxRcEx, uint16_t *out) {
    for (int64_t i = 0; i < s.length; i++) {
        uint64_t r = 0;
        if (rl) r = s.repeatCounts[i] >= maxRcEx ? maxRcEx - 1 : s.repeatCounts[i];
        out[i] = (uint16_t)((r << 8) | base_code(s.bases[i]));
    }
}

bool same_string(const mp_rle_string &a, const mp_rle_string &b) {
    if (a.length != b.length) return false;
    return (a.length == 0) || (std::memcmp(a.bases, b.bases, (size_t)a.length) == 0 &&
                               std::memcmp(a.repeatCounts, b.repeatCounts, sizeof(uint64_t) * a.length) == 0);
}

uint64_t string_hash(const mp_rle_string &s) {
    uint64_t h = 0xcbf29ce484222325ULL ^ (uint64_t)s.length;
    for (int64_t i = 0; i < s.length; i++) {
        h = (h ^ (unsigned char)s.bases[i]) * 0x100000001b3ULL;
        h = (h ^ s.repeatCounts[i]) * 0x100000001b3ULL;
    }
    return h;
}

bool valid(const mp_rle_string &s) { return s.length >= 0 && (s.length == 0 || (s.bases && s.repeatCounts)); }

}  // namespace

extern "C" mp_status mp_bubble_allele_supports(mp_engine *e, const mp_model *m, const mp_band_params *p,
                                               int64_t nBubbles, const int64_t *alleleStart,
                                               const mp_rle_string *alleles, const int64_t *readStart,
                                               const mp_rle_string *reads, const int32_t *readStrand,
                                               int32_t includeRepeatCounts, int64_t maxRepeatCountExclusive,
                                               float *supports, int64_t *pairsOut) {
    if (!e || !m || !p || nBubbles < 0 || (nBubbles > 0 && (!alleleStart || !readStart || !supports)) ||
        maxRepeatCountExclusive < 1 || maxRepeatCountExclusive > 256) {
        mp::set_error("mp_bubble_allele_supports: bad argument");
        return MP_EINVAL;
    }
    if (pairsOut) *pairsOut = 0;
    if (nBubbles == 0) return MP_OK;
    if (alleleStart[0] != 0 || readStart[0] != 0) {
        mp::set_error("mp_bubble_allele_supports: alleleStart[0] and readStart[0] must be 0");
        return MP_EINVAL;
    }
    for (int64_t b = 0; b < nBubbles; b++)
        if (alleleStart[b + 1] < alleleStart[b] || readStart[b + 1] < readStart[b]) {
            mp::set_error("mp_bubble_allele_supports: start arrays must be non-decreasing");
            return MP_EINVAL;
        }
    const int64_t nA = alleleStart[nBubbles], nR = readStart[nBubbles];
    if ((nA > 0 && !alleles) || (nR > 0 && (!reads || !readStrand))) {
        mp::set_error("mp_bubble_allele_supports: null string array");
        return MP_EINVAL;
    }
    for (int64_t i = 0; i < nA; i++)
        if (!valid(alleles[i])) {
            mp::set_error("mp_bubble_allele_supports: bad allele string " + std::to_string(i));
            return MP_EINVAL;
        }
    for (int64_t i = 0; i < nR; i++)
        if (!valid(reads[i])) {
            mp::set_error("mp_bubble_allele_supports: bad read string " + std::to_string(i));
            return MP_EINVAL;
        }
    const bool rl = includeRepeatCounts != 0;
    const uint64_t cap = (uint64_t)maxRepeatCountExclusive;
    try {
        // each read -> the first read of its bubble with the same run-length string
        std::vector<int64_t> firstOf(nR);
        std::vector<uint64_t> rHash(nR);
        std::vector<int64_t> distinct;  // reads that are computed, in order
        for (int64_t b = 0; b < nBubbles; b++) {
            const size_t d0 = distinct.size();
            for (int64_t k = readStart[b]; k < readStart[b + 1]; k++) {
                rHash[k] = string_hash(reads[k]);
                int64_t f = -1;
                for (size_t q = d0; q < distinct.size() && f < 0; q++) {
                    const int64_t o = distinct[q];
                    if (rHash[o] == rHash[k] && same_string(reads[o], reads[k])) f = o;
                }
                if (f < 0) {
                    distinct.push_back(k);
                    f = k;
                }
                firstOf[k] = f;
            }
        }
        // symbol strings in one pool, then one pair per (allele, distinct read) of each bubble
        std::vector<int64_t> aOff(nA + 1, 0), rOff(nR + 1, 0);
        for (int64_t i = 0; i < nA; i++) aOff[i + 1] = aOff[i] + alleles[i].length;
        for (int64_t k = 0; k < nR; k++) rOff[k + 1] = rOff[k] + (firstOf[k] == k ? reads[k].length : 0);
        std::vector<uint16_t> aSym((size_t)std::max<int64_t>(aOff[nA], 1)), rSym((size_t)std::max<int64_t>(rOff[nR], 1));
        for (int64_t i = 0; i < nA; i++) symbols(alleles[i], rl, cap, aSym.data() + aOff[i]);
        for (int64_t k : distinct) symbols(reads[k], rl, cap, rSym.data() + rOff[k]);
        std::vector<mp_pair> pairs;
        std::vector<int64_t> pairSlot;  // output index of each pair
        std::vector<int64_t> outBase(nBubbles + 1, 0);
        for (int64_t b = 0; b < nBubbles; b++)
            outBase[b + 1] = outBase[b] + (alleleStart[b + 1] - alleleStart[b]) * (readStart[b + 1] - readStart[b]);
        size_t di = 0;
        for (int64_t b = 0; b < nBubbles; b++) {
            const int64_t rn = readStart[b + 1] - readStart[b];
            const size_t d0 = di;
            while (di < distinct.size() && distinct[di] < readStart[b + 1]) di++;
            for (int64_t j = alleleStart[b]; j < alleleStart[b + 1]; j++)
                for (size_t q = d0; q < di; q++) {
                    const int64_t k = distinct[q];
                    mp_pair pr{};
                    pr.sX = aSym.data() + aOff[j];
                    pr.lX = alleles[j].length;
                    pr.sY = rSym.data() + rOff[k];
                    pr.lY = reads[k].length;
                    pr.anchors = nullptr;
                    pr.nAnchors = 0;
                    pr.strand = readStrand[k] ? 1 : 0;
                    pairs.push_back(pr);
                    pairSlot.push_back(outBase[b] + (j - alleleStart[b]) * rn + (k - readStart[b]));
                }
        }
        if (pairsOut) *pairsOut = (int64_t)pairs.size();
        std::vector<double> lp(std::max<size_t>(pairs.size(), 1));
        if (!pairs.empty()) {
            mp_batch *bt = nullptr;
            mp_status st = mp_batch_create(e, m, pairs.data(), (int64_t)pairs.size(), p, MP_MODE_FORWARD_ONLY, 0,
                                           &bt);
            if (st) return st;
            std::vector<int64_t> counts(3 * pairs.size());
            if (!(st = mp_batch_run(bt))) st = mp_batch_fetch_counts(bt, counts.data(), lp.data());
            mp_batch_free(bt);
            if (st) return st;
        }
        for (size_t q = 0; q < pairs.size(); q++) supports[pairSlot[q]] = (float)lp[q];
        // repeated read substrings take their first occurrence's values
        for (int64_t b = 0; b < nBubbles; b++) {
            const int64_t rn = readStart[b + 1] - readStart[b];
            for (int64_t k = readStart[b]; k < readStart[b + 1]; k++) {
                if (firstOf[k] == k) continue;
                for (int64_t j = 0; j < alleleStart[b + 1] - alleleStart[b]; j++)
                    supports[outBase[b] + j * rn + (k - readStart[b])] =
                        supports[outBase[b] + j * rn + (firstOf[k] - readStart[b])];
            }
        }
    } catch (const std::bad_alloc &) {
        mp::set_error("mp_bubble_allele_supports: out of host memory");
        return MP_ENOMEM;
    }
    return MP_OK;
}
