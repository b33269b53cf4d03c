/*
 * mp_bubble.h -- C ABI of the bubble graph's per-read x per-allele forward
 * likelihoods, batched over many bubbles in one forward-only GPU launch.
 *
 * Replaces the likelihood loop of bubbleGraph_constructFromPoa
 * (impl/bubbleGraph.c:895-928): for every bubble, every allele j and every
 * read substring k, alleleReadSupports[j * readNo + k] =
 * computeForwardProbability(allele_j, read_k, no anchors, p, strand
 * machine of read k, no ragged ends) (pairwiseAligner.c:813-865), stored as
 * float.  As in the reference, a read substring equal to an earlier one of
 * the same bubble (same run-length string) is not computed again: it takes
 * the earlier read's values, computed with the earlier read's strand.
 * Construction of the bubbles themselves (candidate weights, anchors, read
 * substrings, candidate alleles) stays with the caller.
 */
#ifndef MP_BUBBLE_H
#define MP_BUBBLE_H

#include <stdint.h>

#include "mp_pairhmm.h"
#include "mp_poa.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Bubble b has alleles [alleleStart[b], alleleStart[b+1]) of `alleles` and
   reads [readStart[b], readStart[b+1]) of `reads` / `readStrand` (1 =
   forward machine).  Strings become symbols as rleString_constructSymbolString
   does (stateMachine.c:124-146): base A/C/G/T (either case) = 0..3, anything
   else 4; with includeRepeatCounts the run length, capped at
   maxRepeatCountExclusive - 1, goes into the high byte.  Output: bubble b's
   alleleNo x readNo floats, allele-major, bubbles concatenated in order
   (supports has sum over bubbles of alleleNo * readNo entries).  pairs, when
   not null, receives the number of forward computations run (after the
   per-bubble read deduplication). */
mp_status mp_bubble_allele_supports(mp_engine *e, const mp_model *m, const mp_band_params *p, int64_t nBubbles,
                                    const int64_t *alleleStart, const mp_rle_string *alleles,
                                    const int64_t *readStart, const mp_rle_string *reads,
                                    const int32_t *readStrand, int32_t includeRepeatCounts,
                                    int64_t maxRepeatCountExclusive, float *supports, int64_t *pairs);

#ifdef __cplusplus
}
#endif

#endif /* MP_BUBBLE_H */
