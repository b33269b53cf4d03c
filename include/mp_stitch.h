/*
 * mp_stitch.h -- C ABI of chunk stitching: the overlap alignment between
 * neighbouring polished chunks (removeOverlap) for many boundaries in one
 * GPU batch, and the contig merge built on it (mergeContigChunks).
 *
 * removeOverlap (impl/polisher.c:1937-2006): align the last approxOverlap
 * characters of the prefix string against the first approxOverlap of the
 * suffix string with getAlignedPairs (default three-state nucleotide machine,
 * stateMachine3_constructNucleotide; no anchors; both ends ragged; plain
 * symbols without run lengths), take the first match tuple of maximal weight
 * and cut there: prefix kept up to x + offset (exclusive), suffix kept from y
 * (inclusive); no tuple: keep both whole, weight 0.
 */
#ifndef MP_STITCH_H
#define MP_STITCH_H

#include <stdint.h>

#include "mp_pairhmm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* removeOverlap for n independent (prefix, suffix) pairs in one batch.
   Strings are plain characters (non-ACGT read as N); lengths exclude any
   terminator.  Outputs per pair: prefixCropEnd, suffixCropStart and the
   overlap match weight (floor(p * 1e7) of the chosen tuple, 0 if none). */
mp_status mp_remove_overlaps(mp_engine *e, const mp_band_params *p, int64_t n, const char *const *prefix,
                             const int64_t *prefixLen, const char *const *suffix, const int64_t *suffixLen,
                             const int64_t *approxOverlap, int64_t *prefixCropEnd, int64_t *suffixCropStart,
                             int64_t *overlapWeight);

/* mergeContigChunks (chunker.c:55-118): chunks in contig order, each trimmed
   against its predecessor with removeOverlap(previous polished string,
   chunk, overlap); an empty chunk with no overlap becomes missingSpacer, a
   non-empty one with no overlap (overlap > 0) gets "NNNNNNNNNN" before it.
   All boundaries are aligned in one batch; a boundary whose predecessor was
   cropped into its overlap window is re-aligned with the cropped string, so
   the result equals the reference's sequential merge.  Writes the merged
   string (not terminated) to out when outCap is large enough; *outLen gets
   its length either way (MP_EINVAL with *outLen set when out is too small). */
mp_status mp_merge_chunks(mp_engine *e, const mp_band_params *p, int64_t nChunks, const char *const *chunks,
                          const int64_t *chunkLen, int64_t overlap, const char *missingSpacer, char *out,
                          int64_t outCap, int64_t *outLen);

#ifdef __cplusplus
}
#endif

#endif /* MP_STITCH_H */
