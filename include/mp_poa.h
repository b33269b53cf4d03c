/*
 * mp_poa.h -- C ABI of the POA augmentation that consumes the pair-HMM's
 * posterior lists: the reference graph of a chunk and its per-read
 * augmentation with match weights, complete inserts and complete deletes.
 *
 * Replaces the augment half of poa_realign (impl/polisher.c:656-703):
 * poa_getReferenceGraph (polisher.c:118-131) and poa_augment
 * (polisher.c:316-534) with addToInserts / addToDeletes (177-235) and the
 * left shift getShift / getMaxCommonSuffixLength / rleString_rotateString
 * (269-314, 1508-1520).  The graph is read back through the accessors below
 * (the fields of PoaNode / PoaInsert / PoaDelete / PoaBaseObservation,
 * inc/margin.h:599-627).
 *
 * Weights are doubles summed in read order, as the reference sums them; with
 * integer tuple weights (the posterior lists' floor(p * 1e7)) every sum is
 * exact.  Insert and delete lists keep the reference's order: an entry is
 * appended when first seen, reads in order, complete inserts in the order
 * the reference enumerates them.
 */
#ifndef MP_POA_H
#define MP_POA_H

#include <stdint.h>

#include "mp_pairhmm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* RleString (inc/margin.h): run bases as characters (A/C/G/T in either case;
   anything else counts as N) and run lengths >= 1. */
typedef struct {
    const char *bases;
    const uint64_t *repeatCounts;
    int64_t length; /* number of runs */
} mp_rle_string;

/* PoaBaseObservation (inc/margin.h:623-627). */
typedef struct {
    int64_t readNo;
    int64_t offset;
    double weight;
} mp_poa_observation;

typedef struct mp_poa mp_poa;

/* Sizes of a graph's contents, for sizing the read-out buffers. */
typedef struct {
    int64_t nodes;                /* reference length + 1 (node 0 is the prefix 'N'/1 node) */
    int64_t maxRepeatCount;       /* exclusive bound of repeatCountWeights */
    int64_t inserts;              /* PoaInsert entries over all nodes */
    int64_t insertSymbols;        /* total run count of the insert strings */
    int64_t deletes;              /* PoaDelete entries over all nodes */
    int64_t nodeObservations;     /* base observations over all nodes */
    int64_t insertObservations;
    int64_t deleteObservations;
} mp_poa_sizes;

/* poa_getReferenceGraph (polisher.c:118-131): node 0 = 'N' with repeat count
   1, then one node per reference run (base upper-cased, non-ACGT -> 'N'),
   all weights zero.  maxRepeatCount is exclusive (poa_realign passes
   MAXIMUM_REPEAT_LENGTH = 51 with run-length encoding, 2 without). */
mp_status mp_poa_reference_graph(const mp_rle_string *reference, int64_t maxRepeatCount, mp_poa **out);
void mp_poa_free(mp_poa *poa);

/* poa_augment (polisher.c:316-534) for one read: the read's match list
   (weights into baseWeights / repeatCountWeights of node x + 1, one base
   observation each), its gap-Y list as inserts and its gap-X list as deletes
   (complete runs, left-shifted, deduplicated per node).  readStrand: 1 =
   forward (weightForwardStrand).  compareRepeatCounts =
   polishParams->poaConstructCompareRepeatCounts.  Coordinates outside the
   reference / read give MP_EINVAL (the reference asserts). */
mp_status mp_poa_augment(mp_poa *poa, const mp_rle_string *read, int32_t readStrand, int64_t readNo,
                         const mp_aligned_pair *matches, int64_t nMatches, const mp_aligned_pair *inserts,
                         int64_t nInserts, const mp_aligned_pair *deletes, int64_t nDeletes,
                         int32_t compareRepeatCounts);

/* The whole augment loop of poa_realign for every pair of a posterior batch
   that has run (MP_MODE_POSTERIORS): pair i is read pairReadNo[i] of graph
   poas[pairPoa[i]], with run-length string reads[i] (uncapped repeat counts)
   and the strand it was created with; its lists go in without leaving the
   library.  Each graph takes its pairs in batch order, so a graph sees the
   same sequence of poa_augment calls as the reference's loop; graphs are
   processed on up to `threads` host threads (0 = all cores).  The pairs'
   reference must be the graph's reference (list x coordinates index it). */
mp_status mp_batch_poa_augment(mp_batch *b, mp_poa *const *poas, int64_t nPoas, const int64_t *pairPoa,
                               const int64_t *pairReadNo, const mp_rle_string *reads, int32_t compareRepeatCounts,
                               int32_t threads);

/* poa_estimateRepeatCountsUsingBayesianModel (polisher.c:1557-1566) with
   repeatSubMatrix_getMLRepeatCount (1846-1865): for every node after the
   prefix node, the maximum-likelihood run length of its base given its base
   observations, each contributing weight / 1e7 times the model's repeat-count
   log-probability (repeatCountSubstitutionMatrix of the read's strand; an N
   node uses the A row) of the read's run length at the observed offset,
   capped at MAXIMUM_REPEAT_LENGTH - 1; candidates run from the smallest to
   the largest (capped) observed run length, the first maximum wins, and 0 (no
   observation) becomes 1.  reads[readNo] / readStrand[readNo] are the reads
   the graph was augmented with (uncapped run lengths).  Writes nodes - 1
   counts (the polished reference's run lengths). */
mp_status mp_poa_estimate_repeat_counts(const mp_poa *poa, const mp_model *m, const mp_rle_string *reads,
                                        const int32_t *readStrand, int64_t nReads, uint64_t *repeatCounts,
                                        int32_t threads);

/* The same for nPoas graphs at once (graphs processed in parallel): graph g's
   reads are reads[readStart[g] .. readStart[g + 1]) (indexed by readNo from
   readStart[g]), its counts go to repeatCounts[g] (nodes - 1 entries). */
mp_status mp_poas_estimate_repeat_counts(const mp_poa *const *poas, int64_t nPoas, const mp_model *m,
                                         const mp_rle_string *reads, const int32_t *readStrand,
                                         const int64_t *readStart, uint64_t *const *repeatCounts, int32_t threads);

/* poa_getAnchorAlignments (polisher.c:536-592): anchor pairs for the next
   realignment from the graph's base observations.  Nodes in order (node i
   maps to consensus position poaToConsensusMap[i - 1], or i - 1 when the map
   is null; -1 = unaligned), each node's observations in order: an
   observation with weight / 1e7 > anchorParams[0] becomes (consensus
   position, read offset, expansion) of its read, expansion = anchorParams[1]
   raised to anchorParams[k + 1] for every further threshold anchorParams[k]
   (k = 2, 4, ...) the weight reaches, appended only when both coordinates
   exceed the read's previous anchor.  counts[r] receives read r's anchor
   count (r < nReads); when anchors is not null the anchors (3 int64 each)
   are written read after read, and capacity (in anchors) must hold them all. */
mp_status mp_poa_anchor_alignments(const mp_poa *poa, const int64_t *poaToConsensusMap, int64_t nReads,
                                   const double *anchorParams, int64_t nAnchorParams, int64_t *counts,
                                   int64_t *anchors, int64_t capacity);

/* poa_getConsensus (polisher.c:1040-1275) with getMaxWeight (1025-1038):
   forward log-probabilities through the graph (match / insert / delete
   transitions weighted by the node's base and indel weights, the reference's
   polynomial logAdd), then a greedy traceback from the end; each visited
   node contributes its heaviest base (and, with run-length encoding, its
   heaviest run length) unless the reference's scaled by
   referenceBasePenalty still beats it.  Writes the consensus as an expanded
   string (not terminated) when outCap suffices (*outLen always; MP_EINVAL
   when out is too small) and, when poaToConsensusMap is not null, nodes - 1
   entries mapping node i to the position of its base in the run-length (or
   plain) consensus, -1 where none.  A reference run length at or beyond the
   graph's maxRepeatCount reads as weight 0 (the reference reads past its
   array there).  The result is kept with the graph until it changes, so a
   size query followed by the fill computes it once (calls on one graph must
   not overlap). */
mp_status mp_poa_consensus(const mp_poa *poa, int32_t useRunLengthEncoding, double referenceBasePenalty, char *out,
                           int64_t outCap, int64_t *outLen, int64_t *poaToConsensusMap);

/* The graph's weight totals (polisher.c:706-758): scores[0] =
   poa_getReferenceNodeTotalMatchWeight, [1] =
   poa_getReferenceNodeTotalDisagreementWeight, [2] = poa_getInsertTotalWeight
   (weight x run count), [3] = poa_getDeleteTotalWeight (weight x length);
   poa_realignIterative's score is [0] - ([1] + [2] + [3]). */
mp_status mp_poa_scores(const mp_poa *poa, double *scores);

mp_status mp_poa_get_sizes(const mp_poa *poa, mp_poa_sizes *sizes);

/* Per node: base, repeatCount, baseWeights[5] (A C G T N) and
   repeatCountWeights[maxRepeatCount], row-major by node.  Any pointer may be null. */
mp_status mp_poa_get_nodes(const mp_poa *poa, char *bases, uint64_t *repeatCounts, double *baseWeights,
                           double *repeatCountWeights);

/* Every PoaInsert, grouped by node (ascending) and in each node's list order:
   its node, strand weights, and its run-length string as symbolStart /
   symbolLength into bases / repeatCounts (insertSymbols entries).  Any
   pointer may be null. */
mp_status mp_poa_get_inserts(const mp_poa *poa, int64_t *node, double *weightForwardStrand,
                             double *weightReverseStrand, int64_t *symbolStart, int64_t *symbolLength, char *bases,
                             uint64_t *repeatCounts);

/* Every PoaDelete, grouped by node and in list order.  Any pointer may be null. */
mp_status mp_poa_get_deletes(const mp_poa *poa, int64_t *node, int64_t *length, double *weightForwardStrand,
                             double *weightReverseStrand);

/* Observations of kind 0 (nodes), 1 (inserts, in mp_poa_get_inserts order) or
   2 (deletes): start[k]..start[k+1] are entity k's observations in append
   order; start has (entities + 1) entries. */
mp_status mp_poa_get_observations(const mp_poa *poa, int32_t kind, int64_t *start, mp_poa_observation *obs);

#ifdef __cplusplus
}
#endif

#endif /* MP_POA_H */
