/*
 * mp_pairhmm.h -- C ABI of the MI355X pair-HMM realignment engine.
 *
 * Drop-in for the hot path of MarginPolish's read realignment: the banded
 * pair-HMM forward-backward with run-length-aware emissions and posterior
 * match / gap-X / gap-Y extraction.  Plain C types only (pointers, sizes,
 * POD structs); no torch or HIP types cross this boundary.
 *
 * Coordinates follow the reference: sX is the reference (x), sY the read (y);
 * symbols are uint16 (run length << 8 | base, base in {A,C,G,T,N} = 0..4,
 * stateMachine.c:116-146); anchors are (x, y, expansion) triples in sequence
 * coordinates with strictly increasing x and y and even expansion
 * (pairwiseAligner.c:120-173).  Output tuples are (weight, x, y) with
 * weight = floor(min(p,1) * 1e7) for posteriors p >= threshold and
 * coordinates minus one, exactly as addPosteriorProb (pairwiseAligner.c:586-595);
 * gap-X tuples may have y == -1 and gap-Y tuples x == -1.
 *
 * Errors never abort: every entry returns an mp_status and
 * mp_last_error() gives a thread-local message.
 */
#ifndef MP_PAIRHMM_H
#define MP_PAIRHMM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    MP_OK = 0,
    MP_EINVAL = 1,        /* bad argument (null pointer, negative length)          */
    MP_EPARAMS = 2,       /* params JSON: unreadable, unknown key, missing section */
    MP_EANCHORS = 3,      /* anchors not strictly increasing / out of range / odd  */
    MP_EHIP = 4,          /* HIP runtime error                                     */
    MP_ENOMEM = 5,        /* host or device allocation failed                      */
    MP_EUNSUPPORTED = 6,  /* band wider than the kernel supports                   */
    MP_ENODEVICE = 7      /* no gfx950 device visible                              */
} mp_status;

/* Output lists of one pair. */
enum { MP_LIST_MATCH = 0, MP_LIST_GAPX = 1, MP_LIST_GAPY = 2 };

/* What a batch computes per pair. */
enum {
    MP_MODE_POSTERIORS = 0,   /* match + gapX + gapY posteriors (getAlignedPairsWithIndelsUsingAnchors) */
    MP_MODE_MATCH_ONLY = 1,   /* match posteriors only (getAlignedPairsUsingAnchors)                     */
    MP_MODE_FORWARD_ONLY = 2, /* forward log-likelihood only (computeForwardProbability): static band of
                                 diagonalExpansion, no large-gap split, total at the last anti-diagonal   */
    MP_MODE_EXPECTATIONS = 3  /* expected transition counts + likelihood (getExpectationsUsingAnchors)   */
};

/* Batch flags. */
enum {
    MP_BATCH_CROP_REFERENCE = 1, /* crop sX per pair as poa_realign's caller does (polisher.c:605-654) */
    MP_BATCH_FP64 = 2,           /* FP64 arithmetic on the per-cell-exponent kernel (no reference counterpart:
                                    the precision comparison path); bands up to 1216 cells, no expectations */
    MP_BATCH_NO_GRAPH = 4,       /* enqueue each run's launches directly instead of replaying the batch's
                                    hipGraphs (captured at the first run, re-captured after mp_batch_set_model) */
    MP_BATCH_EXACT_CHAIN = 8     /* posterior modes: run the traceback windows of a sub-problem one after the
                                    other on one forward chain, as the reference does.  Without it a band of at
                                    most 64 cells with several windows runs each window as a job of its own,
                                    whose forward sweep restarts 512 anti-diagonals before the window from a
                                    flat vector (posteriors are invariant to the forward's scale and its
                                    direction forgets the start within ~100 anti-diagonals of an anchored
                                    band); the log-likelihood is chained in FP64 across the windows */,
    MP_BATCH_REFERENCE_ARITHMETIC = 16 /* the reference's own arithmetic on the GPU: FP64 log space with its
                                    polynomial logAdd (pairwiseAligner.c:279-299), its operation order, windows and
                                    every-10th-diagonal total (687-808), so weights are the reference's (up to the
                                    last bit of exp); three-state machines, posterior and match-only modes, any
                                    band width.  For callers whose integer decisions must be the reference's
                                    (chunk stitching, consensus identity); one workgroup per sub-problem */
};

typedef struct mp_model mp_model;          /* both strands' state machines + tables */
typedef struct mp_engine mp_engine;        /* one device, one stream, reusable buffers */
typedef struct mp_batch mp_batch;          /* a device-resident batch of pairs */

/* (weight, x, y): the stIntTuple of the reference lists. */
typedef struct {
    int64_t weight;
    int32_t x;
    int32_t y;
} mp_aligned_pair;

/* PairwiseAlignmentParameters (pairwiseAligner.h:28-49), the fields the path reads. */
typedef struct {
    double threshold;                  /* 0.01 */
    int64_t minDiagsBetweenTraceBack;  /* 10000 in the shipped params */
    int64_t traceBackDiagonals;        /* 40 */
    int64_t diagonalExpansion;         /* 10 */
    int64_t splitMatrixBiggerThanThis; /* 250e6 */
    int32_t dynamicAnchorExpansion;    /* 1 */
} mp_band_params;

/* One read/reference pair, the arguments of
 * getAlignedPairsWithIndelsUsingAnchors (pairwiseAligner.h:75-77). */
typedef struct {
    const uint16_t *sX;
    int64_t lX;
    const uint16_t *sY;
    int64_t lY;
    const int64_t *anchors;  /* nAnchors (x, y, expansion) triples; may be NULL when nAnchors == 0 */
    int64_t nAnchors;
    int32_t strand;          /* 1: stateMachineForForwardStrandRead, 0: ...ReverseStrandRead */
    int32_t raggedLeft;
    int32_t raggedRight;
} mp_pair;

const char *mp_last_error(void);
const char *mp_version(void);

/* ---- model: replaces params_readParams / polishParams_jsonParse
 *      (parser.c:154-408, 503-528) and hmm_getStateMachine
 *      (stateMachine.c:676-689) for the keys on this path.  Same JSON schema;
 *      an unknown key in the polish section is an error as in parser.c:374. */
mp_status mp_model_load_json(const char *path, mp_model **out);
/* stateMachine3_constructNucleotide(threeState) (stateMachine.c:611-620):
 * default transitions and nucleotide emissions, no run-length term. */
mp_status mp_model_default_nucleotide(mp_model **out);
mp_status mp_model_band_params(const mp_model *m, mp_band_params *out);
/* Resolved natural-log tables of one strand (for inspection and tests):
 * trans9 = {matchContinue, matchFromGapX, matchFromGapY, gapOpenX, gapOpenY,
 *           gapExtendX, gapExtendY, gapSwitchToX, gapSwitchToY}. */
mp_status mp_model_tables(const mp_model *m, int32_t strand, double *trans9, double *match16, double *gapX4,
                          double *gapY4, int32_t *rle);
/* repeatSubMatrix_getLogProb (polisher.c:1756-1775): log10 P(observed | underlying, base, strand). */
double mp_model_repeat_log10(const mp_model *m, int32_t base, int32_t strand, int32_t observed, int32_t underlying);
int32_t mp_model_max_repeat_count(const mp_model *m);  /* maximumRepeatLength (51) */
/* stateMachine3_loadAsymmetric's transition part (stateMachine.c:632-668), as
 * hmm_getStateMachine (676-689) applies an Hmm's trained transitions: replace
 * both strands' natural-log transitions (order of mp_model_tables' trans9).
 * pinEnd != 0 keeps the end-state vector (stateMachine.c:537-557) of the
 * transitions in force at the first pinned call, so that the derivative of a
 * pair's log-likelihood in each log transition is exactly its expected count
 * (the parameter-fitting harness); pinEnd = 0 unpins. */
mp_status mp_model_set_transitions(mp_model *m, const double *logTrans9, int32_t pinEnd);
/* Five-state machine: no counterpart in the reference, whose StateMachineType
 * has the 3-state types 2 and 3 only (stateMachine.h:96-99).  States M, short
 * gap X, short gap Y, long gap X, long gap Y; a gap state of either length
 * emits like the 3-state gap of its side, and a cell's gap-X (gap-Y)
 * posterior is the sum over its two gap-X (gap-Y) states.  Boundary vectors:
 * start in M; ragged start in any gap state; end through each state's
 * transition to M; ragged end: M at exp of the mean of the two short-gap open
 * logs, each gap state at its own extension.  The JSON loader reads such a
 * machine from an hmm of "type": 5 with 25 "transitions" [from * 5 + to].
 * Five-state models run every job on a one-wavefront kernel with a per-cell
 * exponent (band widths up to 2048 cells; wider is MP_EUNSUPPORTED) in the
 * posterior, match-only and forward-only modes; expected counts are
 * MP_EUNSUPPORTED.
 * mp_model_set_transitions5 makes the model five-state with these natural-log
 * transitions [from * 5 + to] on both strands; mp_model_transitions5 reads
 * them back; mp_model_num_states returns 3 or 5 (0 for a null model). */
/* Emission part of stateMachine3_loadAsymmetric / hmm_getStateMachine
 * (stateMachine.c:632-689), for the parameter-fitting harness: replace the
 * forward strand's natural-log emissions (match16 = joint P(x base, y base)
 * [x * 4 + y], gapX4, gapY4); the reverse strand takes their reverse
 * complement (nucleotideEmissions_reverseComplement, stateMachine.c:454-470). */
mp_status mp_model_set_emissions(mp_model *m, const double *logMatch16, const double *logGapX4,
                                 const double *logGapY4);
/* repeatSubMatrix_setLogProb (polisher.c:1777-1794) for every entry: rows 0-3
 * of the repeat table, log10 P(observed | underlying, base) [base][underlying]
 * [observed] (51 x 51 per base; NULL keeps the table), and enable = the
 * run-length term in the match emission (rleNucleotideEmissions_construct,
 * parser.c:391-401) on both strands. */
mp_status mp_model_set_repeat_log10(mp_model *m, const double *log10Repeat, int32_t enable);
mp_status mp_model_set_transitions5(mp_model *m, const double *logTrans25);
mp_status mp_model_transitions5(const mp_model *m, double *logTrans25);
int32_t mp_model_num_states(const mp_model *m);
void mp_model_free(mp_model *m);

/* ---- host geometry (no device needed) ---- */
/* band_construct / band_constructDynamic (pairwiseAligner.c:120-226): xmyL and
 * xmyR of every anti-diagonal 0..lX+lY (arrays of lX+lY+1). */
mp_status mp_band_construct(const int64_t *anchors, int64_t nAnchors, int64_t lX, int64_t lY, int64_t expansion,
                            int32_t dynamic, int32_t *xmyL, int32_t *xmyR);
/* getSplitPoints (pairwiseAligner.c:899-926): writes up to cap (x1, y1, x2, y2)
 * tuples, returns the number of split blocks (or -1 on bad arguments). */
int64_t mp_split_points(const int64_t *anchors, int64_t nAnchors, int64_t lX, int64_t lY, int64_t maxArea,
                        int32_t raggedLeft, int32_t raggedRight, int64_t *out, int64_t cap);
/* Traceback windows of getPosteriorProbsWithBanding (pairwiseAligner.c:715-783)
 * for a band: up to cap (tracebackDiagonal, tracedBackTo, tracedBackFrom)
 * triples; returns the number of windows. */
int64_t mp_traceback_windows(const int32_t *xmyL, const int32_t *xmyR, int64_t lXlY, const mp_band_params *p,
                             int32_t *out, int64_t cap);

/* ---- engine ---- */
mp_status mp_engine_create(int32_t device, mp_engine **out);
void mp_engine_free(mp_engine *e);
/* Compute units of the engine's device (for launch sizing reports). */
int32_t mp_engine_cu_count(const mp_engine *e);

/* ---- per-pair entry: getAlignedPairsWithIndelsUsingAnchors /
 *      getAlignedPairsUsingAnchors (pairwiseAligner.c:1088-1130), including
 *      the large-gap split (getSplitPoints, 875-991).  One kernel launch per
 *      call.  Lists are written into caller-owned arrays; *count receives the
 *      list length; when it exceeds *cap nothing is truncated silently: the
 *      call returns MP_OK with counts set and the arrays filled only up to cap,
 *      so the caller can retry with larger arrays.  logProb receives the
 *      pair's forward log-likelihood (the sum over its split sub-problems). */
mp_status mp_pairhmm_posteriors(mp_engine *e, const mp_model *m, const mp_pair *pair, const mp_band_params *p,
                                int32_t mode, mp_aligned_pair *matches, int64_t matchCap, int64_t *matchCount,
                                mp_aligned_pair *gapX, int64_t gapXCap, int64_t *gapXCount,
                                mp_aligned_pair *gapY, int64_t gapYCap, int64_t *gapYCount, double *logProb);

/* ---- batched entry: every pair of one POA iteration (all chunks) in one
 *      persistent kernel launch.  poa_realign (polisher.c:656-703) loops per
 *      read; the batch replaces that loop. ---- */

/* Host preparation + upload: split points, cropping (flag), band geometry,
 * traceback windows, band-area job ordering, output sizing. */
mp_status mp_batch_create(mp_engine *e, const mp_model *m, const mp_pair *pairs, int64_t npairs,
                          const mp_band_params *p, int32_t mode, int32_t flags, mp_batch **out);
/* Re-read the model's transitions and emission tables into an existing batch
 * (same pairs, bands and outputs): the per-iteration model update of
 * parameter fitting.  No counterpart in the reference, whose state machine is
 * a pointer read at every cell. */
mp_status mp_batch_set_model(mp_batch *b, const mp_model *m);
/* One persistent launch over every pair (outputs stay on the device). */
mp_status mp_batch_run(mp_batch *b);
/* Autotuner (hardware tiling only; no reference counterpart): times this batch
 * under each launch geometry -- narrow-kernel workgroups of 512 or 256 threads
 * (the same resident wavefronts), five-state workgroups of 8/4/2/1 waves --
 * `repeats` runs each (best of), keeps the fastest for this batch and caches it
 * in the engine for later batches of the same shape (mode, kernel family,
 * dominant band class, widest band, log2 job count).  Results are identical
 * under every geometry; the batch is left run.  Outputs may be NULL. */
mp_status mp_batch_autotune(mp_batch *b, int32_t repeats, int32_t *narrowThreads, int32_t *wideWaves,
                            float *kernelMs);
/* The same work as one kernel launch per pair (the plain per-read path). */
mp_status mp_batch_run_plain(mp_batch *b);
/* Device time of the last run's pair-HMM kernel launch(es), milliseconds (HIP events on the engine's stream). */
double mp_batch_kernel_ms(const mp_batch *b);
/* Device time of the last run's list assembly and log-likelihood pass, which
 * follows the kernels in every run (the reference's lists per pair, built on
 * the device; milliseconds). */
double mp_batch_list_ms(const mp_batch *b);
/* Sum of the pairs' log-likelihoods of the last run, accumulated in FP64 on
 * the device in a fixed order (each pair's own value as mp_batch_fetch_counts
 * reports it). */
mp_status mp_batch_fetch_total_log_prob(mp_batch *b, double *total);
/* Diagnostics (no reference counterpart): device jobs of the batch (split
 * sub-problems; a sub-problem run as window jobs counts once per window),
 * posterior-record capacity, and the records the last run wrote.  Outputs
 * may be NULL. */
mp_status mp_batch_stats(const mp_batch *b, int64_t *jobs, int64_t *recordCapacity, int64_t *records);
/* The last run's record-region re-runs (no counterpart in the reference): a
 * job that writes more posterior records than its region holds makes the run
 * size every region exactly and run the batch again.  reruns: how many times
 * that happened; overflowJobs: jobs that were over; missingRecords: the
 * records beyond their regions.  Outputs may be NULL. */
mp_status mp_batch_rerun_stats(const mp_batch *b, int32_t *reruns, int64_t *overflowJobs, int64_t *missingRecords);
/* Host time (ms) of mp_batch_create's phases (no counterpart in the
 * reference): [0] checks, reference crop and split points, [1] symbol
 * encoding and jobs, [2] bands, windows and classes, [3] window jobs,
 * [4] device order and record regions, [5] slots, [6] upload and device
 * buffers, [7] list buffers.  Writes min(n, 8) entries. */
mp_status mp_batch_prep_ms(const mp_batch *b, double *ms, int32_t n);
int64_t mp_batch_cells(const mp_batch *b);          /* band cells over all pairs (integer band geometry) */
int64_t mp_batch_npairs(const mp_batch *b);
/* Copy results to the host.  counts: npairs*3 list lengths; logProb: npairs
 * (may be NULL).  Lists are reported per pair in the reference order and, with
 * MP_BATCH_CROP_REFERENCE, in uncropped reference coordinates. */
mp_status mp_batch_fetch_counts(mp_batch *b, int64_t *counts, double *logProb);
/* Copy list `list` of every pair, concatenated in pair order (sizes from
 * mp_batch_fetch_counts).  out must hold sum(counts[:, list]) entries. */
mp_status mp_batch_fetch_list(mp_batch *b, int32_t list, mp_aligned_pair *out);
/* Expectation mode: per pair the 9 expected transition counts T[from*3+to]
 * (M=0, X=1, Y=2) summed over the posterior diagonals of every window and
 * split, and the reference's per-diagonal likelihood sum (pairwiseAligner.c:674). */
mp_status mp_batch_fetch_expectations(mp_batch *b, double *trans, double *likelihood);
/* Expected emission counts of an expectation-mode batch, summed over its
 * pairs, on the forward-strand parameters: match16 [x base * 4 + y base],
 * gapX4, gapY4 and the repeat table's rows 0-3 [base][underlying][observed]
 * (4 x 51 x 51; the derivative of the log-likelihood in a log10 entry is
 * 2.3025 times its count).  Each is the sum over cells on posterior
 * anti-diagonals of the posterior of the state that reads the entry there;
 * with one traceback window per pair it is the derivative of the summed
 * log-likelihood in the natural-log parameter.  No counterpart in the
 * reference, whose emission expectations are disabled (pairwiseAligner.c:
 * 359-362).  Any output may be NULL. */
mp_status mp_batch_fetch_emission_counts(mp_batch *b, double *match16, double *gapX4, double *gapY4,
                                         double *repeat);
void mp_batch_free(mp_batch *b);

/* computeForwardProbability (pairwiseAligner.c:813-865) for one pair. */
mp_status mp_pairhmm_forward_log_prob(mp_engine *e, const mp_model *m, const mp_pair *pair, const mp_band_params *p,
                                      double *logProb);
/* No counterpart in the reference: computeForwardProbability for one pair
 * with no device allocation, for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * result is mp_pairhmm_forward_log_prob's, from the same kernel.  Its device
 * memory is a 16 MiB scratch array per GPU, held until the process ends:
 * pairs whose tables, band entries and symbols exceed it (about 1.9 M
 * anti-diagonals), bands wider than 64 cells and five-state models return
 * MP_EUNSUPPORTED.  Calls on one GPU share that scratch and the engine's
 * stream: calls that overlap on one GPU race. */
mp_status mp_pairhmm_forward_log_prob_noalloc(mp_engine *e, const mp_model *m, const mp_pair *pair,
                                              const mp_band_params *p, double *logProb);
/* getExpectationsUsingAnchors (pairwiseAligner.c:1151-1156) for one pair:
 * trans9 receives the expected transition counts, likelihood the hmm->likelihood increment. */
mp_status mp_pairhmm_expectations(mp_engine *e, const mp_model *m, const mp_pair *pair, const mp_band_params *p,
                                  double *trans9, double *likelihood);

#ifdef __cplusplus
}
#endif
#endif
