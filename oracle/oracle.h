/*
 * oracle.h -- C interface of the CPU oracle (test infrastructure only; the
 * shipped library never includes or links it).
 */
#ifndef MP_ORACLE_H
#define MP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_LOGADD_POLY = 0, OR_LOGADD_EXACT = 1 };
enum { OR_OK = 0, OR_EBAND = 1, OR_EUNSUPPORTED = 2 };

/* A resolved 3-state machine in natural-log space (StateMachine3 fields,
 * stateMachine.c:490-500) plus the emission tables it reads. */
typedef struct {
    double matchContinue, matchFromGapX, matchFromGapY;
    double gapOpenX, gapOpenY, gapExtendX, gapExtendY, gapSwitchToX, gapSwitchToY;
    double match[16], gapX[4], gapY[4];
    int rle;                    /* run-length match term on (rleNucleotideEmissions_construct) */
    int strand;                 /* 1 forward, 0 reverse (selects the repeat-table row) */
    int maxRepeat;              /* repeat-table side, 51 */
    const double *repeatLog10;  /* [5][maxRepeat][maxRepeat] log10, [base][underlying][observed] */
    /* Five-state extension (no reference counterpart): nStates == 5 selects
     * states M, short X, short Y, long X, long Y with natural-log transitions
     * trans5[from * 5 + to]; the 3-state fields are then unused.  0 or 3
     * select the 3-state machine above. */
    int nStates;
    double trans5[25];
} or_model;

/* PairwiseAlignmentParameters (pairwiseAligner.h:28-49), the fields on the path. */
typedef struct {
    double threshold;
    int64_t minDiagsBetweenTraceBack;
    int64_t traceBackDiagonals;
    int64_t diagonalExpansion;
    int64_t splitMatrixBiggerThanThis;
    int dynamicAnchorExpansion;
} or_params;

typedef struct or_result or_result;

double or_logadd(double x, double y, int mode);
int or_band(const int64_t *anchors, int64_t n, int64_t lX, int64_t lY, int64_t expansion, int dynamic,
            int64_t *xmyL, int64_t *xmyR);
int64_t or_split_points(const int64_t *anchors, int64_t n, int64_t lX, int64_t lY, int64_t maxArea,
                        int raggedL, int raggedR, int64_t *out, int64_t cap);
or_result *or_posteriors(const or_model *m, const uint16_t *sX, int64_t lX, const uint16_t *sY, int64_t lY,
                         const int64_t *anchors, int64_t n, const or_params *p, int raggedL, int raggedR,
                         int withIndels, int mode);
or_result *or_expectations(const or_model *m, const uint16_t *sX, int64_t lX, const uint16_t *sY, int64_t lY,
                           const int64_t *anchors, int64_t n, const or_params *p, int raggedL, int raggedR,
                           int mode);
double or_forward_prob(const or_model *m, const uint16_t *sX, int64_t lX, const uint16_t *sY, int64_t lY,
                       const int64_t *anchors, int64_t n, const or_params *p, int raggedL, int raggedR, int mode);
int or_full_dp(const or_model *m, const uint16_t *sX, int64_t lX, const uint16_t *sY, int64_t lY,
               int64_t expansion, int mode, double *fwdTotal, double *bwdTotal, double *cutTotals, double *postM);

int or_result_status(const or_result *r);
int64_t or_result_count(const or_result *r, int list);
void or_result_copy(const or_result *r, int list, int64_t *out);
double or_result_total(const or_result *r);
double or_result_likelihood(const or_result *r);
int64_t or_result_cells(const or_result *r);
void or_result_expectations(const or_result *r, double *out9);
/* expected emission counts (expectations runs), forward-strand parameters:
 * match[16] [x*4+y], gapX[4], gapY[4], repeat[4][51][51] [base][underlying][observed] */
void or_result_emissions(const or_result *r, double *out);
void or_result_free(or_result *r);

int64_t or_batch_cells(const or_model *models, int64_t npairs,
                       const uint16_t *symX, const int64_t *offX, const int64_t *lenX,
                       const uint16_t *symY, const int64_t *offY, const int64_t *lenY,
                       const int64_t *anchors, const int64_t *anchorOff, const int64_t *anchorN,
                       const int *strand, const or_params *p, int mode, int nthreads, int64_t *outCounts);

#ifdef __cplusplus
}
#endif
#endif
