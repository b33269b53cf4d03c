/*
 * pairhmm_oracle.c -- CPU oracle for the banded pair-HMM realignment path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this file; only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py use it, as the checker and as the timed CPU baseline.
 *
 * It restates, from the reference C sources, the FP64 log-space algorithm
 * that MarginPolish runs per read (reference files cited per function):
 *   impl/pairwiseAligner.c   bands, logAdd, cells, windowed forward-backward,
 *                            total-probability cut, posterior extraction,
 *                            split points, forward-only, expectations
 *   impl/stateMachine.c      3-state cell update and boundary vectors,
 *                            nucleotide and run-length (RLE) emissions
 *   impl/polisher.c          repeat-count table indexing
 *
 * Two arithmetic modes share every line of control flow:
 *   OR_LOGADD_POLY  ("O_faithful") the reference's 4-piece cubic lookup
 *                   (pairwiseAligner.c:279-299), so the numbers, the window
 *                   cadence and the output order match the reference;
 *   OR_LOGADD_EXACT ("O_exact") max + log1p(exp(-|d|)), the FP64 oracle the
 *                   GPU posteriors are measured against.
 *
 * Parity pins (tests/test_oracle_*.py): band geometry vs test_bands, split
 * points vs test_getSplitPoints, AGCG/AGTTCG posterior set vs
 * test_diagonalDPCalculations, forward==backward totals, brute-force path
 * enumeration on tiny inputs, removeOverlap crop (6,2).
 *
 * Five-state machine (M, short X, short Y, long X, long Y): no reference
 * counterpart; the same driver with a generic 5x5 cell update, pinned by
 * brute-force path enumeration, by forward == backward totals, by the
 * per-x / per-y posterior sums of a single-window global alignment, and by
 * reducing to the 3-state machine when the long-gap states are unreachable.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

#define LOG_ZERO (-INFINITY)
#define PROB_ONE_INT 10000000   /* PAIR_ALIGNMENT_PROB_1, pairwiseAligner.h:22 */

enum { ST_M = 0, ST_X = 1, ST_Y = 2 };
enum { S5_M = 0, S5_XS = 1, S5_YS = 2, S5_XL = 3, S5_YL = 4 };
#define MAX_NS 5

static inline int nstates(const or_model *m) { return m->nStates == 5 ? 5 : 3; }
/* state kind: 0 match (from (x-1, y-1)), 1 gap X (from (x-1, y)), 2 gap Y (from (x, y-1)) */
static inline int kind5(int s) { return s == 0 ? 0 : ((s & 1) ? 1 : 2); }

/* ------------------------------------------------------------------------ */
/* logAdd (pairwiseAligner.c:279-299)                                        */
/* ------------------------------------------------------------------------ */

/* The reference evaluates the cubic with float-suffixed literals in double
 * arithmetic; the (double)(float) casts reproduce exactly that rounding. */
#define FC(v) ((double)(float)(v))

static inline double poly_log1pexp(double x) {
    if (x <= FC(1.00))
        return ((FC(-0.009350833524763) * x + FC(0.130659527668286)) * x + FC(0.498799810682272)) * x + FC(0.693203116424741);
    if (x <= FC(2.50))
        return ((FC(-0.014532321752540) * x + FC(0.139942324101744)) * x + FC(0.495635523139337)) * x + FC(0.692140569840976);
    if (x <= FC(4.50))
        return ((FC(-0.004605031767994) * x + FC(0.063427417320019)) * x + FC(0.695956496475118)) * x + FC(0.514272634594009);
    return ((FC(-0.000458661602210) * x + FC(0.009695946122598)) * x + FC(0.930734667215156)) * x + FC(0.168037164329057);
}

static inline double logadd_poly(double x, double y) {
    if (x < y)
        return (x == LOG_ZERO || y - x >= 7.5) ? y : poly_log1pexp(y - x) + x;
    return (y == LOG_ZERO || x - y >= 7.5) ? x : poly_log1pexp(x - y) + y;
}

static inline double logadd_exact(double x, double y) {
    if (x < y) {
        if (x == LOG_ZERO) return y;
        return y + log1p(exp(x - y));
    }
    if (y == LOG_ZERO) return x;
    return x + log1p(exp(y - x));
}

double or_logadd(double x, double y, int mode) {
    return mode == OR_LOGADD_EXACT ? logadd_exact(x, y) : logadd_poly(x, y);
}

/* ------------------------------------------------------------------------ */
/* Emissions (stateMachine.c:360-380, 702-736; polisher.c:1756-1775)         */
/* ------------------------------------------------------------------------ */

typedef struct {
    const or_model *m;
    int mode;
} Ctx;

static inline double LADD(const Ctx *c, double x, double y) {
    return c->mode == OR_LOGADD_EXACT ? logadd_exact(x, y) : logadd_poly(x, y);
}

static inline double gap_prob(const double *g, int base) {
    return base >= 4 ? -1.386294361 : g[base];
}

static inline double emit_gap_x(const or_model *m, uint16_t s) {
    return gap_prob(m->gapX, m->rle ? (s & 255) : s);
}

static inline double emit_gap_y(const or_model *m, uint16_t s) {
    return gap_prob(m->gapY, m->rle ? (s & 255) : s);
}

static inline double emit_match(const or_model *m, uint16_t sx, uint16_t sy) {
    if (!m->rle) {
        if (sx >= 4 || sy >= 4) return -2.772588722;
        return m->match[sx * 4 + sy];
    }
    int xb = sx & 255, yb = sy & 255;
    double e = (xb >= 4 || yb >= 4) ? -2.772588722 : m->match[xb * 4 + yb];
    /* repeatSubMatrix_getLogProb(base=xb, strand, observed=rl(y), underlying=rl(x)) */
    int b = xb == 4 ? 0 : xb;
    int row = m->strand ? b : 3 - b;
    int R = m->maxRepeat;
    double lp = m->repeatLog10[(int64_t)row * R * R + (int64_t)(sx >> 8) * R + (sy >> 8)];
    return e + 2.3025 * lp;
}

/* ------------------------------------------------------------------------ */
/* Band geometry (pairwiseAligner.c:22-269)                                  */
/* ------------------------------------------------------------------------ */

static inline int64_t xcoord(int64_t xay, int64_t xmy) { return (xay + xmy) / 2; }
static inline int64_t ycoord(int64_t xay, int64_t xmy) { return (xay - xmy) / 2; }

static int64_t fix_parity(int64_t xay, int64_t xmy) {
    return (xay + xmy) % 2 == 0 ? xmy : xmy + 1;
}

static void push_if_less(int64_t *xmy, int64_t i, int64_t j, int64_t k) {
    if (i < j) *xmy += 2 * (j - i) * k;
}

static int band_diag(int64_t xay, int64_t xL, int64_t yL, int64_t xU, int64_t yU,
                     int64_t *outL, int64_t *outR) {
    int64_t l = fix_parity(xay, xL - yL);
    int64_t r = fix_parity(xay, xU - yU);
    push_if_less(&l, xcoord(xay, l), xL, 1);
    push_if_less(&l, yL, ycoord(xay, l), 1);
    push_if_less(&r, xU, xcoord(xay, r), -1);
    push_if_less(&r, ycoord(xay, r), yU, -1);
    *outL = l;
    *outR = r;
    if ((xay + l) % 2 != 0 || (xay + r) % 2 != 0 || l > r) return -1;
    return 0;
}

static int64_t clampc(int64_t z, int64_t lz) { return z < 0 ? 0 : (z > lz ? lz : z); }

/* anchors: n triples (x, y, expansion), sequence coordinates.
 * dynamic != 0: band_constructDynamic (120-173), else band_construct(expansion). */
int or_band(const int64_t *anchors, int64_t n, int64_t lX, int64_t lY,
            int64_t expansion, int dynamic, int64_t *xmyL, int64_t *xmyR) {
    int64_t D = lX + lY;
    int64_t ai = 0, xay = 0, pxay = 0, pxmy = 0, nxay = 0, nxmy = 0;
    int64_t xL = 0, yL = 0, xU = 0, yU = 0;
    int64_t e = dynamic ? 0 : expansion;
    while (xay <= D) {
        if (band_diag(xay, xL, yL, xU, yU, &xmyL[xay], &xmyR[xay]) != 0) return -1;
        if (nxay == xay++) {
            pxay = nxay;
            pxmy = nxmy;
            int64_t x = lX, y = lY;
            if (ai < n) {
                x = anchors[3 * ai] + 1;
                y = anchors[3 * ai + 1] + 1;
                if (dynamic) e = anchors[3 * ai + 2];
                ai++;
            }
            nxay = x + y;
            nxmy = x - y;
            xL = clampc(xcoord(pxay, pxmy - e), lX);
            yL = clampc(ycoord(nxay, nxmy - e), lY);
            xU = clampc(xcoord(nxay, nxmy + e), lX);
            yU = clampc(ycoord(pxay, pxmy + e), lY);
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Split points (pairwiseAligner.c:875-926)                                  */
/* ------------------------------------------------------------------------ */

static int split_step(int64_t *x1, int64_t *y1, int64_t x2, int64_t y2, int64_t x3, int64_t y3,
                      int64_t *out, int64_t *nout, int64_t cap, int64_t maxArea, int skip) {
    int64_t lx = x3 - x2, ly = y3 - y2;
    if (lx * ly > maxArea) {
        int64_t maxLen = (int64_t)sqrt((double)maxArea);
        int64_t hX = lx / 2 > maxLen ? maxLen : lx / 2;
        int64_t hY = ly / 2 > maxLen ? maxLen : ly / 2;
        if (!skip) {
            if (*nout < cap) {
                out[4 * *nout] = *x1; out[4 * *nout + 1] = *y1;
                out[4 * *nout + 2] = x2 + hX; out[4 * *nout + 3] = y2 + hY;
            }
            (*nout)++;
        }
        *x1 = x3 - hX;
        *y1 = y3 - hY;
        return 1;
    }
    return 0;
}

int64_t or_split_points(const int64_t *anchors, int64_t n, int64_t lX, int64_t lY,
                        int64_t maxArea, int raggedL, int raggedR, int64_t *out, int64_t cap) {
    int64_t x1 = 0, y1 = 0, x2 = 0, y2 = 0, nout = 0;
    for (int64_t i = 0; i < n; i++) {
        int64_t x3 = anchors[3 * i], y3 = anchors[3 * i + 1];
        split_step(&x1, &y1, x2, y2, x3, y3, out, &nout, cap, maxArea, raggedL && i == 0);
        x2 = x3 + 1;
        y2 = y3 + 1;
    }
    if (!split_step(&x1, &y1, x2, y2, lX, lY, out, &nout, cap, maxArea, raggedL && n == 0) || !raggedR) {
        if (nout < cap) {
            out[4 * nout] = x1; out[4 * nout + 1] = y1; out[4 * nout + 2] = lX; out[4 * nout + 3] = lY;
        }
        nout++;
    }
    return nout;
}

/* ------------------------------------------------------------------------ */
/* DP storage: one row of w*3 doubles per anti-diagonal, as DpDiagonal       */
/* (pairwiseAligner.c:379-517); rows live for the whole sub-problem.         */
/* ------------------------------------------------------------------------ */

typedef struct {
    int ns;                 /* states per cell, 3 or 5 */
    int64_t D;              /* last anti-diagonal index (lX + lY) */
    const int64_t *L, *R;   /* band */
    int64_t *off;           /* cell offset of each diagonal */
    double *F, *B;
    unsigned char *hasF, *hasB;
} Dp;

static double *cellp(const Dp *dp, double *base, int64_t xay, int64_t xmy) {
    if (xay < 0 || xay > dp->D) return NULL;
    if (xmy < dp->L[xay] || xmy > dp->R[xay]) return NULL;
    return base + dp->ns * (dp->off[xay] + (xmy - dp->L[xay]) / 2);
}

static inline uint16_t xsym(const uint16_t *sX, int64_t xay, int64_t xmy) {
    int64_t x = xcoord(xay, xmy);
    return x > 0 ? sX[x - 1] : 4;
}

static inline uint16_t ysym(const uint16_t *sY, int64_t xay, int64_t xmy) {
    int64_t y = ycoord(xay, xmy);
    return y > 0 ? sY[y - 1] : 4;
}

/* stateMachine3_cellCalculate (stateMachine.c:559-581) with doTransitionForward
 * (pairwiseAligner.c:311-314). */
/* Five-state cell update: each target state t reads the cell its kind steps
 * from, with that kind's emission, through T5[s][t] from every state s. */
static void cell_forward5(const Ctx *c, double *cur, const double *lo, const double *mid, const double *up,
                          uint16_t cx, uint16_t cy) {
    const or_model *m = c->m;
    for (int t = 0; t < 5; t++) {
        const int k = kind5(t);
        const double *src = k == 0 ? mid : (k == 1 ? lo : up);
        if (!src) continue;
        const double e = k == 0 ? emit_match(m, cx, cy) : (k == 1 ? emit_gap_x(m, cx) : emit_gap_y(m, cy));
        for (int s = 0; s < 5; s++) cur[t] = LADD(c, cur[t], src[s] + (e + m->trans5[s * 5 + t]));
    }
}

static void cell_backward5(const Ctx *c, const double *cur, double *lo, double *mid, double *up,
                           uint16_t cx, uint16_t cy) {
    const or_model *m = c->m;
    for (int t = 0; t < 5; t++) {
        const int k = kind5(t);
        double *src = k == 0 ? mid : (k == 1 ? lo : up);
        if (!src) continue;
        const double e = k == 0 ? emit_match(m, cx, cy) : (k == 1 ? emit_gap_x(m, cx) : emit_gap_y(m, cy));
        for (int s = 0; s < 5; s++) src[s] = LADD(c, src[s], cur[t] + (e + m->trans5[s * 5 + t]));
    }
}

static void cell_forward(const Ctx *c, double *cur, const double *lo, const double *mid, const double *up,
                         uint16_t cx, uint16_t cy) {
    const or_model *m = c->m;
    if (m->nStates == 5) {
        cell_forward5(c, cur, lo, mid, up, cx, cy);
        return;
    }
    if (lo) {
        double e = emit_gap_x(m, cx);
        cur[ST_X] = LADD(c, cur[ST_X], lo[ST_M] + (e + m->gapOpenX));
        cur[ST_X] = LADD(c, cur[ST_X], lo[ST_X] + (e + m->gapExtendX));
        cur[ST_X] = LADD(c, cur[ST_X], lo[ST_Y] + (e + m->gapSwitchToX));
    }
    if (mid) {
        double e = emit_match(m, cx, cy);
        cur[ST_M] = LADD(c, cur[ST_M], mid[ST_M] + (e + m->matchContinue));
        cur[ST_M] = LADD(c, cur[ST_M], mid[ST_X] + (e + m->matchFromGapX));
        cur[ST_M] = LADD(c, cur[ST_M], mid[ST_Y] + (e + m->matchFromGapY));
    }
    if (up) {
        double e = emit_gap_y(m, cy);
        cur[ST_Y] = LADD(c, cur[ST_Y], up[ST_M] + (e + m->gapOpenY));
        cur[ST_Y] = LADD(c, cur[ST_Y], up[ST_Y] + (e + m->gapExtendY));
        cur[ST_Y] = LADD(c, cur[ST_Y], up[ST_X] + (e + m->gapSwitchToY));
    }
}

/* Same cell update with doTransitionBackward (pairwiseAligner.c:321-324). */
static void cell_backward(const Ctx *c, const double *cur, double *lo, double *mid, double *up,
                          uint16_t cx, uint16_t cy) {
    const or_model *m = c->m;
    if (m->nStates == 5) {
        cell_backward5(c, cur, lo, mid, up, cx, cy);
        return;
    }
    if (lo) {
        double e = emit_gap_x(m, cx);
        lo[ST_M] = LADD(c, lo[ST_M], cur[ST_X] + (e + m->gapOpenX));
        lo[ST_X] = LADD(c, lo[ST_X], cur[ST_X] + (e + m->gapExtendX));
        lo[ST_Y] = LADD(c, lo[ST_Y], cur[ST_X] + (e + m->gapSwitchToX));
    }
    if (mid) {
        double e = emit_match(m, cx, cy);
        mid[ST_M] = LADD(c, mid[ST_M], cur[ST_M] + (e + m->matchContinue));
        mid[ST_X] = LADD(c, mid[ST_X], cur[ST_M] + (e + m->matchFromGapX));
        mid[ST_Y] = LADD(c, mid[ST_Y], cur[ST_M] + (e + m->matchFromGapY));
    }
    if (up) {
        double e = emit_gap_y(m, cy);
        up[ST_M] = LADD(c, up[ST_M], cur[ST_Y] + (e + m->gapOpenY));
        up[ST_Y] = LADD(c, up[ST_Y], cur[ST_Y] + (e + m->gapExtendY));
        up[ST_X] = LADD(c, up[ST_X], cur[ST_Y] + (e + m->gapSwitchToY));
    }
}

/* Boundary vectors (stateMachine.c:518-557).  Five states: start in M;
 * ragged start in any gap state; end through each state's transition to M;
 * ragged end: M at the mean of the short-gap open logs, each gap state at its
 * own extension. */
static void start_vec(const or_model *m, int ragged, double *v) {
    const int ns = nstates(m);
    for (int s = 0; s < ns; s++) v[s] = ragged ? (s == 0 ? LOG_ZERO : 0.0) : (s == 0 ? 0.0 : LOG_ZERO);
}

static void end_vec(const or_model *m, int ragged, double *v) {
    if (m->nStates == 5) {
        for (int s = 0; s < 5; s++) {
            if (ragged) v[s] = s == 0 ? (m->trans5[S5_XS] + m->trans5[S5_YS]) / 2.0 : m->trans5[s * 5 + s];
            else v[s] = m->trans5[s * 5 + S5_M];
        }
        return;
    }
    if (ragged) {
        v[0] = (m->gapOpenX + m->gapOpenY) / 2.0;
        v[1] = m->gapExtendX;
        v[2] = m->gapExtendY;
    } else {
        v[0] = m->matchContinue;
        v[1] = m->matchFromGapX;
        v[2] = m->matchFromGapY;
    }
}

static void diag_fill(const Dp *dp, double *base, int64_t xay, const double *v) {
    int64_t w = (dp->R[xay] - dp->L[xay]) / 2 + 1;
    double *p = base + dp->ns * dp->off[xay];
    for (int64_t i = 0; i < w; i++)
        for (int s = 0; s < dp->ns; s++) p[dp->ns * i + s] = v[s];
}

static void diag_zero(const Dp *dp, double *base, int64_t xay) {
    double z[MAX_NS] = {LOG_ZERO, LOG_ZERO, LOG_ZERO, LOG_ZERO, LOG_ZERO};
    diag_fill(dp, base, xay, z);
}

/* diagonalCalculationForward (pairwiseAligner.c:540-560) */
static void diag_forward(const Ctx *c, Dp *dp, const uint16_t *sX, const uint16_t *sY, int64_t xay) {
    for (int64_t xmy = dp->L[xay]; xmy <= dp->R[xay]; xmy += 2) {
        double *cur = cellp(dp, dp->F, xay, xmy);
        const double *lo = (xay >= 1 && dp->hasF[xay - 1]) ? cellp(dp, dp->F, xay - 1, xmy - 1) : NULL;
        const double *mid = (xay >= 2 && dp->hasF[xay - 2]) ? cellp(dp, dp->F, xay - 2, xmy) : NULL;
        const double *up = (xay >= 1 && dp->hasF[xay - 1]) ? cellp(dp, dp->F, xay - 1, xmy + 1) : NULL;
        cell_forward(c, cur, lo, mid, up, xsym(sX, xay, xmy), ysym(sY, xay, xmy));
    }
}

/* diagonalCalculationBackward (pairwiseAligner.c:562-565) */
static void diag_backward(const Ctx *c, Dp *dp, const uint16_t *sX, const uint16_t *sY, int64_t xay) {
    for (int64_t xmy = dp->L[xay]; xmy <= dp->R[xay]; xmy += 2) {
        double *cur = cellp(dp, dp->B, xay, xmy);
        double *lo = (xay >= 1 && dp->hasB[xay - 1]) ? cellp(dp, dp->B, xay - 1, xmy - 1) : NULL;
        double *mid = (xay >= 2 && dp->hasB[xay - 2]) ? cellp(dp, dp->B, xay - 2, xmy) : NULL;
        double *up = (xay >= 1 && dp->hasB[xay - 1]) ? cellp(dp, dp->B, xay - 1, xmy + 1) : NULL;
        cell_backward(c, cur, lo, mid, up, xsym(sX, xay, xmy), ysym(sY, xay, xmy));
    }
}

static double cell_dot(const Ctx *c, const double *a, const double *b) {
    double t = a[0] + b[0];
    t = LADD(c, t, a[1] + b[1]);
    t = LADD(c, t, a[2] + b[2]);
    if (c->m->nStates == 5) {
        t = LADD(c, t, a[3] + b[3]);
        t = LADD(c, t, a[4] + b[4]);
    }
    return t;
}

static double diag_dot(const Ctx *c, const Dp *dp, const double *A, int64_t xa, const double *Bv, int64_t xb) {
    double t = LOG_ZERO;
    for (int64_t xmy = dp->L[xa]; xmy <= dp->R[xa]; xmy += 2) {
        const double *a = A + dp->ns * (dp->off[xa] + (xmy - dp->L[xa]) / 2);
        const double *b = Bv + dp->ns * (dp->off[xb] + (xmy - dp->L[xb]) / 2);
        t = LADD(c, t, cell_dot(c, a, b));
    }
    return t;
}

/* diagonalCalculationTotalProbability (pairwiseAligner.c:567-584) */
static double diag_total(const Ctx *c, Dp *dp, const uint16_t *sX, const uint16_t *sY, int64_t xay) {
    double t = diag_dot(c, dp, dp->F, xay, dp->B, xay);
    if (xay + 1 <= dp->D && dp->hasB[xay + 1] && xay >= 1 && dp->hasF[xay - 1]) {
        int64_t d1 = xay + 1;
        int64_t w = (dp->R[d1] - dp->L[d1]) / 2 + 1;
        double *mt = (double *)malloc(sizeof(double) * dp->ns * (size_t)w);
        for (int64_t i = 0; i < dp->ns * w; i++) mt[i] = LOG_ZERO;
        double acc = LOG_ZERO;
        for (int64_t xmy = dp->L[d1], i = 0; xmy <= dp->R[d1]; xmy += 2, i++) {
            const double *mid = cellp(dp, dp->F, xay - 1, xmy);
            cell_forward(c, mt + dp->ns * i, NULL, mid, NULL, xsym(sX, d1, xmy), ysym(sY, d1, xmy));
        }
        for (int64_t xmy = dp->L[d1], i = 0; xmy <= dp->R[d1]; xmy += 2, i++) {
            const double *b = cellp(dp, dp->B, d1, xmy);
            acc = LADD(c, acc, cell_dot(c, mt + dp->ns * i, b));
        }
        free(mt);
        t = LADD(c, t, acc);
    }
    return t;
}

/* ------------------------------------------------------------------------ */
/* Output lists                                                              */
/* ------------------------------------------------------------------------ */

typedef struct {
    int64_t *v;   /* triples (weight, x, y) */
    int64_t n, cap;
} List;

static void list_push(List *l, int64_t w, int64_t x, int64_t y) {
    if (l->n == l->cap) {
        l->cap = l->cap ? 2 * l->cap : 64;
        l->v = (int64_t *)realloc(l->v, sizeof(int64_t) * 3 * (size_t)l->cap);
    }
    l->v[3 * l->n] = w; l->v[3 * l->n + 1] = x; l->v[3 * l->n + 2] = y;
    l->n++;
}

#define OR_EMIT_N (16 + 4 + 4 + 4 * 51 * 51)

struct or_result {
    List lists[3];        /* match, gapX, gapY */
    double expect[9];     /* transition expectations T[from*3+to] */
    double emit[OR_EMIT_N]; /* expected emission counts, forward-strand parameters: match[x*4+y],
                               gapX[4], gapY[4], repeat[base][underlying][observed] */
    double likelihood;    /* per-diagonal accumulated total (expectation hack) */
    double totalLogProb;  /* sum over sub-problems of the last window's total (forward log-likelihood) */
    int64_t cells;        /* band cells of all sub-problems */
    int status;
};

/* addPosteriorProb (pairwiseAligner.c:586-595) */
static void add_post(List *l, int64_t x, int64_t y, double p, double thr) {
    if (p >= thr) {
        if (p > 1.0) p = 1.0;
        p = floor(p * PROB_ONE_INT);
        list_push(l, (int64_t)p, x - 1, y - 1);
    }
}

enum { PF_MATCH_ONLY = 0, PF_ALL = 1, PF_EXPECT = 2 };

/* diagonalCalculationPosteriorMatchProbs / diagonalCalculationPosteriorProbs /
 * diagonalCalculationExpectations (pairwiseAligner.c:597-677) */
static void diag_posterior(const Ctx *c, Dp *dp, const uint16_t *sX, const uint16_t *sY, int64_t xay,
                           double total, double thr, int kind, List *sub, or_result *res) {
    if (kind == PF_EXPECT) {
        const or_model *m = c->m;
        res->likelihood += total;
        for (int64_t xmy = dp->L[xay]; xmy <= dp->R[xay]; xmy += 2) {
            const double *cur = cellp(dp, dp->B, xay, xmy);
            const double *lo = (xay >= 1 && dp->hasF[xay - 1]) ? cellp(dp, dp->F, xay - 1, xmy - 1) : NULL;
            const double *mid = (xay >= 2 && dp->hasF[xay - 2]) ? cellp(dp, dp->F, xay - 2, xmy) : NULL;
            const double *up = (xay >= 1 && dp->hasF[xay - 1]) ? cellp(dp, dp->F, xay - 1, xmy + 1) : NULL;
            uint16_t cx = xsym(sX, xay, xmy), cy = ysym(sY, xay, xmy);
            /* updateExpectations (pairwiseAligner.c:347-369) in cellCalculate order */
#define EXPT(from_, to_, fromv, tov, e, t) res->expect[(from_) * 3 + (to_)] += exp((fromv) + (tov) + ((e) + (t)) - total)
            if (lo) {
                double e = emit_gap_x(m, cx);
                EXPT(ST_M, ST_X, lo[ST_M], cur[ST_X], e, m->gapOpenX);
                EXPT(ST_X, ST_X, lo[ST_X], cur[ST_X], e, m->gapExtendX);
                EXPT(ST_Y, ST_X, lo[ST_Y], cur[ST_X], e, m->gapSwitchToX);
            }
            if (mid) {
                double e = emit_match(m, cx, cy);
                EXPT(ST_M, ST_M, mid[ST_M], cur[ST_M], e, m->matchContinue);
                EXPT(ST_X, ST_M, mid[ST_X], cur[ST_M], e, m->matchFromGapX);
                EXPT(ST_Y, ST_M, mid[ST_Y], cur[ST_M], e, m->matchFromGapY);
            }
            if (up) {
                double e = emit_gap_y(m, cy);
                EXPT(ST_M, ST_Y, up[ST_M], cur[ST_Y], e, m->gapOpenY);
                EXPT(ST_Y, ST_Y, up[ST_Y], cur[ST_Y], e, m->gapExtendY);
                EXPT(ST_X, ST_Y, up[ST_X], cur[ST_Y], e, m->gapSwitchToY);
            }
#undef EXPT
            /* expected emission counts: the posterior of each state at this cell,
             * on the forward-strand parameter that state's emission reads (no
             * reference counterpart: emission expectations are disabled at
             * pairwiseAligner.c:359-362) */
            const double *f = cellp(dp, dp->F, xay, xmy);
            int64_t x = xcoord(xay, xmy), y = ycoord(xay, xmy);
            int st = m->strand;
            int bx = m->rle ? (cx & 255) : cx, by = m->rle ? (cy & 255) : cy;
            double pM = exp((f[ST_M] + cur[ST_M]) - total);
            double pX = exp((f[ST_X] + cur[ST_X]) - total);
            double pY = exp((f[ST_Y] + cur[ST_Y]) - total);
            if (x > 0 && y > 0) {
                if (bx < 4 && by < 4) res->emit[st ? bx * 4 + by : (3 - bx) * 4 + (3 - by)] += pM;
                if (m->rle) {
                    int b = bx >= 4 ? 0 : bx;
                    int row = st ? b : 3 - b;
                    res->emit[24 + (row * 51 + (cx >> 8)) * 51 + (cy >> 8)] += pM;
                }
            }
            if (x > 0 && bx < 4) res->emit[16 + (st ? bx : 3 - bx)] += pX;
            if (y > 0 && by < 4) res->emit[20 + (st ? by : 3 - by)] += pY;
        }
        return;
    }
    for (int64_t xmy = dp->L[xay]; xmy <= dp->R[xay]; xmy += 2) {
        int64_t x = xcoord(xay, xmy), y = ycoord(xay, xmy);
        const double *f = cellp(dp, dp->F, xay, xmy);
        const double *b = cellp(dp, dp->B, xay, xmy);
        if (x > 0 && y > 0) add_post(&sub[0], x, y, exp((f[ST_M] + b[ST_M]) - total), thr);
        if (kind == PF_ALL && dp->ns == 5) {
            /* gap posteriors of the five-state machine: short + long */
            if (x > 0)
                add_post(&sub[1], x, y, exp((f[S5_XS] + b[S5_XS]) - total) + exp((f[S5_XL] + b[S5_XL]) - total), thr);
            if (y > 0)
                add_post(&sub[2], x, y, exp((f[S5_YS] + b[S5_YS]) - total) + exp((f[S5_YL] + b[S5_YL]) - total), thr);
        } else if (kind == PF_ALL) {
            if (x > 0) add_post(&sub[1], x, y, exp((f[ST_X] + b[ST_X]) - total), thr);
            if (y > 0) add_post(&sub[2], x, y, exp((f[ST_Y] + b[ST_Y]) - total), thr);
        }
    }
}

/* getPosteriorProbsWithBanding (pairwiseAligner.c:687-808) */
static int banded_fb(const Ctx *c, const int64_t *anchors, int64_t n, const uint16_t *sX, int64_t lX,
                     const uint16_t *sY, int64_t lY, const or_params *p, int raggedL, int raggedR,
                     int kind, List *sub, or_result *res) {
    int64_t D = lX + lY;
    if (D == 0) return 0;
    int64_t *L = (int64_t *)malloc(sizeof(int64_t) * (size_t)(D + 1));
    int64_t *R = (int64_t *)malloc(sizeof(int64_t) * (size_t)(D + 1));
    if (or_band(anchors, n, lX, lY, p->diagonalExpansion, p->dynamicAnchorExpansion, L, R) != 0) {
        free(L); free(R);
        return OR_EBAND;
    }
    Dp dp;
    dp.D = D; dp.L = L; dp.R = R;
    dp.off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(D + 2));
    int64_t tot = 0;
    for (int64_t d = 0; d <= D; d++) { dp.off[d] = tot; tot += (R[d] - L[d]) / 2 + 1; }
    dp.off[D + 1] = tot;
    res->cells += tot;
    dp.ns = nstates(c->m);
    dp.F = (double *)malloc(sizeof(double) * dp.ns * (size_t)tot);
    dp.B = (double *)malloc(sizeof(double) * dp.ns * (size_t)tot);
    dp.hasF = (unsigned char *)calloc((size_t)D + 1, 1);
    dp.hasB = (unsigned char *)calloc((size_t)D + 1, 1);

    double v[MAX_NS];
    start_vec(c->m, raggedL, v);
    diag_fill(&dp, dp.F, 0, v);
    dp.hasF[0] = 1;

    int64_t tracedBackTo = 0;
    for (int64_t xay = 1;; xay++) {
        diag_zero(&dp, dp.F, xay);
        dp.hasF[xay] = 1;
        diag_forward(c, &dp, sX, sY, xay);
        int atEnd = xay == D;
        int64_t width = (R[xay] - L[xay]) / 2 + 1;
        int tbPoint = xay >= tracedBackTo + p->minDiagsBetweenTraceBack && width <= p->diagonalExpansion * 2 + 1;
        if (atEnd || tbPoint) {
            end_vec(c->m, atEnd && raggedR, v);
            diag_fill(&dp, dp.B, xay, v);
            dp.hasB[xay] = 1;
            if (xay > tracedBackTo + 1) { diag_zero(&dp, dp.B, xay - 1); dp.hasB[xay - 1] = 1; }
            int64_t from = xay - (atEnd ? 0 : p->traceBackDiagonals + 1);
            double total = LOG_ZERO;
            int64_t count = 0;
            for (int64_t d2 = xay; d2 > tracedBackTo; d2--) {
                if (d2 > tracedBackTo + 2) { diag_zero(&dp, dp.B, d2 - 2); dp.hasB[d2 - 2] = 1; }
                if (d2 > tracedBackTo + 1) diag_backward(c, &dp, sX, sY, d2);
                if (d2 <= from) {
                    if (count++ % 10 == 0) {
                        total = diag_total(c, &dp, sX, sY, d2);
                        if (count == 1 && atEnd) res->totalLogProb += total;
                    }
                    diag_posterior(c, &dp, sX, sY, d2, total, p->threshold, kind, sub, res);
                    if (d2 < from || atEnd) dp.hasF[d2] = 0;
                }
                if (d2 + 1 <= D) dp.hasB[d2 + 1] = 0;
            }
            dp.hasB[tracedBackTo + 1] = 0;
            dp.hasF[tracedBackTo] = 0;
            tracedBackTo = from;
        }
        if (atEnd) break;
    }
    free(dp.F); free(dp.B); free(dp.hasF); free(dp.hasB); free(dp.off); free(L); free(R);
    return 0;
}

/* getPosteriorProbsWithBandingSplittingAlignmentsByLargeGaps (942-991) plus
 * the pop/append coordinate correction of 1068-1086. */
static or_result *run_split(const or_model *m, const uint16_t *sX, int64_t lX, const uint16_t *sY, int64_t lY,
                            const int64_t *anchors, int64_t n, const or_params *p, int raggedL, int raggedR,
                            int kind, int mode) {
    Ctx c = {m, mode};
    or_result *res = (or_result *)calloc(1, sizeof(or_result));
    if (kind == PF_EXPECT && nstates(m) == 5) {
        res->status = OR_EUNSUPPORTED;  /* expected counts: 3-state machine only */
        return res;
    }
    int64_t cap = 16, ns;
    int64_t *sp = NULL;
    for (;;) {
        sp = (int64_t *)realloc(sp, sizeof(int64_t) * 4 * (size_t)cap);
        ns = or_split_points(anchors, n, lX, lY, p->splitMatrixBiggerThanThis, raggedL, raggedR, sp, cap);
        if (ns <= cap) break;
        cap = ns;
    }
    int64_t j = 0;
    int64_t *sub_anchors = (int64_t *)malloc(sizeof(int64_t) * 3 * (size_t)(n > 0 ? n : 1));
    List sub[3];
    memset(sub, 0, sizeof(sub));
    for (int64_t i = 0; i < ns && res->status == 0; i++) {
        int64_t x1 = sp[4 * i], y1 = sp[4 * i + 1], x2 = sp[4 * i + 2], y2 = sp[4 * i + 3];
        int64_t k = 0;
        while (j < n) {
            int64_t x = anchors[3 * j], y = anchors[3 * j + 1];
            if (x + y >= x2 + y2) break;
            sub_anchors[3 * k] = x - x1; sub_anchors[3 * k + 1] = y - y1; sub_anchors[3 * k + 2] = anchors[3 * j + 2];
            k++; j++;
        }
        int st = banded_fb(&c, sub_anchors, k, sX + x1, x2 - x1, sY + y1, y2 - y1, p,
                           raggedL || i > 0, raggedR || i < ns - 1, kind, sub, res);
        if (st) res->status = st;
        if (kind != PF_EXPECT) {
            for (int li = 0; li < 3; li++) {
                while (sub[li].n > 0) {
                    sub[li].n--;
                    int64_t *t = sub[li].v + 3 * sub[li].n;
                    list_push(&res->lists[li], t[0], t[1] + x1, t[2] + y1);
                }
            }
        }
    }
    for (int li = 0; li < 3; li++) free(sub[li].v);
    free(sub_anchors);
    free(sp);
    return res;
}

or_result *or_posteriors(const or_model *m, const uint16_t *sX, int64_t lX, const uint16_t *sY, int64_t lY,
                         const int64_t *anchors, int64_t n, const or_params *p, int raggedL, int raggedR,
                         int withIndels, int mode) {
    return run_split(m, sX, lX, sY, lY, anchors, n, p, raggedL, raggedR, withIndels ? PF_ALL : PF_MATCH_ONLY, mode);
}

or_result *or_expectations(const or_model *m, const uint16_t *sX, int64_t lX, const uint16_t *sY, int64_t lY,
                           const int64_t *anchors, int64_t n, const or_params *p, int raggedL, int raggedR, int mode) {
    return run_split(m, sX, lX, sY, lY, anchors, n, p, raggedL, raggedR, PF_EXPECT, mode);
}

/* computeForwardProbability (pairwiseAligner.c:813-865): static band with
 * p->diagonalExpansion, full forward, total at the last diagonal. */
double or_forward_prob(const or_model *m, const uint16_t *sX, int64_t lX, const uint16_t *sY, int64_t lY,
                       const int64_t *anchors, int64_t n, const or_params *p, int raggedL, int raggedR, int mode) {
    Ctx c = {m, mode};
    int64_t D = lX + lY;
    if (D == 0) return 0.0;
    int64_t *L = (int64_t *)malloc(sizeof(int64_t) * (size_t)(D + 1));
    int64_t *R = (int64_t *)malloc(sizeof(int64_t) * (size_t)(D + 1));
    if (or_band(anchors, n, lX, lY, p->diagonalExpansion, 0, L, R) != 0) { free(L); free(R); return NAN; }
    Dp dp;
    dp.D = D; dp.L = L; dp.R = R;
    dp.off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(D + 2));
    int64_t tot = 0;
    for (int64_t d = 0; d <= D; d++) { dp.off[d] = tot; tot += (R[d] - L[d]) / 2 + 1; }
    dp.off[D + 1] = tot;
    dp.ns = nstates(m);
    dp.F = (double *)malloc(sizeof(double) * dp.ns * (size_t)tot);
    dp.B = (double *)malloc(sizeof(double) * dp.ns * (size_t)tot);
    dp.hasF = (unsigned char *)calloc((size_t)D + 1, 1);
    dp.hasB = (unsigned char *)calloc((size_t)D + 1, 1);
    double v[MAX_NS];
    start_vec(m, raggedL, v);
    diag_fill(&dp, dp.F, 0, v);
    dp.hasF[0] = 1;
    for (int64_t xay = 1; xay <= D; xay++) {
        diag_zero(&dp, dp.F, xay);
        dp.hasF[xay] = 1;
        diag_forward(&c, &dp, sX, sY, xay);
    }
    end_vec(m, raggedR, v);
    diag_fill(&dp, dp.B, D, v);
    dp.hasB[D] = 1;
    double t = diag_total(&c, &dp, sX, sY, D);
    free(dp.F); free(dp.B); free(dp.hasF); free(dp.hasB); free(dp.off); free(L); free(R);
    return t;
}

/* Full-matrix forward and backward without windows, as the reference test
 * test_diagonalDPCalculations drives them (pairwiseAlignerTest.c:256-339):
 * returns forward total, backward total and the cut total of every diagonal. */
int or_full_dp(const or_model *m, const uint16_t *sX, int64_t lX, const uint16_t *sY, int64_t lY,
               int64_t expansion, int mode, double *fwdTotal, double *bwdTotal, double *cutTotals,
               double *postM /* (lX+1)*(lY+1) posterior match matrix, may be NULL */) {
    Ctx c = {m, mode};
    int64_t D = lX + lY;
    int64_t *L = (int64_t *)malloc(sizeof(int64_t) * (size_t)(D + 1));
    int64_t *R = (int64_t *)malloc(sizeof(int64_t) * (size_t)(D + 1));
    if (or_band(NULL, 0, lX, lY, expansion, 0, L, R) != 0) { free(L); free(R); return OR_EBAND; }
    Dp dp;
    dp.D = D; dp.L = L; dp.R = R;
    dp.off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(D + 2));
    int64_t tot = 0;
    for (int64_t d = 0; d <= D; d++) { dp.off[d] = tot; tot += (R[d] - L[d]) / 2 + 1; }
    dp.ns = nstates(m);
    dp.F = (double *)malloc(sizeof(double) * dp.ns * (size_t)tot);
    dp.B = (double *)malloc(sizeof(double) * dp.ns * (size_t)tot);
    dp.hasF = (unsigned char *)calloc((size_t)D + 1, 1);
    dp.hasB = (unsigned char *)calloc((size_t)D + 1, 1);
    for (int64_t d = 0; d <= D; d++) {
        diag_zero(&dp, dp.F, d); diag_zero(&dp, dp.B, d);
        dp.hasF[d] = dp.hasB[d] = 1;
    }
    double v[MAX_NS];
    start_vec(m, 0, v);
    diag_fill(&dp, dp.F, 0, v);
    end_vec(m, 0, v);
    diag_fill(&dp, dp.B, D, v);
    for (int64_t d = 1; d <= D; d++) diag_forward(&c, &dp, sX, sY, d);
    for (int64_t d = D; d > 0; d--) diag_backward(&c, &dp, sX, sY, d);
    double e[MAX_NS], s[MAX_NS];
    end_vec(m, 0, e);
    start_vec(m, 0, s);
    double *fc = cellp(&dp, dp.F, D, lX - lY);
    double *bc = cellp(&dp, dp.B, 0, 0);
    *fwdTotal = cell_dot(&c, fc, e);
    *bwdTotal = cell_dot(&c, bc, s);
    for (int64_t d = 0; d <= D; d++) cutTotals[d] = diag_total(&c, &dp, sX, sY, d);
    if (postM) {
        for (int64_t i = 0; i < (lX + 1) * (lY + 1); i++) postM[i] = 0.0;
        for (int64_t d = 1; d <= D; d++)
            for (int64_t xmy = L[d]; xmy <= R[d]; xmy += 2) {
                int64_t x = xcoord(d, xmy), y = ycoord(d, xmy);
                double *f = cellp(&dp, dp.F, d, xmy), *b = cellp(&dp, dp.B, d, xmy);
                postM[x * (lY + 1) + y] = exp(f[0] + b[0] - *fwdTotal);
            }
    }
    free(dp.F); free(dp.B); free(dp.hasF); free(dp.hasB); free(dp.off); free(L); free(R);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Result accessors                                                          */
/* ------------------------------------------------------------------------ */

int or_result_status(const or_result *r) { return r->status; }
int64_t or_result_count(const or_result *r, int list) { return r->lists[list].n; }
void or_result_copy(const or_result *r, int list, int64_t *out) {
    memcpy(out, r->lists[list].v, sizeof(int64_t) * 3 * (size_t)r->lists[list].n);
}
double or_result_total(const or_result *r) { return r->totalLogProb; }
double or_result_likelihood(const or_result *r) { return r->likelihood; }
int64_t or_result_cells(const or_result *r) { return r->cells; }
void or_result_expectations(const or_result *r, double *out9) { memcpy(out9, r->expect, sizeof(double) * 9); }
void or_result_emissions(const or_result *r, double *out) { memcpy(out, r->emit, sizeof(double) * OR_EMIT_N); }
void or_result_free(or_result *r) {
    if (!r) return;
    for (int i = 0; i < 3; i++) free(r->lists[i].v);
    free(r);
}

/* ------------------------------------------------------------------------ */
/* Batch driver for the timed CPU baseline: OpenMP over pairs, outputs are    */
/* counted and discarded (the per-pair entry returns them for parity).       */
/* ------------------------------------------------------------------------ */

int64_t or_batch_cells(const or_model *models /* [2]: reverse, forward strand */, int64_t npairs,
                       const uint16_t *symX, const int64_t *offX, const int64_t *lenX,
                       const uint16_t *symY, const int64_t *offY, const int64_t *lenY,
                       const int64_t *anchors, const int64_t *anchorOff, const int64_t *anchorN,
                       const int *strand, const or_params *p, int mode, int nthreads,
                       int64_t *outCounts /* npairs*3, may be NULL */) {
    int64_t cells = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : cells)
#endif
    for (int64_t i = 0; i < npairs; i++) {
        or_result *r = or_posteriors(&models[strand[i] ? 1 : 0], symX + offX[i], lenX[i], symY + offY[i], lenY[i],
                                     anchors + 3 * anchorOff[i], anchorN[i], p, 0, 0, 1, mode);
        cells += r->cells;
        if (outCounts)
            for (int l = 0; l < 3; l++) outCounts[3 * i + l] = r->lists[l].n;
        or_result_free(r);
    }
    return cells;
}
