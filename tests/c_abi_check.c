/* A C caller of the library's C ABI, compiled with gcc against include/ by
 * tests/test_host_cpu.py: prints each ABI struct's size and field offsets
 * (compared there with the ctypes mirrors in marginpolish_amd/_lib.py) and
 * calls the host entry points that need no GPU (params loading, band, split
 * points, traceback windows). */
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>

#include "mp_bubble.h"
#include "mp_pairhmm.h"
#include "mp_poa.h"
#include "mp_stitch.h"

#define OFF(T, f) (long)offsetof(T, f)

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    printf("mp_pair %ld %ld %ld %ld %ld %ld %ld %ld %ld %ld\n", (long)sizeof(mp_pair), OFF(mp_pair, sX),
           OFF(mp_pair, lX), OFF(mp_pair, sY), OFF(mp_pair, lY), OFF(mp_pair, anchors), OFF(mp_pair, nAnchors),
           OFF(mp_pair, strand), OFF(mp_pair, raggedLeft), OFF(mp_pair, raggedRight));
    printf("mp_band_params %ld %ld %ld %ld %ld %ld %ld\n", (long)sizeof(mp_band_params),
           OFF(mp_band_params, threshold), OFF(mp_band_params, minDiagsBetweenTraceBack),
           OFF(mp_band_params, traceBackDiagonals), OFF(mp_band_params, diagonalExpansion),
           OFF(mp_band_params, splitMatrixBiggerThanThis), OFF(mp_band_params, dynamicAnchorExpansion));
    printf("mp_aligned_pair %ld %ld %ld %ld\n", (long)sizeof(mp_aligned_pair), OFF(mp_aligned_pair, weight),
           OFF(mp_aligned_pair, x), OFF(mp_aligned_pair, y));
    printf("mp_rle_string %ld %ld %ld %ld\n", (long)sizeof(mp_rle_string), OFF(mp_rle_string, bases),
           OFF(mp_rle_string, repeatCounts), OFF(mp_rle_string, length));
    printf("mp_poa_sizes %ld\n", (long)sizeof(mp_poa_sizes));
    printf("version %s\n", mp_version());

    mp_model *m = NULL;
    if (mp_model_load_json(argv[1], &m) != MP_OK) {
        printf("load failed: %s\n", mp_last_error());
        return 1;
    }
    mp_band_params bp;
    if (mp_model_band_params(m, &bp) != MP_OK) return 1;
    printf("params %g %ld %ld %ld %ld %d\n", bp.threshold, (long)bp.minDiagsBetweenTraceBack,
           (long)bp.traceBackDiagonals, (long)bp.diagonalExpansion, (long)bp.splitMatrixBiggerThanThis,
           bp.dynamicAnchorExpansion);

    /* band of a 6 x 5 problem with three anchors (test_bands' shape) */
    const int64_t anchors[] = {1, 0, 0, 2, 1, 0, 3, 3, 0};
    int32_t L[12], R[12];
    if (mp_band_construct(anchors, 3, 6, 5, 2, 0, L, R) != MP_OK) return 1;
    printf("band");
    for (int d = 0; d <= 11; d++) printf(" %d:%d", L[d], R[d]);
    printf("\n");

    /* split points of two distant anchor runs */
    const int64_t far[] = {10, 10, 0, 11, 11, 0, 5000, 5000, 0, 5001, 5001, 0};
    int64_t sp[64];
    const int64_t ns = mp_split_points(far, 4, 6000, 6000, 1000000, 0, 0, sp, 16);
    printf("splits %ld", (long)ns);
    for (int64_t i = 0; i < 4 * ns && i < 64; i++) printf(" %ld", (long)sp[i]);
    printf("\n");
    mp_model_free(m);
    return 0;
}
