/*
 * hdm_oracle.h - CPU oracle for the Hausdorff dilation path.  TEST
 * INFRASTRUCTURE ONLY: the test suite, the smoke check and the benchmark's CPU
 * baseline leg compare results against it.  The product library never links or
 * calls it.
 *
 * Two restatements of the reference algorithm (hello.cu, mainPassKernel,
 * 1066-1896):
 *   oracle_naive  - one byte per voxel, whole-volume 6-connectivity dilation
 *                   per iteration, first-hit recording exactly as
 *                   hello.cu:1590-1626; for small volumes.
 *   oracle_tiled  - the same algorithm the device engine runs: masks packed
 *                   into 64-bit x-words, 64x8x8 tiles with empty / active /
 *                   full state, a queue of the tiles still changing, first-hit
 *                   histograms; OpenMP over queued tiles.  The timed CPU
 *                   baseline.
 *
 * Both return the first-hit histograms hist_fp[t], hist_fn[t] (t = 1..) from
 * which every reported number follows (oracle_finish).
 */
#ifndef HDM_ORACLE_H
#define HDM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_result {
  int32_t hd, hd_fp, hd_fn, iterations;
  uint64_t sum, count, sum_fp, cnt_fp, sum_fn, cnt_fn;
  uint64_t n_gold, n_segm, n_fp, n_fn;
  double avg_hd;
  uint64_t queue_total; /* tiled only: sum of queued (tile, mask) entries */
  uint32_t queue_max;
  uint32_t tiles;
} oracle_result;

#define ORACLE_DTYPE_U8 0
#define ORACLE_DTYPE_I32 1

/* hist_len >= nx+ny+nz+2 is always enough.  map (optional, nx*ny*nz) receives
 * d(v) on covered FP/FN voxels.  Returns 0, -1 (invalid), -2 (empty mask),
 * -8 (stuck). */
int oracle_naive(const void *gold, const void *segm, int dtype, int32_t label, int nx, int ny,
                 int nz, double robustness, uint32_t *hist_fp, uint32_t *hist_fn,
                 int hist_len, uint16_t *map, oracle_result *out);

/* threads <= 0: OpenMP default.  crop: restrict dilation to the bounding box
 * of non-empty tiles (widened by one tile in x and y for the 18-neighbour
 * element, whose shortest paths can leave the box by one voxel there). */
int oracle_tiled(const void *gold, const void *segm, int dtype, int32_t label, int nx, int ny,
                 int nz, double robustness, int threads, int crop, uint32_t *hist_fp,
                 uint32_t *hist_fn, int hist_len, uint16_t *map, oracle_result *out);

/* The same two with a structuring element: conn 6 (the reference's cross,
 * hello.cu:373-376, 1310-1542; what the functions above use), 18 (the
 * reference's second engine with CHEBYSHEV 1: the six faces, the four
 * diagonals in the xy plane and the eight corners,
 * forBench/HausdorffDistance.cu:37-60) or 26 (the full 3x3x3 cube).
 * -1 for any other conn. */
int oracle_naive_se(const void *gold, const void *segm, int dtype, int32_t label, int nx,
                    int ny, int nz, double robustness, int conn, uint32_t *hist_fp,
                    uint32_t *hist_fn, int hist_len, uint16_t *map, oracle_result *out);
int oracle_tiled_se(const void *gold, const void *segm, int dtype, int32_t label, int nx,
                    int ny, int nz, double robustness, int conn, int threads, int crop,
                    uint32_t *hist_fp, uint32_t *hist_fn, int hist_len, uint16_t *map,
                    oracle_result *out);

/* Fill out (hd, sums, counts, avg) from the histograms and totals. */
void oracle_finish(const uint32_t *hist_fp, const uint32_t *hist_fn, int hist_len,
                   double robustness, oracle_result *out);

int oracle_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
