/*
 * rtod_oracle.h -- CPU restatement of RTOD's per-slide outlier detection.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path and
 * the timed CPU baseline; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product path (rtod_amd/)
 * never links or calls it.
 *
 * Contract (shared with the HIP path):
 *   - a window point p is an OUTLIER iff  #{q in window : d2(p,q) <= r2} <= K,
 *     the count including p itself
 *     (reference: src/ods/optixScan.cu:158-199 counts the ray's own box,
 *      src/ods/optixScan.cpp:1095-1099 selects count <= K,
 *      src/ods/optixScan.cpp:817-838 brute-force checker, "include itself").
 *   - d2 is FP32:  dx = p.x - q.x; d2 = dx*dx; d2 = fmaf(dy,dy,d2);
 *     d2 = fmaf(dz,dz,d2);  r2 = R*R in FP32.  (The reference tests FP64
 *     `|dx| < R` / `sqdist < R2`, optixScan.cu:170-189; semantics mode 1 of
 *     rtod_o_brute restates that rule so disagreements can be counted.)
 *   - ring semantics: global index g lives in ring slot g mod W; slide t
 *     (t = 1, 2, ...) holds x[tS, tS+W)  (optixScan.cpp:1019-1126).
 *   - outliers are reported as ring-slot ids sorted ascending (the reference
 *     reports the same set in arbitrary order, optixScan.cpp:1097).
 *
 * Parity status of each function is stated next to it and in the tests that
 * pin it (tests/test_oracle.py).
 */
#ifndef RTOD_ORACLE_H
#define RTOD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* FP32 squared distance with explicit FMA (the contract predicate). */
float rtod_o_d2(const float *p, const float *q, int dim);

/* Brute force over one window `win` ([W][dim], slot order).
 * semantics 0: contract predicate (FP32 d2 <= r2).
 * semantics 1: the reference's FP64 strict rule (1-D |dx| < R, else
 *              sqdist < R*R in double) -- src/ods/optixScan.cu:170-189.
 * Writes sorted outlier slots, optional per-slot counts capped at K+1.
 * Returns the number of outliers.  Follows optixScan.cpp:817-838. */
int64_t rtod_o_brute(const float *win, int64_t W, int dim, float R, int K,
                     int semantics, int nthreads, int32_t *out_slots,
                     int32_t *out_counts);

/* 1-D: sorted-window oracle.  For sorted values the neighbour set of x_i is a
 * contiguous index interval [lo, hi] (the rounded predicate is monotone in
 * |x_i - x_j|), so count = hi - lo + 1 from two binary searches. */
int64_t rtod_o_sorted1d(const float *win, int64_t W, float R, int K,
                        int32_t *out_slots, int32_t *out_counts);

/* Grid oracle: the reference's grid-filter algorithm restated with exact
 * semantics (optixScan.cpp:947-1017 cell map + dense rule, optixScan.cu:158-199
 * count with early exit).  Cells of side R(1-1e-5)/sqrt(d) keyed by
 * floor(x / c) in double; a cell holding >= K+1 points is dense (all of them
 * are inliers); every other point is range-counted over the stencil cells
 * with early exit at K+1.  Cells are per-cell FIFOs of ring slots updated
 * incrementally each slide. */
typedef struct rtod_o_grid rtod_o_grid;
rtod_o_grid *rtod_o_grid_create(int dim, int64_t W, int64_t S, float R, int K,
                                int nthreads);
int rtod_o_grid_prime(rtod_o_grid *g, const float *pts);
int64_t rtod_o_grid_slide(rtod_o_grid *g, const float *new_pts,
                          int32_t *out_slots);
/* stats: [0] candidates last slide, [1] active cells, [2] pair tests last
 * slide, [3] dense cells last slide, [4] hits last slide: the sum over the
 * candidates of min(count, K + 1) (the reference's DEBUG_INFO hit counts,
 * optixScan.cpp:765-815, with the count stopping at K + 1) */
void rtod_o_grid_stats(const rtod_o_grid *g, int64_t *out5);
void rtod_o_grid_destroy(rtod_o_grid *g);

/* Cell key of a point as the oracle computes it (packing documented in the
 * source); exposed for tests. */
uint64_t rtod_o_cell_key(const float *p, int dim, float R);

/* Drive T slides of stream xs[(W + T*S) * dim].
 * method 0 grid, 1 brute (contract), 2 sorted1d, 3 brute (reference FP64).
 * n_out[T] per-slide outlier counts; out[] concatenated sorted slot lists
 * (capacity cap).  *us_per_slide (nullable) = mean wall time of the slide
 * calls (method 0) or of the per-window detections.  Returns total listed,
 * or -1 on overflow / bad argument. */
int64_t rtod_o_run_stream(int method, const float *xs, int dim, int64_t W,
                          int64_t S, int64_t T, float R, int K, int nthreads,
                          int32_t *n_out, int32_t *out, int64_t cap,
                          double *us_per_slide);

/* Contract (FP32 d2 <= r2) against the reference's FP64 strict rule
 * (optixScan.cu:170-189) on one window: full neighbour counts under both
 * rules (1-D: sorted-window intervals; else adjacent cells of side
 * R (1 + 1e-4)).  out3 (nullable): [0] points whose outlier status differs,
 * [1] points whose counts differ, [2] ordered pairs the rules disagree on.
 * Returns out3[0], or -1 on bad arguments. */
int64_t rtod_o_semantics_diff(const float *win, int64_t W, int dim, float R, int K,
                              int nthreads, int64_t *out3);
/* The same over slides 1..T of stream xs (every `every`-th slide, from slide
 * 1); per_slide[T] (nullable) gets each checked slide's status differences
 * (-1: not checked); out3 the totals.  Returns the total of out3[0]. */
int64_t rtod_o_semantics_diff_stream(const float *xs, int dim, int64_t W, int64_t S, int64_t T,
                                     int64_t every, float R, int K, int nthreads,
                                     int64_t *per_slide, int64_t *out3);

/* ------------------------------------------------------------------ */
/* Learned scheduler's predictor: MLP 6 -> 32 (ReLU) -> 32 (ReLU) -> 1     */
/* (sigmoid), loss = mean squared error.  Parameters, flat (RTOD_O_MLP_NP  */
/* floats): W1[6][32] b1[32] W2[32][32] b2[32] W3[32] b3.                  */
/* bf16 != 0 rounds the operands of the two hidden layers' products       */
/* (inputs, W1, h1, W2 forward; h1, d2, W2, x, d1 backward) to bfloat16    */
/* (round to nearest even) and accumulates in FP32, as the GPU's MFMA path */
/* does; the output layer (h2 . W3) and all bias/elementwise terms stay    */
/* FP32.  bf16 == 0 is plain FP32.  The GPU kernels are checked against    */
/* this; this is checked against central differences of the FP64 loss.    */
/* ------------------------------------------------------------------ */
#define RTOD_O_MLP_NP 1313
void rtod_o_mlp_forward(const float *params, const float *X /*[n][6]*/, int n, int bf16,
                        float *pred /*[n]*/);
/* gradient of the mean loss over the n samples; returns the loss */
float rtod_o_mlp_grad(const float *params, const float *X, const float *y, int n, int bf16,
                      float *grad /*[NP]*/);
double rtod_o_mlp_loss64(const double *params, const float *X, const float *y, int n);
float rtod_o_bf16_round(float x);

#ifdef __cplusplus
}
#endif
#endif
