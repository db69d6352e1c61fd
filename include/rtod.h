/*
 * rtod.h -- C-ABI of the MI355X-native RTOD per-slide outlier detector.
 *
 * Plain C types only: pointers, sizes, status codes.  No exception crosses
 * this boundary; every call returns an rtod_status.  The library owns all
 * device memory, its HIP stream and its HIP graph; the caller owns every host
 * array it passes.  One context per host thread; contexts are independent.
 *
 * Semantics (identical to the reference's ods slide loop):
 *   - window of W points, slide of S points, W % S == 0 (the reference uses
 *     unit_num = W / S and silently corrupts otherwise, optixScan.cpp:1037);
 *   - global point g lives in ring slot g mod W; rtod_prime loads slots
 *     0..W-1, slide t writes slots ((t-1) mod (W/S))*S + [0,S)
 *     (optixScan.cpp:1020-1033, :1053-1057, :1106);
 *   - a window point p is an outlier iff #{q : d2(p,q) <= R*R} <= K, the
 *     count including p (optixScan.cu:158-199, optixScan.cpp:1095-1099), with
 *     d2 in FP32: dx*dx, then fmaf(dy,dy,.), fmaf(dz,dz,.);
 *   - outliers are returned as ring slots, sorted ascending (the reference's
 *     h_outlier_list holds the same set unordered, optixScan.cpp:1097).
 */
#ifndef RTOD_H
#define RTOD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RTOD_API __attribute__((visibility("default")))

typedef enum rtod_status {
    RTOD_OK = 0,
    RTOD_ERR_INVALID_ARGUMENT = 1, /* bad config / pointer / size            */
    RTOD_ERR_UNSUPPORTED = 2,      /* valid but not supported by this build  */
    RTOD_ERR_STATE = 3,            /* call out of order (e.g. slide before prime) */
    RTOD_ERR_HIP = 4,              /* HIP runtime error (device lost, ...)   */
    RTOD_ERR_NOMEM = 5,            /* device or pinned host allocation failed */
    RTOD_ERR_INTERNAL = 6          /* device-side consistency check failed   */
} rtod_status;

typedef enum rtod_index_kind {
    RTOD_INDEX_AUTO = 0,         /* the fastest measured: the slide-run index
                                    where it applies (1-D), else the cell
                                    stencil                                  */
    RTOD_INDEX_CELL_STENCIL = 1, /* cell-sorted window + sampled key table    */
    RTOD_INDEX_LBVH = 2,         /* Morton-code LBVH over candidate boxes     */
    RTOD_INDEX_SLIDE_RUNS = 3    /* 1-D: each slide bucketed by sub-cell on
                                    arrival, the window = the m latest runs,
                                    nothing merged.  Applies with no ablation,
                                    the plain schedule (or the learned one
                                    without RTOD_FLAG_LEARN_ALWAYS: in 1-D its
                                    cost rule runs the plain kernels),
                                    S <= 16384, (W/S) * ceil(S/8192) <= 256
                                    and W <= 2^18; requested where it does
                                    not apply, rtod_create fails with
                                    RTOD_ERR_UNSUPPORTED                     */
} rtod_index_kind;

typedef enum rtod_sched_kind {
    RTOD_SCHED_PLAIN = 0,  /* candidates range-counted in cell order          */
    RTOD_SCHED_LEARNED = 1 /* MLP-predicted ordering/bucketing, exact verify  */
} rtod_sched_kind;

/* Ablations of the reference's OPTIMIZATION macro (optixScan.h:22-24): the
 * outlier sets never change, only the work done to find them. */
typedef enum rtod_ablation {
    RTOD_ABL_NONE = 0,        /* OPTIMIZATION 2 (default): grid filter, then
                                 the candidates range-counted (cell stencil)
                                 or an index over the candidates queried from
                                 every window point (LBVH, the reference's
                                 ray-BVH inversion)                          */
    RTOD_ABL_FILTER_ONLY = 1, /* OPTIMIZATION 1: grid filter, the candidates
                                 queried over an index of the whole window
                                 (LBVH; with the cell stencil this is the
                                 default computation)                        */
    RTOD_ABL_NO_FILTER = 2    /* OPTIMIZATION 0: no grid filter, every window
                                 point (or launch_ray_num rays) counted over
                                 an index of the whole window                */
} rtod_ablation;

/* flags */
#define RTOD_FLAG_PHASE_TIMES 0x1 /* record HIP events between the slide's
                                     kernels and fill rtod_times per phase  */
#define RTOD_FLAG_LEARN_ALWAYS 0x2 /* learned schedule: run the learned
                                      kernels on every slide, the cost rule
                                      (rtod_stats.plain_kernel_slides) off */

/* Replaces the compile-time DIMENSION/MK/OPTIMIZATION/UPDATE_GAS_TYPE macros
 * and the runtime flags --window --slide --R --K --launch_ray_num of the
 * reference (optixScan.h:4-28, optixScan.cpp:149-212).  dim in {1,2,3};
 * 1 <= S <= W, W % S == 0; K >= 1; R finite with 1e-18 <= R <= 1e18.  A
 * zero-initialised tail (ablation, refit, launch_ray_num) is the default. */
typedef struct rtod_config {
    int32_t dim;
    int32_t window;
    int32_t slide;
    float R;
    int32_t K;
    int32_t index_kind;   /* rtod_index_kind */
    int32_t sched_kind;   /* rtod_sched_kind */
    float miss_threshold; /* learned scheduler: fall back to the plain
                             schedule when the miss rate exceeds this      */
    int32_t flags;        /* RTOD_FLAG_*                                   */
    int32_t device;       /* HIP device ordinal                            */
    int32_t ablation;     /* rtod_ablation (0: the default path)           */
    int32_t refit;        /* UPDATE_GAS_TYPE 0 (optixScan.cpp:335-376): the
                             whole-window LBVH keeps the topology built at
                             rtod_prime and only refits its boxes each slide
                             (LBVH with RTOD_ABL_FILTER_ONLY / NO_FILTER) */
    int32_t launch_ray_num; /* --launch_ray_num (optixScan.cpp:200-206,
                             738-745), RTOD_ABL_NO_FILTER only: rays cast,
                             ray i from ring slot i mod W; 0 = W.  The list
                             then holds the ray indices whose count is <= K
                             (capacity max(W, launch_ray_num))            */
} rtod_config;

/* Per-slide time breakdown in milliseconds, the reference Timer buckets
 * (timer.h:16-24, printed by showTimeNew timer.h:102-117).  Filled on every
 * rtod_slide / rtod_slide_device call; the device parts come from the
 * start stamps (device wall clock) of the slide's kernels. */
typedef struct rtod_times {
    double copy_h2d;       /* the new slide copied into pinned staging (the
                              device reads it from there inside the ingest) */
    double prepare_cell;   /* ingest: cell keys, LDS radix sort, run merge;
                              classify: dense fast-accept, candidate units  */
    double build_index;    /* cell-stencil index (the cell-sorted window
                              merge) / LBVH build                          */
    double detect_outlier; /* range counts with early exit                */
    double copy_d2h;       /* outlier compaction + list into host memory,
                              and the copy into the caller's array         */
    double total;          /* host wall clock, entry to list in host mem  */
    double launch;         /* total minus the five buckets: kernel launch,
                              dispatch and the host's completion polling   */
} rtod_times;

/* Counters mirroring the reference's DEBUG_INFO / "BVH Node / slide" output
 * (optixScan.cpp:1176-1187, optixScan.cu:71-74). */
typedef struct rtod_stats {
    int64_t slides;
    int64_t candidates_last;    /* points not fast-accepted, last slide    */
    int64_t candidates_total;
    int64_t pair_tests_last;    /* predicate evaluations, last slide       */
    int64_t pair_tests_total;
    int64_t outliers_last;
    int64_t outliers_total;
    int64_t device_bytes;       /* device memory owned by the context      */
    int64_t host_pinned_bytes;  /* pinned host memory owned by the context */
    int64_t predictor_misses;   /* learned scheduler: mispredicted cell
                                   units (a unit: <= 4 candidates of one
                                   sparse cell in 1-D, <= 8 in 2-D/3-D,
                                   counted together)                      */
    int64_t predictor_points;   /* learned scheduler: predicted units      */
    int64_t fallbacks;          /* learned scheduler: slides not ordered by
                                   prediction -- a miss rate over the
                                   threshold two slides before (or no
                                   prediction yet), or the cost rule below */
    int64_t train_steps;        /* learned scheduler: optimiser steps      */
    double train_loss_last;     /* learned scheduler: mean loss, last slide */
    int32_t index_kind;         /* index actually used                     */
    int32_t sched_kind;         /* schedule actually used last slide       */
    int64_t hits_last;          /* neighbours found, last slide: the sum
                                   over candidates of min(count, K + 1)
                                   (DEBUG_INFO "Hit / slide",
                                   optixScan.cpp:765-815, :1181-1186; exact
                                   despite the early exits)               */
    int64_t hits_total;
    int64_t plain_kernel_slides; /* learned scheduler: slides run on the
                                   plain kernels by the cost rule, which
                                   skips inference where prediction cannot
                                   shorten the slide: 1-D (a unit's near
                                   stencil is its whole stencil), or no
                                   more cell units on the last collected
                                   slide than the detect keeps resident
                                   wavefronts (each unit has a wavefront
                                   of its own; the slowest unit sets the
                                   time).  Included in fallbacks.         */
} rtod_stats;

typedef struct rtod_ctx rtod_ctx;

/* Allocates the device ring, sorted-window buffers, batch ring, pinned
 * staging and output buffers, and the stream.  Replaces data_h2d,
 * initialize_params and initialize_cell (optixScan.cpp:232-238, :674-724,
 * :960-965). */
RTOD_API rtod_status rtod_create(const rtod_config *cfg, rtod_ctx **out);

/* Loads the first W points ([W][dim] row-major) into ring slots 0..W-1 and
 * builds the cell structure.  Untimed, like the reference's initial window
 * load (optixScan.cpp:1020-1033).  May be called again to restart. */
RTOD_API rtod_status rtod_prime(rtod_ctx *ctx, const float *pts, int64_t n_pts);

/* One slide: new_pts [S][dim] row-major; writes the window's outliers
 * (ring slots, ascending; ray indices with launch_ray_num > W) to
 * outlier_slots (capacity >= max(W, launch_ray_num)) and their number
 * to *n_outliers; times may be NULL.  Replaces the body of the while loop of
 * detect_outlier (optixScan.cpp:1045-1110). */
RTOD_API rtod_status rtod_slide(rtod_ctx *ctx, const float *new_pts, int32_t *outlier_slots,
                                int32_t *n_outliers, rtod_times *times);

/* rtod_slide with the new slide already in HBM: dev_new_pts is a device
 * pointer to [S][dim] floats, valid until the call returns (e.g. a stream
 * uploaded once).  The reference copies each slide H2D inside its timed loop
 * (optixScan.cpp:1051-1058); this entry point is that loop with the copy
 * hoisted out by the caller.  Rejects host pointers. */
RTOD_API rtod_status rtod_slide_device(rtod_ctx *ctx, const float *dev_new_pts,
                                       int32_t *outlier_slots, int32_t *n_outliers,
                                       rtod_times *times);

/* n_slides consecutive slides from HBM: slide i takes the S points at
 * dev_pts + i*S*dim.  No counterpart in the reference, whose loop
 * (optixScan.cpp:1045-1110) runs one slide per iteration; this runs the same
 * slides software-pipelined on one stream.  With the cell-stencil index and
 * S <= 8192 a slide's detect shares a launch with the previous slide's
 * outlier list, the next slide's window merge, the run merge of the slide
 * two ahead and the tile sorts of the slide three ahead; its classify is a
 * launch of its own before it (learned schedule) or the first workgroups of
 * that launch, which the detect waits for (plain schedule; the wait is
 * bounded, 2^20 polls or the environment variable RTOD_CLASSIFY_WAIT_SPINS
 * read at rtod_create, and past it the call fails with error flag 8).  (Otherwise a
 * slide's launches carry the next slide's sort and the previous slide's list
 * where they can.)  Up to four slides are queued,
 * and the host collects each slide's list as it completes.  On the slide-run
 * index (1-D) the slides go in batches instead, the first 32 slides and then
 * 128 each (64 from a rtod_host_alloc buffer, 32 when rtod_slides stages them
 * from pageable memory): a
 * launch detects one batch and sorts the next, with at most two launches
 * queued.  After a failed
 * call the context must be primed again.  Results equal those of n_slides
 * rtod_slide_device calls:
 * slide i's outliers (ring slots, ascending) are
 * out_slots[out_offsets[i] .. out_offsets[i+1]); out_offsets holds
 * n_slides + 1 entries and out_slots at least n_slides * max(W,
 * launch_ray_num) (slots_cap). */
RTOD_API rtod_status rtod_slides_device(rtod_ctx *ctx, const float *dev_pts, int32_t n_slides,
                                        int32_t *out_offsets, int32_t *out_slots,
                                        int64_t slots_cap);

/* rtod_slides_device with the slides in host memory: pts [n_slides*S][dim]
 * row-major, slide i at pts + i*S*dim.  Each slide is copied into one of the
 * context's pinned staging slots just before the launch that reads it (its
 * ingest, up to three slides ahead of the slide being detected), and the
 * kernels read it from there over PCIe: the reference's per-slide H2D copy
 * (optixScan.cpp:1051-1058) inside the pipelined run, overlapping the
 * detection of earlier slides.  Outputs, errors and pipelining as
 * rtod_slides_device; the results equal those of n_slides rtod_slide calls.
 * No counterpart in the reference, whose loop (optixScan.cpp:1045-1110)
 * runs one slide per iteration. */
RTOD_API rtod_status rtod_slides(rtod_ctx *ctx, const float *pts, int32_t n_slides,
                                 int32_t *out_offsets, int32_t *out_slots, int64_t slots_cap);

/* Pinned (page-locked) host memory mapped for every device: bytes > 0, the
 * buffer at *out.  A slide that rtod_slide finds inside such a buffer, or a
 * run of slides that rtod_slides finds wholly inside one, is read by the
 * kernels in place over PCIe, with no copy into the context's staging; the
 * caller leaves it unchanged until the call returns.  rtod_host_free
 * releases a buffer from rtod_host_alloc (NULL: nothing to do; any other
 * pointer: RTOD_ERR_INVALID_ARGUMENT).  No counterpart in the reference,
 * whose loop copies each slide from pageable memory
 * (optixScan.cpp:1051-1058); for callers that can receive their points
 * straight into such a buffer (a socket or file read, a ring of slides). */
RTOD_API rtod_status rtod_host_alloc(int64_t bytes, void **out);
RTOD_API rtod_status rtod_host_free(void *p);

/* Phase-level entry points: the same slide split where the reference's loop
 * body calls out.  Call in this order; rtod_slide == all five.
 *   rtod_copy_new_points  <- cudaMemcpy of the new slide (optixScan.cpp:1051-1058)
 *   rtod_prepare_cell     <- prepare_c_non_points_queue   (optixScan.cpp:967-1017)
 *   rtod_build_index      <- rebuild_gas / kGenAABB       (optixScan.cpp:378-426,
 *                                                          aabb.cu:35-45)
 *   rtod_detect_outlier   <- launch                       (optixScan.cpp:727-750)
 *   rtod_copy_outlier     <- D2H + select count <= K      (optixScan.cpp:1083-1101) */
RTOD_API rtod_status rtod_copy_new_points(rtod_ctx *ctx, const float *new_pts);
RTOD_API rtod_status rtod_prepare_cell(rtod_ctx *ctx);
RTOD_API rtod_status rtod_build_index(rtod_ctx *ctx);
RTOD_API rtod_status rtod_detect_outlier(rtod_ctx *ctx);
RTOD_API rtod_status rtod_copy_outlier(rtod_ctx *ctx, int32_t *outlier_slots, int32_t *n_outliers);

/* Counters (reference DEBUG_INFO counters and [Mem] lines,
 * optixScan.cpp:66-93, :752-815, :1170-1187). */
RTOD_API rtod_status rtod_get_stats(const rtod_ctx *ctx, rtod_stats *out);

/* Average device duration (ms) per launch of each kernel recorded with HIP
 * events on the context's stream while RTOD_FLAG_PHASE_TIMES is set (with
 * rtod_slides_device the flag also waits on every slide: a measuring mode;
 * its launch per slide is "slide_stage" (plain schedule) and its two launches
 * per slide "classify_stage" and "detect_stage" (learned schedule), otherwise
 * its window merge and detect launches that carry the next slide's sort are
 * "merge_window_ingest" and "detect_runs").
 * names[i] (static strings) / avg_ms[i] / launches[i] for i < *n (in: capacity). */
RTOD_API rtod_status rtod_kernel_times(const rtod_ctx *ctx, const char **names, double *avg_ms,
                                       int64_t *launches, int32_t *n);
RTOD_API rtod_status rtod_reset_kernel_times(rtod_ctx *ctx);

/* Kernel durations (ms, device wall clock) of the last slide run one per
 * call (rtod_slide / rtod_slide_device), from the start stamp each of its
 * kernels' first workgroup records: out[0] ingest (tile sorts and batch
 * merges), [1] window merge, [2] classify, [3] LBVH build (0 with the cell
 * stencil), [4] detect, [5] epilogue (list into host memory); n (in) >= 6.
 * The stamps rtod_times is built from; RTOD_ERR_STATE before a slide.  No
 * reference counterpart beyond Timer's buckets (timer.h:16-24). */
RTOD_API rtod_status rtod_stage_times(const rtod_ctx *ctx, double *out, int32_t n);

/* ---- Learned scheduler (RTOD_SCHED_LEARNED) ------------------------------
 * No reference counterpart: the reference schedules every range count the
 * same way.  The predictor is an MLP 6 -> 32 -> 32 -> 1 whose RTOD_MLP_NPARAMS
 * FP32 parameters are laid out W1[6][32] b1[32] W2[32][32] b2[32] W3[32] b3
 * (row index = input unit).  It only orders the exact range counts; outlier
 * sets are identical to RTOD_SCHED_PLAIN's on every slide. */
#define RTOD_MLP_NPARAMS 1313

/* Copies the context's current predictor parameters (n == RTOD_MLP_NPARAMS). */
RTOD_API rtod_status rtod_get_predictor(rtod_ctx *ctx, float *params, int32_t n);

/* Replaces them, clears the optimiser state; frozen != 0 stops online
 * training (for A/B runs against a fixed or untrained predictor). */
RTOD_API rtod_status rtod_set_predictor(rtod_ctx *ctx, const float *params, int32_t n,
                                        int32_t frozen);

/* Predictor forward pass on the GPU's matrix cores, the device code the
 * detector runs: X [n][6] features -> out [n] in (0,1).  Host arrays. */
RTOD_API rtod_status rtod_mlp_forward(int32_t device, const float *params, const float *X,
                                      int32_t n, float *out);

/* Gradient of the mean squared error over n (1..128) samples, by the
 * detector's training kernels: grad [RTOD_MLP_NPARAMS], *loss = mean error. */
RTOD_API rtod_status rtod_mlp_grad(int32_t device, const float *params, const float *X,
                                   const float *y, int32_t n, float *grad, float *loss);

/* rtod_mlp_grad that allocates nothing: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code that forbids hipMalloc,
 * tight serving or training loops): the parameters, samples, gradient and
 * loss are staged in a fixed device scratch area of the library (14592
 * bytes per GPU, held until the process ends) instead of per-call
 * allocations, then the same kernel runs.  n in [1, 128]; larger n is
 * rejected (RTOD_ERR_INVALID_ARGUMENT).  Every call on one GPU shares its
 * area: calls that overlap on one GPU (from several threads) race. */
RTOD_API rtod_status rtod_mlp_grad_static(int32_t device, const float *params, const float *X,
                                          const float *y, int32_t n, float *grad,
                                          float *loss);

/* ---- Batched streams ----------------------------------------------------
 * No reference counterpart: the reference runs one stream per process.  A
 * batch advances n independent streams (each its own dim, W, S, R, K; plain
 * schedule, cell-stencil index) by one slide per step: each phase of the
 * slide is ONE launch covering every stream, and the step's launches are
 * replayed from a hipGraph.  Per stream the semantics are exactly
 * rtod_slide's. */
typedef struct rtod_batch rtod_batch;

typedef struct rtod_batch_tuning {
    int32_t detect_points_per_wg; /* window points per detect workgroup     */
    int32_t learned_dims;         /* learned batches: bit d-1 set, the d-D
                                     streams run the learned kernels (2-D
                                     and 3-D only; autotune chooses)       */
    float cell_scale[3];          /* cell side multiplier by dim (1-D in
                                     [1/16,1], 2-D in [1/4,1], 3-D 1)       */
    double us_per_step;           /* autotune: measured mean step time      */
} rtod_batch_tuning;

/* cfgs[n]: per-stream configurations (sched PLAIN, index AUTO or
 * CELL_STENCIL; device and flags ignored, `device` used). */
RTOD_API rtod_status rtod_batch_create(const rtod_config *cfgs, int32_t n, int32_t device,
                                       rtod_batch **out);
/* pts[i]: the first W_i points of stream i ([W_i][dim_i], host). */
RTOD_API rtod_status rtod_batch_prime(rtod_batch *b, const float *const *pts);
/* One slide of every stream: new_pts[i] [S_i][dim_i] (host).  Outliers of
 * stream i (ring slots, ascending) are out_slots[out_offsets[i] ..
 * out_offsets[i+1]); out_offsets has n+1 entries, out_slots capacity sum W_i. */
RTOD_API rtod_status rtod_batch_step(rtod_batch *b, const float *const *new_pts,
                                     int32_t *out_offsets, int32_t *out_slots);
/* Device-resident streams: stream i's points [n_points[i]][dim_i] stay in HBM
 * (device pointers); slide t reads rows W_i + (t-1) S_i ...  NULL unbinds. */
RTOD_API rtod_status rtod_batch_bind_device(rtod_batch *b, const float *const *dev_streams,
                                            const int64_t *n_points);
RTOD_API rtod_status rtod_batch_prime_device(rtod_batch *b);
RTOD_API rtod_status rtod_batch_step_device(rtod_batch *b, int32_t *out_offsets,
                                            int32_t *out_slots);
/* The autotuner: replays the first `slides` slides of the bound device
 * streams under each candidate (detect workgroup size x cell scale of the
 * 1-D/2-D streams), keeps the fastest, and leaves the batch primed from
 * slide 0 with it.  Results do not depend on the tuning. */
RTOD_API rtod_status rtod_batch_autotune(rtod_batch *b, int32_t slides, rtod_batch_tuning *out);
RTOD_API rtod_status rtod_batch_get_tuning(const rtod_batch *b, rtod_batch_tuning *out);
/* Sets a tuning; the batch must be primed again afterwards. */
RTOD_API rtod_status rtod_batch_set_tuning(rtod_batch *b, const rtod_batch_tuning *t);
RTOD_API rtod_status rtod_batch_stream_stats(const rtod_batch *b, int32_t i, rtod_stats *out);
RTOD_API void rtod_batch_destroy(rtod_batch *b);

/* Frees everything (reference cleanup, optixScan.cpp:883-932). */
RTOD_API void rtod_destroy(rtod_ctx *ctx);

RTOD_API const char *rtod_status_str(rtod_status s);

/* Last error message of this thread (static storage, never NULL). */
RTOD_API const char *rtod_last_error(void);

/* Library version string. */
RTOD_API const char *rtod_version(void);

#ifdef __cplusplus
}
#endif
#endif
