// rtod_lib.hip -- context management and the C-ABI of include/rtod.h.
//
// One context = one window stream on one GPU: its own HIP stream, device
// arena and pinned staging.  A slide is four direct kernel launches of the
// plain path (rtod_plain.h) and one stream synchronisation.  Per-slide
// quantities (ring offset, ping-pong parity, batch-ring entries, staging
// buffer) travel as kernel arguments.  Direct launches are used rather than
// a hipGraph: on ROCm 7.2 / MI355X a graph of four small kernels measured
// 16 us per launch against 6 us for the same four direct launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <emmintrin.h>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/rtod.h"
#include "rtod_batch.h"
#include "rtod_lbvh.h"
#include "rtod_runs.h"

using namespace rtod;

namespace {

// the last error message of this thread, in a fixed buffer (setting it
// allocates nothing)
thread_local char g_last_error[1024] = "";

void set_error(const char *s)
{
    std::snprintf(g_last_error, sizeof(g_last_error), "%s", s);
}
void set_error(const std::string &s) { set_error(s.c_str()); }
void set_error_hip(const char *what, hipError_t e)
{
    std::snprintf(g_last_error, sizeof(g_last_error), "%s: %s", what, hipGetErrorString(e));
}

#define RTOD_HIP(call)                                                                            \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            set_error(std::string(#call) + ": " + hipGetErrorString(e_));                         \
            return RTOD_ERR_HIP;                                                                  \
        }                                                                                         \
    } while (0)

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// Pinned host buffers from rtod_host_alloc: [base, base + bytes) and the
// device's address of base (mapped memory)
struct PinnedBuf {
    const char *base;
    size_t bytes;
    const char *dev;
};
std::mutex g_pinned_mu;
std::vector<PinnedBuf> g_pinned;
std::atomic<int> g_pinned_n{0};

// The device's address of host range [p, p + bytes) when it lies inside one
// buffer from rtod_host_alloc, else nullptr (no lock while there is none).
const float *pinned_device_ptr(const void *p, size_t bytes)
{
    if (g_pinned_n.load(std::memory_order_acquire) == 0)
        return nullptr;
    const char *q = static_cast<const char *>(p);
    std::lock_guard<std::mutex> lk(g_pinned_mu);
    for (const PinnedBuf &b : g_pinned)
        if (q >= b.base && bytes <= b.bytes && (size_t)(q - b.base) <= b.bytes - bytes)
            return reinterpret_cast<const float *>(b.dev + (q - b.base));
    return nullptr;
}

// Copies n floats into pinned staging with streaming (non-temporal) stores:
// the lines go to memory without first being read for ownership, and the
// device reads them over PCIe from there.  The fence orders the stores before
// the launch that reads the slot.  On MI355X's host this staged a 20 KB slide
// about 0.1 us faster per slide in the pipelined loop than std::memcpy; one
// slide per call, read by the launch right after the copy, it measured 0.5 us
// slower, so the one-slide calls keep std::memcpy.
void stream_copy(float *dst, const float *src, size_t n)
{
    size_t k = 0;
    for (; k < n && (reinterpret_cast<uintptr_t>(dst + k) & 15) != 0; ++k)
        dst[k] = src[k];
    for (; k + 16 <= n; k += 16) {
        const __m128i a = _mm_loadu_si128(reinterpret_cast<const __m128i *>(src + k));
        const __m128i b = _mm_loadu_si128(reinterpret_cast<const __m128i *>(src + k + 4));
        const __m128i e = _mm_loadu_si128(reinterpret_cast<const __m128i *>(src + k + 8));
        const __m128i f = _mm_loadu_si128(reinterpret_cast<const __m128i *>(src + k + 12));
        _mm_stream_si128(reinterpret_cast<__m128i *>(dst + k), a);
        _mm_stream_si128(reinterpret_cast<__m128i *>(dst + k + 4), b);
        _mm_stream_si128(reinterpret_cast<__m128i *>(dst + k + 8), e);
        _mm_stream_si128(reinterpret_cast<__m128i *>(dst + k + 12), f);
    }
    for (; k < n; ++k)
        dst[k] = src[k];
    _mm_sfence();
}

enum KernelId {
    kKIngest = 0,
    kKMergePass,
    kKMergeWindow,
    kKClassify,
    kKDetect,
    kKLbvhBuild,
    kKLbvhQuery,
    kKMergeWindowIngest, // pipelined: window merge + the next slide's tile sorts
    kKDetectRuns,        // pipelined: detect + the next slide's run merge
    kKClassifyStage,     // two launches per slide (learned): classify
    kKDetectStage,       // two launches per slide (learned): detect + epilogue +
                         // window merge + run merge + tile sorts
    kKSlideStage,        // one launch per slide (plain): classify + the above
    kKRunSort,           // slide-run index: a slide's runs sorted (two launches per slide)
    kKRunDetect,         // slide-run index: a slide detected (two launches per slide),
                         // or a pipelined run's last launch (its last slides detected)
    kKRunSlide,          // slide-run index: one slide sorted and detected, one launch
    kKRunStage,          // slide-run index, pipelined: the detects of kRunBatch slides
                         // and the sorts of the next kRunBatch
    kKCount
};
const char *kKernelNames[kKCount] = {"ingest_sort",         "merge_pass",  "merge_window",
                                     "classify",            "detect",      "lbvh_build",
                                     "lbvh_query",          "merge_window_ingest",
                                     "detect_runs",         "classify_stage",
                                     "detect_stage",        "slide_stage",
                                     "run_sort",            "run_detect",
                                     "run_slide",           "run_stage"};

// LDS-sort tile shapes (threads x items per thread).  A batch of up to
// kRunMax * 1024 points is sorted as tiles of 1024 points on as many
// workgroups (one CU sorts a tile; a whole 8k-point tile on one CU costs
// several microseconds per 4-bit pass), then merged by one k_merge_runs
// launch; up to kRunMergeCap points, as tiles of 2048.  Larger batches use
// the largest tile that fits 160 KB of LDS (coordinates/counters pool, keys,
// 16-bit values) and pairwise merge passes.
struct SortShape {
    int nt, it;
};
constexpr SortShape kSortShapes[6] = {{256, 2},  {256, 4},  {512, 4},
                                      {1024, 4}, {1024, 8}, {256, 8}};
constexpr int kBatchSortShape = 3; // batched streams: the 1024 x 4 tile
static_assert(kSortShapes[kBatchSortShape].nt == kBatchSortNT &&
                  kSortShapes[kBatchSortShape].it == kBatchSortIT,
              "batched ingest tile");
constexpr int kRunTileShape = 1;  // 1024-point tiles merged by k_merge_runs
constexpr int kRunTileShape2 = 5; // 2048-point tiles merged by k_merge_runs
static_assert(kSortShapes[kRunTileShape].nt * kSortShapes[kRunTileShape].it == kRunLen &&
                  kSortShapes[kRunTileShape2].nt * kSortShapes[kRunTileShape2].it ==
                      2 * kRunLen,
              "runs merged by k_merge_runs");

int pick_sort_shape(int S, int dim)
{
    const int one = kSortShapes[0].nt * kSortShapes[0].it;
    if (S <= one)
        return 0;
    if (S <= kRunMax * kRunLen)
        return kRunTileShape;
    if (S <= kRunMergeCap)
        return kRunTileShape2;
    return dim == 3 ? 3 : 4; // (1024, 8) x 3 floats exceeds LDS
}

// k_merge_runs covers a batch of several tiles (of 1024 or 2048 points, at
// most kRunMax) in one launch
bool use_run_merge(int S, int CH)
{
    return (CH == kRunLen || CH == 2 * kRunLen) && S > CH && S <= kRunMax * CH;
}

// the next slide's ingest stage can ride in this slide's core launches:
// 256-thread sort tiles (the merge's workgroup size), at most one merge
// launch (k_merge_runs, the detect's workgroup size)
bool ride_along(int sort_shape, int S, int CH)
{
    return kSortShapes[sort_shape].nt == 256 && (S <= CH || use_run_merge(S, CH));
}

} // namespace

struct rtod_ctx {
    rtod_config cfg{};
    int dim = 1, W = 0, S = 0, m = 0, K = 0;
    int sort_shape = 0, CH = 512;
    // detect workgroups: twice the three per CU that are resident, so the
    // dispatcher hands a finished workgroup's place to a waiting one (the
    // large 3-D window: 127.9 -> 123.2 us per slide; four times: 122.4)
    int detect_blocks = 1536;
    int units_last = -1;     // cell units of the last collected slide (-1: none yet)
    // learned schedule, the cost rule: a slide runs the learned kernels only
    // with dim > 1 and more units on the last collected slide than
    // learn_min_units (the detect's resident wavefronts); the choice of
    // slide t is made once, before its classify is launched, and kept in
    // learn_of[t % kLearnRing] for its later launches (the epilogue)
    static constexpr int kLearnRing = 8;
    int learn_min_units = 0;
    int8_t learn_of[kLearnRing] = {};
    int64_t learn_t[kLearnRing] = {-1, -1, -1, -1, -1, -1, -1, -1};
    bool learn_new = false;  // the last collected slide ran the learned kernels
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = true; // false: a batch member on the batch's stream
    hipStream_t train_stream = nullptr; // learned: the predictor's training launches
    hipEvent_t train_ev = nullptr;
    bool train_pending = false;
    int64_t train_t_next = 0; // the pending training launch skips the replay
                              // slots of kReplaySkip slides from this one
    Dev d{};
    float *d_prime = nullptr; // device copy of the initial window
    float *h_in = nullptr;   // stage_slots pinned staging slots of S*dim floats
    int stage_slots = 2;     // slide t is staged in slot t % stage_slots
    HostOut *h_out = nullptr;
    int n_hout = kHostOuts;  // host-out buffers: slide t writes t % n_hout
    // 1-D slide-run index (rtod_runs.h): the window as the m latest sorted
    // runs; every other path of the context is unused
    bool runs = false;
    RunsDev rd{};
    const float *dev_src = nullptr; // device-resident new slide of this call
    bool lbvh = false;
    bool learned = false;
    bool refit = false;      // whole-window LBVH: boxes refit each slide
    int ray_num = 0;         // launch_ray_num (no grid filter; 0: W)
    int list_cap = 0;        // entries a slide's list may hold: max(W, ray_num)
    Lbvh L{};
    int lbvh_passes = 0;
    int lbvh_group = 64;     // LBVH walks: queries per wavefront (divides 64)
    std::vector<void *> dev_allocs;
    int64_t dev_bytes = 0, pinned_bytes = 0;
    bool primed = false;
    int64_t fail_at = -1;  // RTOD_FAIL_AT_SLIDE: collecting this slide reports a
                           // failure (a host-side test of the error paths)
    bool staged = false;   // phase API: new points copied, kernels pending
    int phase = 0;         // phase API progress
    int64_t host_t = 0;
    hipEvent_t ev[8] = {};
    bool ride = false; // pipelined slides: next ingest rides in core launches
    double wall_khz = 0.0; // the device wall clock's rate (ticks per ms)
    double ksum[kKCount] = {};
    int64_t kcnt[kKCount] = {};
    // stats
    int64_t cand_last = 0, cand_total = 0, tests_last = 0, tests_total = 0;
    int64_t hits_last = 0, hits_total = 0;
    int64_t out_last = 0, out_total = 0;
    int64_t misses_total = 0, pred_points = 0, fallbacks = 0, train_steps = 0;
    int64_t plain_slides = 0; // learned: slides on the plain kernels (cost rule)
    double train_loss_last = 0.0;
    bool used_pred_last = false;
};

namespace {

template <typename T> rtod_status dalloc(rtod_ctx *c, T **p, size_t n)
{
    void *q = nullptr;
    size_t bytes = n * sizeof(T);
    if (bytes == 0)
        bytes = 16;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        return RTOD_ERR_NOMEM;
    }
    c->dev_allocs.push_back(q);
    c->dev_bytes += (int64_t)bytes;
    *p = (T *)q;
    return RTOD_OK;
}

#define RTOD_TRY(x)                                                                               \
    do {                                                                                          \
        rtod_status s_ = (x);                                                                     \
        if (s_ != RTOD_OK)                                                                        \
            return s_;                                                                            \
    } while (0)

template <int DIM> void launch_ingest(rtod_ctx *c, hipStream_t s, const Step &st, int mode,
                                      const float *src, int grid, int dst_batch)
{
    const dim3 g(grid);
    switch (c->sort_shape) {
    case 0:
        hipLaunchKernelGGL((k_ingest_sort<DIM, 256, 2>), g, dim3(256), 0, s, c->d, st, mode, src,
                           dst_batch);
        break;
    case 1:
        hipLaunchKernelGGL((k_ingest_sort<DIM, 256, 4>), g, dim3(256), 0, s, c->d, st, mode, src,
                           dst_batch);
        break;
    case 2:
        hipLaunchKernelGGL((k_ingest_sort<DIM, 512, 4>), g, dim3(512), 0, s, c->d, st, mode, src,
                           dst_batch);
        break;
    case 3:
        hipLaunchKernelGGL((k_ingest_sort<DIM, 1024, 4>), g, dim3(1024), 0, s, c->d, st, mode,
                           src, dst_batch);
        break;
    case 5:
        hipLaunchKernelGGL((k_ingest_sort<DIM, 256, 8>), g, dim3(256), 0, s, c->d, st, mode, src,
                           dst_batch);
        break;
    default:
        if constexpr (DIM < 3)
            hipLaunchKernelGGL((k_ingest_sort<DIM, 1024, 8>), g, dim3(1024), 0, s, c->d, st, mode,
                               src, dst_batch);
        break;
    }
}

void launch_ingest_dim(rtod_ctx *c, hipStream_t s, const Step &st, int mode, const float *src,
                       int grid, int dst_batch)
{
    if (c->dim == 1)
        launch_ingest<1>(c, s, st, mode, src, grid, dst_batch);
    else if (c->dim == 2)
        launch_ingest<2>(c, s, st, mode, src, grid, dst_batch);
    else
        launch_ingest<3>(c, s, st, mode, src, grid, dst_batch);
}

void launch_merge_pass(rtod_ctx *c, hipStream_t s, const Step &st, const uint64_t *sk,
                       const int32_t *ss, const float *sx, size_t sxs, const int32_t *sp,
                       uint64_t *dk, int32_t *ds, float *dx, size_t dxs, int32_t *dp, int n_total,
                       int seg_len, int L, int dst_batch, int mode)
{
    const dim3 g((n_total + 255) / 256), b(256);
    if (c->dim == 1)
        hipLaunchKernelGGL(k_merge_pass<1>, g, b, 0, s, c->d, st, sk, ss, sx, sxs, sp, dk, ds, dx,
                           dxs, dp, n_total, seg_len, L, dst_batch, mode);
    else if (c->dim == 2)
        hipLaunchKernelGGL(k_merge_pass<2>, g, b, 0, s, c->d, st, sk, ss, sx, sxs, sp, dk, ds, dx,
                           dxs, dp, n_total, seg_len, L, dst_batch, mode);
    else
        hipLaunchKernelGGL(k_merge_pass<3>, g, b, 0, s, c->d, st, sk, ss, sx, sxs, sp, dk, ds, dx,
                           dxs, dp, n_total, seg_len, L, dst_batch, mode);
}

// chunk-sorted runs of length c->CH into sorted segments of S (the
// batches) in the batch ring: one k_merge_runs launch, or pairwise merge
// passes for batches beyond its LDS staging
void launch_batch_merges(rtod_ctx *c, hipStream_t s, const Step &st, int mode, int n_total)
{
    if (use_run_merge(c->S, c->CH)) {
        const dim3 g((n_total / c->S) * run_merge_wgs(c->S)), b(kRunMergeNT);
        if (c->dim == 1)
            hipLaunchKernelGGL(k_merge_runs<1>, g, b, 0, s, c->d, st, mode);
        else if (c->dim == 2)
            hipLaunchKernelGGL(k_merge_runs<2>, g, b, 0, s, c->d, st, mode);
        else
            hipLaunchKernelGGL(k_merge_runs<3>, g, b, 0, s, c->d, st, mode);
        return;
    }
    int src = 0;
    const size_t bxs = (size_t)c->d.nent * c->S;
    for (int L = c->CH; L < c->S; L *= 2) {
        const bool last = 2 * L >= c->S;
        launch_merge_pass(c, s, st, c->d.tkey[src], c->d.tslot[src], c->d.txyz[src],
                          (size_t)c->d.tmpn, nullptr, last ? c->d.bkey : c->d.tkey[src ^ 1],
                          last ? c->d.bslot : c->d.tslot[src ^ 1],
                          last ? c->d.bxyz : c->d.txyz[src ^ 1], last ? bxs : (size_t)c->d.tmpn,
                          nullptr, n_total, c->S, L, last ? 1 : 0, mode);
        src ^= 1;
    }
}

// slide st's window merge; next != nullptr: the next slide's tile sorts
// ride along as extra workgroups (ride_along); prev != nullptr: so does the
// epilogue of the previous slide, whose detect left it (deferred finish)
template <int DIM>
void launch_merge_window_dim(rtod_ctx *c, hipStream_t s, const Step &st, const Step *next,
                             const Step *prev)
{
    const int nmb = (c->W + kMergeT - 1) / kMergeT;
    if (!next && !prev) {
        hipLaunchKernelGGL(k_merge_window<DIM>, dim3(nmb), dim3(kMergeT), 0, s, c->d, st);
        return;
    }
    const int fin = prev ? c->d.fin_parts : 0;
    const dim3 g(fin + nmb + (next ? c->d.cpb : 0));
    const Step nx = next ? *next : Step{}, pv = prev ? *prev : Step{};
    if (c->sort_shape == 0)
        hipLaunchKernelGGL((k_merge_window_ingest<DIM, 2>), g, dim3(256), 0, s, c->d, st, nx, pv,
                           nmb, fin);
    else if (c->sort_shape == kRunTileShape2)
        hipLaunchKernelGGL((k_merge_window_ingest<DIM, 8>), g, dim3(256), 0, s, c->d, st, nx, pv,
                           nmb, fin);
    else
        hipLaunchKernelGGL((k_merge_window_ingest<DIM, 4>), g, dim3(256), 0, s, c->d, st, nx, pv,
                           nmb, fin);
}

void launch_merge_window(rtod_ctx *c, hipStream_t s, const Step &st, const Step *next = nullptr,
                         const Step *prev = nullptr)
{
    if (c->dim == 1)
        launch_merge_window_dim<1>(c, s, st, next, prev);
    else if (c->dim == 2)
        launch_merge_window_dim<2>(c, s, st, next, prev);
    else
        launch_merge_window_dim<3>(c, s, st, next, prev);
}

template <int DIM> void launch_classify_dim(rtod_ctx *c, hipStream_t s, const Step &st)
{
    const int grid = (c->W + 255) / 256;
    if (c->lbvh) // the LBVH index is built over the per-point candidate list
        hipLaunchKernelGGL((k_classify<DIM, kClsPoints>), dim3(grid), dim3(kClassifyNT), 0, s,
                           c->d, st);
    else if (st.learn)
        hipLaunchKernelGGL((k_classify<DIM, kClsLearned>), dim3(grid), dim3(kClassifyNT), 0, s,
                           c->d, st);
    else
        hipLaunchKernelGGL((k_classify<DIM, kClsUnits>), dim3(grid), dim3(kClassifyNT), 0, s,
                           c->d, st);
}

void launch_classify(rtod_ctx *c, hipStream_t s, const Step &st)
{
    if (c->dim == 1)
        launch_classify_dim<1>(c, s, st);
    else if (c->dim == 2)
        launch_classify_dim<2>(c, s, st);
    else
        launch_classify_dim<3>(c, s, st);
}

// slide st's detect; runs != nullptr: that slide's run merge rides along as
// extra workgroups (and the epilogue is left to the next launch); fin: the
// detect runs the slide's epilogue itself
template <int DIM>
void launch_detect_dim(rtod_ctx *c, hipStream_t s, const Step &st, const Step *runs, bool fin)
{
    const int grid = c->detect_blocks;
    if (runs) {
        const dim3 g(grid + run_merge_wgs(c->S));
        const size_t lds = detect_runs_lds(c->S);
        if (st.learn)
            hipLaunchKernelGGL((k_detect_runs<DIM, true>), g, dim3(kDetectNT), lds, s, c->d, st,
                               *runs, grid);
        else
            hipLaunchKernelGGL((k_detect_runs<DIM, false>), g, dim3(kDetectNT), lds, s, c->d, st,
                               *runs, grid);
        return;
    }
    const int f = fin ? 1 : 0;
    if (st.learn)
        hipLaunchKernelGGL((k_detect<DIM, true>), dim3(grid), dim3(kDetectNT), 0, s, c->d, st, f);
    else
        hipLaunchKernelGGL((k_detect<DIM, false>), dim3(grid), dim3(kDetectNT), 0, s, c->d, st, f);
}

void launch_detect(rtod_ctx *c, hipStream_t s, const Step &st, const Step *runs = nullptr,
                   bool fin = true)
{
    if (c->dim == 1)
        launch_detect_dim<1>(c, s, st, runs, fin);
    else if (c->dim == 2)
        launch_detect_dim<2>(c, s, st, runs, fin);
    else
        launch_detect_dim<3>(c, s, st, runs, fin);
}

// slide st's epilogue on its own, after a detect that left it
// (stamps: a slide run one per call, whose epilogue publishes its kernels'
// start stamps)
void launch_finish(rtod_ctx *c, hipStream_t s, const Step &st, bool stamps = false)
{
    hipLaunchKernelGGL(k_finish, dim3(c->d.fin_parts), dim3(256), 0, s, c->d, st,
                       stamps ? 1 : 0);
}

// Stage launches (cell-stencil index, the ingest riding), learned
// schedule: slide st's classify ...
template <int DIM, int MODE>
void launch_classify_stage_dim(rtod_ctx *c, hipStream_t s, const Step &st)
{
    const int ncls = (c->W + kClassifyNT - 1) / kClassifyNT;
    hipLaunchKernelGGL((k_classify_stage<DIM, MODE>), dim3(ncls), dim3(kClassifyNT), 0, s, c->d,
                       st);
}

void launch_classify_stage(rtod_ctx *c, hipStream_t s, const Step &st)
{
    if (c->dim == 1)
        st.learn ? launch_classify_stage_dim<1, kClsLearned>(c, s, st)
                 : launch_classify_stage_dim<1, kClsUnits>(c, s, st);
    else if (c->dim == 2)
        st.learn ? launch_classify_stage_dim<2, kClsLearned>(c, s, st)
                 : launch_classify_stage_dim<2, kClsUnits>(c, s, st);
    else
        st.learn ? launch_classify_stage_dim<3, kClsLearned>(c, s, st)
                 : launch_classify_stage_dim<3, kClsUnits>(c, s, st);
}

// ... and slide st's detect with the epilogue of prev, the window merge of
// merge, the run merge of runs and the tile sorts of ts (tiles of c->CH
// points: kDetectNT threads with CH / kDetectNT items each)
struct StageParts {
    const Step *prev, *merge, *runs, *ts;
};

template <int DIM, bool LEARN, int TSIT>
void launch_detect_stage_t(rtod_ctx *c, hipStream_t s, const Step &st, const StageParts &p)
{
    const int nfin = p.prev ? c->d.fin_parts : 0;
    const int nmb = p.merge ? (c->W + kMergeT - 1) / kMergeT : 0;
    const int nrm = p.runs ? run_merge_wgs(c->S) : 0;
    const int nts = p.ts ? c->d.cpb : 0;
    const int ndet = c->detect_blocks;
    const dim3 g(nfin + nmb + nrm + nts + ndet);
    const size_t lds = detect_stage_lds<DIM, TSIT>(nfin > 0, nmb > 0, nrm > 0, nts > 0, c->S);
    const Step pv = p.prev ? *p.prev : Step{}, mv = p.merge ? *p.merge : Step{},
               rv = p.runs ? *p.runs : Step{}, tv = p.ts ? *p.ts : Step{};
    hipLaunchKernelGGL((k_detect_stage<DIM, LEARN, TSIT>), g, dim3(kDetectNT), lds, s, c->d, st,
                       pv, mv, rv, tv, nfin, nmb, nrm, nts, ndet);
}

template <int DIM, bool LEARN>
void launch_detect_stage_l(rtod_ctx *c, hipStream_t s, const Step &st, const StageParts &p)
{
    // tiles of CH = kDetectNT * TSIT points (ride_along: CH is 512, 1024 or 2048)
    if (c->CH == kDetectNT)
        launch_detect_stage_t<DIM, LEARN, 1>(c, s, st, p);
    else if (c->CH == 2 * kDetectNT)
        launch_detect_stage_t<DIM, LEARN, 2>(c, s, st, p);
    else
        launch_detect_stage_t<DIM, LEARN, 4>(c, s, st, p);
}

template <int DIM>
void launch_detect_stage_dim(rtod_ctx *c, hipStream_t s, const Step &st, const StageParts &p)
{
    if (st.learn)
        launch_detect_stage_l<DIM, true>(c, s, st, p);
    else
        launch_detect_stage_l<DIM, false>(c, s, st, p);
}

// one launch per slide (plain schedule): slide st's classify in front of
// the detect stage's parts
template <int DIM, int TSIT>
void launch_slide_stage_t(rtod_ctx *c, hipStream_t s, const Step &st, const StageParts &p)
{
    const int ncls = (c->W + kDetectNT - 1) / kDetectNT;
    const int nfin = p.prev ? c->d.fin_parts : 0;
    const int nmb = p.merge ? (c->W + kMergeT - 1) / kMergeT : 0;
    const int nrm = p.runs ? run_merge_wgs(c->S) : 0;
    const int nts = p.ts ? c->d.cpb : 0;
    const int ndet = c->detect_blocks;
    const dim3 g(ncls + nfin + nmb + nrm + nts + ndet);
    const size_t lds = slide_stage_lds<DIM, TSIT>(nfin > 0, nmb > 0, nrm > 0, nts > 0, c->S);
    const Step pv = p.prev ? *p.prev : Step{}, mv = p.merge ? *p.merge : Step{},
               rv = p.runs ? *p.runs : Step{}, tv = p.ts ? *p.ts : Step{};
    // detect workgroups staging their sample table before the wait: twice
    // the busy ones of the last collected slide (a few slides back)
    constexpr int NW = kDetectNT / kWave;
    const int pre =
        c->units_last < 0 ? 64 : std::min(ndet, std::max(16, 2 * ((c->units_last + NW - 1) / NW)));
    hipLaunchKernelGGL((k_slide_stage<DIM, TSIT>), g, dim3(kDetectNT), lds, s, c->d, st, pv, mv,
                       rv, tv, ncls, nfin, nmb, nrm, nts, ndet, pre);
}

template <int DIM>
void launch_slide_stage_dim(rtod_ctx *c, hipStream_t s, const Step &st, const StageParts &p)
{
    // tiles of CH = kDetectNT * TSIT points (ride_along: CH is 512, 1024 or 2048)
    if (c->CH == kDetectNT)
        launch_slide_stage_t<DIM, 1>(c, s, st, p);
    else if (c->CH == 2 * kDetectNT)
        launch_slide_stage_t<DIM, 2>(c, s, st, p);
    else
        launch_slide_stage_t<DIM, 4>(c, s, st, p);
}

void launch_slide_stage(rtod_ctx *c, hipStream_t s, const Step &st, const StageParts &p)
{
    if (c->dim == 1)
        launch_slide_stage_dim<1>(c, s, st, p);
    else if (c->dim == 2)
        launch_slide_stage_dim<2>(c, s, st, p);
    else
        launch_slide_stage_dim<3>(c, s, st, p);
}

void launch_detect_stage(rtod_ctx *c, hipStream_t s, const Step &st, const StageParts &p)
{
    if (c->dim == 1)
        launch_detect_stage_dim<1>(c, s, st, p);
    else if (c->dim == 2)
        launch_detect_stage_dim<2>(c, s, st, p);
    else
        launch_detect_stage_dim<3>(c, s, st, p);
}

// LBVH index: codes, tile sort, merges up to U = W, hierarchy, boxes
void launch_lbvh_build(rtod_ctx *c, hipStream_t s, const Step &st)
{
    Lbvh L = c->L;
    // the leaf count: the slide's candidates (device address, read by the
    // kernels), or W for a whole-window tree
    L.n_cand = L.whole ? L.n_whole : &c->d.ctl->cnt[st.t % kSlideSets].n_cand;
    const dim3 gw((c->W + 255) / 256), b(256);
    if (c->dim == 1)
        hipLaunchKernelGGL(k_lbvh_codes<1>, gw, b, 0, s, c->d, st, L);
    else if (c->dim == 2)
        hipLaunchKernelGGL(k_lbvh_codes<2>, gw, b, 0, s, c->d, st, L);
    else
        hipLaunchKernelGGL(k_lbvh_codes<3>, gw, b, 0, s, c->d, st, L);
    constexpr int CH = 1024 * 4;
    hipLaunchKernelGGL((k_lbvh_sort<1024, 4>), dim3((c->W + CH - 1) / CH), dim3(1024), 0, s,
                       c->d, L);
    int src = 1;
    for (int p = 0, Lr = CH; p < c->lbvh_passes; ++p, Lr *= 2) {
        hipLaunchKernelGGL(k_lbvh_merge, gw, b, 0, s, c->d, L, Lr, src);
        src ^= 1;
    }
    hipLaunchKernelGGL(k_lbvh_settle, gw, b, 0, s, c->d, L, src);
    hipLaunchKernelGGL(k_lbvh_build, gw, b, 0, s, c->d, L);
    if (c->dim == 1)
        hipLaunchKernelGGL(k_lbvh_boxes<1>, gw, b, 0, s, c->d, st, L);
    else if (c->dim == 2)
        hipLaunchKernelGGL(k_lbvh_boxes<2>, gw, b, 0, s, c->d, st, L);
    else
        hipLaunchKernelGGL(k_lbvh_boxes<3>, gw, b, 0, s, c->d, st, L);
}

// whole-window LBVH, refit: the boxes recomputed over the topology built at
// rtod_prime (the reference's UPDATE_GAS_TYPE 0, optixScan.cpp:335-376)
void launch_lbvh_refit(rtod_ctx *c, hipStream_t s, const Step &st)
{
    Lbvh L = c->L;
    L.n_cand = L.n_whole;
    const dim3 gw((c->W + 255) / 256), b(256);
    hipLaunchKernelGGL(k_lbvh_unflag, gw, b, 0, s, c->d, L);
    if (c->dim == 1)
        hipLaunchKernelGGL(k_lbvh_boxes<1>, gw, b, 0, s, c->d, st, L);
    else if (c->dim == 2)
        hipLaunchKernelGGL(k_lbvh_boxes<2>, gw, b, 0, s, c->d, st, L);
    else
        hipLaunchKernelGGL(k_lbvh_boxes<3>, gw, b, 0, s, c->d, st, L);
}

// whole-window LBVH: the queries (candidates, or ring slots) count their
// neighbours over it
void launch_lbvh_count(rtod_ctx *c, hipStream_t s, const Step &st, bool stamps)
{
    Lbvh L = c->L;
    L.n_cand = L.n_whole;
    const int cands = c->cfg.ablation == RTOD_ABL_FILTER_ONLY ? 1 : 0;
    const int nq = cands ? c->W : std::min(c->W, c->ray_num > 0 ? c->ray_num : c->W);
    // one wavefront per lbvh_group queries (candidates, or sorted-window points)
    const int G = c->lbvh_group, per = (kLbvhNT / kWave) * G;
    const dim3 g((c->W + per - 1) / per), b(kLbvhNT);
    if (c->dim == 1)
        hipLaunchKernelGGL(k_lbvh_count<1>, g, b, 0, s, c->d, st, L, cands, nq, G);
    else if (c->dim == 2)
        hipLaunchKernelGGL(k_lbvh_count<2>, g, b, 0, s, c->d, st, L, cands, nq, G);
    else
        hipLaunchKernelGGL(k_lbvh_count<3>, g, b, 0, s, c->d, st, L, cands, nq, G);
    launch_finish(c, s, st, stamps);
}

void launch_lbvh_detect(rtod_ctx *c, hipStream_t s, const Step &st, bool stamps = false)
{
    if (c->L.whole) {
        launch_lbvh_count(c, s, st, stamps);
        return;
    }
    Lbvh L = c->L;
    L.n_cand = &c->d.ctl->cnt[st.t % kSlideSets].n_cand;
    // one wavefront per lbvh_group consecutive sorted-window points
    const int G = c->lbvh_group, per = (kLbvhNT / kWave) * G;
    const dim3 gq((c->W + per - 1) / per), bq(kLbvhNT);
    if (c->dim == 1)
        hipLaunchKernelGGL(k_lbvh_query<1>, gq, bq, 0, s, c->d, st, L, G);
    else if (c->dim == 2)
        hipLaunchKernelGGL(k_lbvh_query<2>, gq, bq, 0, s, c->d, st, L, G);
    else
        hipLaunchKernelGGL(k_lbvh_query<3>, gq, bq, 0, s, c->d, st, L, G);
    const dim3 gw((c->W + 255) / 256);
    hipLaunchKernelGGL(k_lbvh_select, gw, dim3(kDetectNT), 0, s, c->d, st, L);
    launch_finish(c, s, st, stamps);
}

// slide t's quantities; src: its new points in HBM, or nullptr for the
// pinned staging buffer
Step make_step_at(const rtod_ctx *c, uint64_t t, const float *src)
{
    Step st{};
    const uint64_t ne = (uint64_t)c->d.nent;
    st.t = t;
    st.u = (int)(t % (uint64_t)c->m);
    st.par = (int)(t & 1);
    st.eold = (int)(t % ne);
    st.enew = (int)(((uint64_t)c->m + t) % ne);
    st.buf = (int)(t % (uint64_t)c->stage_slots);
    const int q = (int)(t % rtod_ctx::kLearnRing);
    st.learn = c->learn_t[q] == (int64_t)t ? c->learn_of[q] : 0;
    st.src = src ? src : c->h_in + (size_t)st.buf * c->S * c->dim;
    return st;
}

// The learned schedule's cost rule for slide t (before its classify is
// launched; later calls keep the first choice).  Inference only pays where
// the prediction can shorten the slide: in 1-D a unit's near stencil is its
// whole stencil, so there is nothing to order; and with no more units than
// the detect keeps resident wavefronts every unit has a wavefront of its
// own, so the slide lasts as long as its slowest unit, which ordering does
// not shorten.  The rule reads the last collected slide's unit count (a
// slide or a few before t); RTOD_FLAG_LEARN_ALWAYS turns it off.
void decide_learn(rtod_ctx *c, int64_t t)
{
    const int q = (int)(t % rtod_ctx::kLearnRing);
    if (c->learn_t[q] == t)
        return;
    bool on = false;
    if (c->learned) {
        if (c->cfg.flags & RTOD_FLAG_LEARN_ALWAYS)
            on = true;
        else
            on = c->dim > 1 && c->units_last > c->learn_min_units;
    }
    c->learn_t[q] = t;
    c->learn_of[q] = on ? 1 : 0;
}

Step make_step(const rtod_ctx *c) { return make_step_at(c, (uint64_t)c->host_t, c->dev_src); }

// the host-out buffer slide t writes
HostOut *host_out_of(const rtod_ctx *c, int64_t t)
{
    return reinterpret_cast<HostOut *>(reinterpret_cast<char *>(c->h_out) +
                                       (size_t)(t % c->n_hout) * (size_t)c->d.h_out_bytes);
}

// The slide's kernel sequence in two stages.  Ingest stage: the new batch
// sorted into its batch-ring entry (ingest, batch merges).  Core stage:
// window merge, classify, [LBVH build], detect.  ev != nullptr: record an
// event before the first kernel and after each one.
void launch_ingest_stage(rtod_ctx *c, hipStream_t s, const Step &st, hipEvent_t *ev)
{
    if (ev)
        (void)hipEventRecord(ev[0], s);
    launch_ingest_dim(c, s, st, kModeSlide, nullptr, c->d.cpb, c->d.cpb == 1 ? 1 : 0);
    if (ev)
        (void)hipEventRecord(ev[1], s);
    if (c->d.cpb > 1)
        launch_batch_merges(c, s, st, kModeSlide, c->S);
}

// next != nullptr (pipelined slides, ride_along): the next slide's ingest
// stage rides in this stage's launches -- its tile sorts in the window
// merge's, its run merge in the detect's (after the LBVH detect on its own).
// Cell-stencil index only: prev != nullptr, the previous slide's epilogue
// rides in the window merge's launch; defer, this slide's epilogue is left
// to the next slide's window merge (a run's last slide must not defer),
// else it follows the detect in a launch of its own.
void launch_core_stage(rtod_ctx *c, hipStream_t s, const Step &st, hipEvent_t *ev,
                       const Step *next, const Step *prev = nullptr, bool defer = false,
                       bool stamps = false)
{
    if (ev)
        (void)hipEventRecord(ev[2], s);
    launch_merge_window(c, s, st, next, prev);
    if (ev)
        (void)hipEventRecord(ev[3], s);
    launch_classify(c, s, st);
    if (ev)
        (void)hipEventRecord(ev[4], s);
    const bool runs = next && c->d.cpb > 1;
    if (c->lbvh) {
        if (c->refit)
            launch_lbvh_refit(c, s, st);
        else
            launch_lbvh_build(c, s, st);
        if (ev)
            (void)hipEventRecord(ev[5], s);
        launch_lbvh_detect(c, s, st, stamps);
        if (ev)
            (void)hipEventRecord(ev[6], s);
        if (runs)
            launch_batch_merges(c, s, *next, kModeSlide, c->S);
    } else {
        launch_detect(c, s, st, runs ? next : nullptr, false);
        if (!defer)
            launch_finish(c, s, st, stamps);
        if (ev) {
            (void)hipEventRecord(ev[5], s);
            (void)hipEventRecord(ev[6], s);
        }
    }
}

void launch_slide(rtod_ctx *c, hipStream_t s, hipEvent_t *ev)
{
    decide_learn(c, c->host_t);
    const Step st = make_step(c);
    launch_ingest_stage(c, s, st, ev);
    launch_core_stage(c, s, st, ev, nullptr, nullptr, false, true);
}

rtod_status validate(const rtod_config *cfg)
{
    if (!cfg) {
        set_error("config is NULL");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (cfg->dim < 1 || cfg->dim > 3) {
        set_error("dim must be 1, 2 or 3");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (cfg->window < 1 || cfg->window > (1 << 28)) {
        set_error("window must be in [1, 2^28]");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (cfg->slide < 1 || cfg->slide > cfg->window || cfg->window % cfg->slide != 0) {
        set_error("slide must divide window (1 <= S <= W, W % S == 0)");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (cfg->K < 1) {
        set_error("K must be >= 1");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (!std::isfinite(cfg->R) || cfg->R < 1e-18f || cfg->R > 1e18f) {
        set_error("R must be finite and in [1e-18, 1e18]");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (cfg->index_kind < 0 || cfg->index_kind > 3 || cfg->sched_kind < 0 ||
        cfg->sched_kind > 1) {
        set_error("unknown index_kind or sched_kind");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (cfg->ablation < RTOD_ABL_NONE || cfg->ablation > RTOD_ABL_NO_FILTER) {
        set_error("unknown ablation");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (cfg->refit != 0 && cfg->refit != 1) {
        set_error("refit must be 0 or 1");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (cfg->launch_ray_num < 0 ||
        (cfg->launch_ray_num > 0 && cfg->ablation != RTOD_ABL_NO_FILTER)) {
        set_error("launch_ray_num (>= 0) applies to RTOD_ABL_NO_FILTER only, as the "
                  "reference's --launch_ray_num applies to OPTIMIZATION 0");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (cfg->refit &&
        (cfg->index_kind != RTOD_INDEX_LBVH || cfg->ablation == RTOD_ABL_NONE)) {
        set_error("refit applies to an LBVH over the whole window (RTOD_INDEX_LBVH with "
                  "RTOD_ABL_FILTER_ONLY or RTOD_ABL_NO_FILTER): the default LBVH holds the "
                  "slide's candidates, which change every slide");
        return RTOD_ERR_UNSUPPORTED;
    }
    if (cfg->sched_kind == RTOD_SCHED_LEARNED && cfg->ablation != RTOD_ABL_NONE) {
        set_error("the learned scheduler orders the default path's range counts");
        return RTOD_ERR_UNSUPPORTED;
    }
    if (cfg->sched_kind == RTOD_SCHED_LEARNED) {
        if (cfg->index_kind == RTOD_INDEX_LBVH) {
            set_error("the learned scheduler orders the cell-stencil range counts; "
                      "it does not apply to the LBVH index");
            return RTOD_ERR_UNSUPPORTED;
        }
        if (cfg->window > (1 << 22)) {
            set_error("the learned scheduler's replay (720 bytes per window point) is "
                      "limited to W <= 2^22");
            return RTOD_ERR_UNSUPPORTED;
        }
        if (!(cfg->miss_threshold >= 0.f && cfg->miss_threshold <= 1.f)) {
            set_error("miss_threshold must be in [0, 1] (0 selects the default 0.25)");
            return RTOD_ERR_INVALID_ARGUMENT;
        }
    }
    return RTOD_OK;
}

// He-uniform hidden weights from a fixed seed, zero biases: the untrained
// network outputs ~0.5 everywhere (every candidate a probable outlier, i.e.
// the plain schedule) until training moves it.
std::vector<float> initial_predictor()
{
    std::vector<float> p(kMlpNP, 0.f);
    uint64_t x = 0x9e3779b97f4a7c15ull;
    auto uni = [&x]() {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        return (double)(x >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    };
    for (int i = 0; i < kMlpIn * kMlpH; ++i)
        p[kOffW1 + i] = (float)(uni() * std::sqrt(6.0 / kMlpIn));
    for (int i = 0; i < kMlpH * kMlpH; ++i)
        p[kOffW2 + i] = (float)(uni() * std::sqrt(6.0 / kMlpH));
    for (int i = 0; i < kMlpH; ++i)
        p[kOffW3 + i] = (float)(uni() * std::sqrt(6.0 / kMlpH) * 0.1);
    return p;
}

// Options of contexts created as members of a batch.
struct CreateOpts {
    int sort_shape = -1;          // -1: picked from S and dim
    double cell_scale = 1.0;      // cell side multiplier (<= 1)
    hipStream_t stream = nullptr; // borrowed stream
    bool trainer = true;          // learned: own training stream (batches share one)
    bool batch_member = false;    // a stream of a batch (its own slide kernels)
};

// The kernel-argument structs as the device pass compiled them: the size of
// every struct that reaches a kernel and the offsets of the Dev fields the
// kernels read last, plus the hash of the source snapshot the library was
// built from (RTOD_SRC_HASH, set by the build).  The build compiles a
// private snapshot of the sources, so both passes see the same headers; a
// library whose passes disagreed anyway (or whose host code was linked with
// another device image) would hand its kernels misplaced pointers, so it is
// refused before any launch.
#ifndef RTOD_SRC_HASH
#define RTOD_SRC_HASH "unhashed"
#endif
constexpr int kAbiFields = 14;
#define RTOD_ABI_FIELDS                                                                           \
    {(int32_t)sizeof(Dev),         (int32_t)sizeof(Step),        (int32_t)sizeof(Ctl),           \
     (int32_t)sizeof(SlideCnt),    (int32_t)sizeof(HostOut),     (int32_t)sizeof(Learn),         \
     (int32_t)sizeof(LearnCtl),    (int32_t)sizeof(LearnShared), (int32_t)sizeof(Lbvh),          \
     (int32_t)sizeof(BatchItem),   (int32_t)sizeof(BatchStream), (int32_t)offsetof(Dev, ctl),    \
     (int32_t)offsetof(Dev, h_out), (int32_t)offsetof(Dev, lr)}
__device__ int32_t g_abi_layout[kAbiFields] = RTOD_ABI_FIELDS;
__device__ char g_src_hash[sizeof(RTOD_SRC_HASH)] = RTOD_SRC_HASH;

rtod_status check_abi_layout()
{
    int32_t dev[kAbiFields] = {};
    RTOD_HIP(hipMemcpyFromSymbol(dev, HIP_SYMBOL(g_abi_layout), sizeof(dev), 0,
                                 hipMemcpyDeviceToHost));
    const int32_t host[kAbiFields] = RTOD_ABI_FIELDS;
    char hash[sizeof(RTOD_SRC_HASH)] = {};
    RTOD_HIP(hipMemcpyFromSymbol(hash, HIP_SYMBOL(g_src_hash), sizeof(hash), 0,
                                 hipMemcpyDeviceToHost));
    if (std::memcmp(dev, host, sizeof(dev)) != 0 ||
        std::memcmp(hash, RTOD_SRC_HASH, sizeof(hash)) != 0) {
        set_error("librtod.so was built from inconsistent sources: kernel argument layout "
                  "or source hash differs between its host and device code; rebuild it");
        return RTOD_ERR_INTERNAL;
    }
    return RTOD_OK;
}

// The slide-run index (rtod_runs.h) serves 1-D streams on the default path:
// index AUTO or SLIDE_RUNS, no ablation, the plain schedule or the learned
// one under its cost rule (which runs 1-D slides on the plain kernels), a
// slide of at most two runs, at most kRunMaxRuns runs per window and a
// window whose outlier bitmap fits LDS.  why: the reason it does not apply.
bool runs_eligible(const rtod_config &cfg, const CreateOpts &opt, std::string *why)
{
    const int rps = (cfg.slide + kRunMaxLen - 1) / kRunMaxLen;
    const char *no = nullptr;
    if (opt.batch_member)
        no = "batched streams run on the cell-stencil index";
    else if (cfg.dim != 1)
        no = "it indexes 1-D streams";
    else if (cfg.index_kind != RTOD_INDEX_AUTO && cfg.index_kind != RTOD_INDEX_SLIDE_RUNS)
        no = "another index was requested";
    else if (cfg.ablation != RTOD_ABL_NONE)
        no = "the ablation modes run on the cell stencil or the LBVH";
    else if (cfg.sched_kind == RTOD_SCHED_LEARNED && (cfg.flags & RTOD_FLAG_LEARN_ALWAYS))
        no = "RTOD_FLAG_LEARN_ALWAYS runs the learned cell-stencil kernels";
    else if (rps > 2)
        no = "slides of more than 16384 points";
    else if ((int64_t)(cfg.window / cfg.slide) * rps > kRunMaxRuns)
        no = "more than 256 runs per window (W / S too large)";
    else if (cfg.window > kRunMaxW)
        no = "windows of more than 2^18 points";
    if (no && why)
        *why = no;
    return no == nullptr;
}

// the predictor's parameters, published copies, optimiser state and shared
// control block, set to the initial network
rtod_status alloc_predictor(rtod_ctx *c)
{
    Learn &lr = c->d.lr;
    RTOD_TRY(dalloc(c, &lr.params, (size_t)kMlpNP));
    RTOD_TRY(dalloc(c, &lr.pub, (size_t)2 * kMlpNP));
    RTOD_TRY(dalloc(c, &lr.adam_m, (size_t)kMlpNP));
    RTOD_TRY(dalloc(c, &lr.adam_v, (size_t)kMlpNP));
    RTOD_TRY(dalloc(c, &lr.ls, 1));
    RTOD_HIP(hipMemsetAsync(lr.ls, 0, sizeof(LearnShared), c->stream));
    std::vector<float> p0 = initial_predictor();
    RTOD_HIP(hipMemcpyAsync(lr.params, p0.data(), sizeof(float) * kMlpNP, hipMemcpyHostToDevice,
                            c->stream));
    for (int h = 0; h < 2; ++h)
        RTOD_HIP(hipMemcpyAsync(lr.pub + (size_t)h * kMlpNP, p0.data(), sizeof(float) * kMlpNP,
                                hipMemcpyHostToDevice, c->stream));
    RTOD_HIP(hipMemsetAsync(lr.adam_m, 0, sizeof(float) * kMlpNP, c->stream));
    RTOD_HIP(hipMemsetAsync(lr.adam_v, 0, sizeof(float) * kMlpNP, c->stream));
    RTOD_HIP(hipStreamSynchronize(c->stream));
    return RTOD_OK;
}

// A context on the slide-run index: the run ring (NE = m + 2 kRunBatch + 1
// entries of rps runs: the window's m slides plus the slides a pipelined
// launch detects and sorts ahead), the prime copy, kRunStageSlots pinned
// staging slots and kRunHostOuts host-out buffers.  A learned context keeps
// its predictor (rtod_get/set_predictor) but, in 1-D, never runs it.
rtod_status runs_create(rtod_ctx *c, const rtod_config *cfg)
{
    c->runs = true;
    RunsDev &r = c->rd;
    r.W = c->W;
    r.S = c->S;
    r.m = c->m;
    r.K = c->K;
    r.rps = (c->S + kRunMaxLen - 1) / kRunMaxLen;
    r.RL = (c->S + r.rps - 1) / r.rps;
    r.NE = c->m + 2 * kRunBatch + 1;
    r.n_hout = kRunHostOuts;
    const float R = cfg->R;
    r.r2 = R * R;
    r.inv_c = 1.0 / ((double)R * (1.0 - 1e-5));
    r.R_in = (double)R * (1.0 - 1e-5);
    r.R_out = (double)R * (1.0 + 1e-5);
    c->d.dim = 1;
    c->d.W = c->W;
    c->d.S = c->S;
    c->d.m = c->m;
    c->d.K = c->K;
    const size_t nrun = (size_t)r.NE * r.rps, RL = (size_t)r.RL;
    RTOD_TRY(dalloc(c, &r.rx, nrun * RL));
    RTOD_TRY(dalloc(c, &r.rslot, nrun * RL));
    RTOD_TRY(dalloc(c, &r.rpre, nrun * (kRunBMax + 1)));
    RTOD_TRY(dalloc(c, &r.rcell, nrun * RL));
    RTOD_TRY(dalloc(c, &r.rcs, nrun * (RL + 1)));
    RTOD_TRY(dalloc(c, &r.meta, nrun));
    RTOD_TRY(dalloc(c, &r.err, 1));
    RTOD_TRY(dalloc(c, &r.tstamp, (size_t)kTsCount));
    RTOD_TRY(dalloc(c, &c->d_prime, (size_t)c->W));
    RTOD_HIP(hipMemsetAsync(r.meta, 0, sizeof(RunMeta) * nrun, c->stream));
    RTOD_HIP(hipMemsetAsync(r.err, 0, sizeof(int32_t), c->stream));
    RTOD_HIP(hipMemsetAsync(r.tstamp, 0, sizeof(uint64_t) * kTsCount, c->stream));
    c->learned = cfg->sched_kind == RTOD_SCHED_LEARNED;
    if (c->learned)
        RTOD_TRY(alloc_predictor(c));
    c->list_cap = c->W;
    {
        int khz = 0;
        RTOD_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device));
        c->wall_khz = khz > 0 ? (double)khz : 100000.0;
    }
    if (const char *e = std::getenv("RTOD_FAIL_AT_SLIDE")) {
        char *end = nullptr;
        const long long v = std::strtoll(e, &end, 10);
        if (end != e && *end == '\0' && v >= 0)
            c->fail_at = v;
    }
    c->stage_slots = kRunStageSlots;
    c->n_hout = kRunHostOuts;
    const size_t in_bytes = (size_t)c->stage_slots * c->S * sizeof(float);
    const size_t out_bytes = (sizeof(HostOut) + (size_t)c->W * sizeof(int32_t) + 255) & ~(size_t)255;
    RTOD_HIP(hipHostMalloc((void **)&c->h_in, in_bytes, hipHostMallocMapped | hipHostMallocCoherent));
    RTOD_HIP(hipHostMalloc((void **)&c->h_out, c->n_hout * out_bytes,
                           hipHostMallocMapped | hipHostMallocCoherent));
    c->pinned_bytes = (int64_t)(in_bytes + c->n_hout * out_bytes);
    std::memset(c->h_in, 0, in_bytes);
    std::memset((void *)c->h_out, 0, c->n_hout * out_bytes);
    c->d.h_in = c->h_in;
    c->d.h_out = c->h_out;
    c->d.h_out_bytes = (int64_t)out_bytes;
    r.h_out = c->h_out;
    r.h_out_bytes = (int64_t)out_bytes;
    RTOD_HIP(hipStreamSynchronize(c->stream));
    return RTOD_OK;
}

rtod_status create_impl(const rtod_config *cfg, rtod_ctx **out, const CreateOpts &opt = {})
{
    RTOD_TRY(validate(cfg));
    rtod_ctx *c = new (std::nothrow) rtod_ctx();
    if (!c)
        return RTOD_ERR_NOMEM;
    *out = c;
    c->cfg = *cfg;
    c->dim = cfg->dim;
    c->W = cfg->window;
    c->S = cfg->slide;
    c->m = c->W / c->S;
    c->K = cfg->K;
    c->device = cfg->device;
    RTOD_HIP(hipSetDevice(c->device));
    RTOD_TRY(check_abi_layout());
    if (opt.stream) {
        c->stream = opt.stream;
        c->own_stream = false;
    } else {
        RTOD_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    }
    for (auto &e : c->ev)
        RTOD_HIP(hipEventCreate(&e));
    {
        std::string why;
        if (runs_eligible(*cfg, opt, &why))
            return runs_create(c, cfg);
        if (cfg->index_kind == RTOD_INDEX_SLIDE_RUNS) {
            set_error("the slide-run index does not apply: " + why);
            return RTOD_ERR_UNSUPPORTED;
        }
    }

    c->sort_shape = opt.sort_shape >= 0 ? opt.sort_shape : pick_sort_shape(c->S, c->dim);
    c->CH = kSortShapes[c->sort_shape].nt * kSortShapes[c->sort_shape].it;
    c->ride = ride_along(c->sort_shape, c->S, c->CH);

    Dev &d = c->d;
    d.dim = c->dim;
    d.W = c->W;
    d.S = c->S;
    d.m = c->m;
    d.K = c->K;
    d.cpb = (c->S + c->CH - 1) / c->CH;
    d.run_len = c->CH;
    int sl = 2; // sample stride >= 4, table <= kMaxSamples entries
    while (((int64_t)c->W + (1ll << sl) - 1) >> sl > kMaxSamples)
        ++sl;
    d.stride_log2 = sl;
    d.nsamp = (int)(((int64_t)c->W + (1ll << sl) - 1) >> sl);
    int bl = 3;
    while (((int64_t)c->S + (1ll << bl) - 1) >> bl > 512)
        ++bl;
    d.bs_log2 = bl;
    d.nbs = (int)(((int64_t)c->S + (1ll << bl) - 1) >> bl);
    const float R = cfg->R;
    d.r2 = R * R;
    const double cside =
        (double)R * (1.0 - 1e-5) / std::sqrt((double)c->dim) * opt.cell_scale;
    d.inv_c = 1.0 / cside;
    d.Rp = (double)R * (1.0 + 1e-5);

    const size_t W = c->W, S = c->S, dim = c->dim, m = c->m;
    // the tile scratch: the prime's segments (W), or two slides' tiles
    const size_t tmpn = W > 2 * S ? W : 2 * S;
    RTOD_TRY(dalloc(c, &d.ring, dim * W));
    d.tmpn = (int)tmpn;
    for (int p = 0; p < 2; ++p) {
        RTOD_TRY(dalloc(c, &d.akey[p], W));
        RTOD_TRY(dalloc(c, &d.aslot[p], W));
        RTOD_TRY(dalloc(c, &d.axyz[p], dim * W));
        RTOD_TRY(dalloc(c, &d.apos[p], W));
        RTOD_TRY(dalloc(c, &d.tkey[p], tmpn));
        RTOD_TRY(dalloc(c, &d.tslot[p], tmpn));
        RTOD_TRY(dalloc(c, &d.txyz[p], dim * tmpn));
    }
    d.nent = (int)m + 3;
    // single-stream slides sorted as several tiles and joined by the run
    // merge keep two slides' tiles (pipelined slides sort a slide's tiles
    // one launch before its run merge, beside the previous slide's run
    // merge); a batch's streams merge tiles in passes
    d.tile_par = (!opt.batch_member && use_run_merge(c->S, c->CH)) ? 1 : 0;
    d.unit_max = opt.batch_member ? kUnitCap : unit_max_single(c->dim);
    d.wait_spins = kClassifyWaitSpins;
    // LBVH walks: 16 queries per wavefront in 2-D and 3-D (a smaller box,
    // fewer leaves per group), 64 in 1-D; RTOD_LBVH_GROUP (16, 32 or 64)
    // overrides
    c->lbvh_group = cfg->dim == 1 ? 64 : 16;
    if (const char *e = std::getenv("RTOD_LBVH_GROUP")) {
        const int v = std::atoi(e);
        if (v == 16 || v == 32 || v == 64)
            c->lbvh_group = v;
    }
    if (const char *e = std::getenv("RTOD_FAIL_AT_SLIDE")) {
        char *end = nullptr;
        const long long v = std::strtoll(e, &end, 10);
        if (end != e && *end == '\0' && v >= 0)
            c->fail_at = v;
    }
    if (const char *e = std::getenv("RTOD_CLASSIFY_WAIT_SPINS")) {
        char *end = nullptr;
        const long v = std::strtol(e, &end, 10);
        if (end != e && *end == '\0' && v >= 0 && v <= kClassifyWaitSpins)
            d.wait_spins = (int)v;
    }
    const size_t ne = (size_t)d.nent;
    RTOD_TRY(dalloc(c, &d.bkey, ne * S));
    RTOD_TRY(dalloc(c, &d.bslot, ne * S));
    RTOD_TRY(dalloc(c, &d.bsamp, ne * (size_t)d.nbs));
    RTOD_TRY(dalloc(c, &d.bxyz, dim * ne * S));
    RTOD_TRY(dalloc(c, &d.samples[0], (size_t)d.nsamp));
    RTOD_TRY(dalloc(c, &d.samples[1], (size_t)d.nsamp));
    d.nsamp4 = d.stride_log2 > 2 ? (int)(((int64_t)c->W + 3) >> 2) : 0;
    if (d.nsamp4) {
        RTOD_TRY(dalloc(c, &d.samples4[0], (size_t)d.nsamp4));
        RTOD_TRY(dalloc(c, &d.samples4[1], (size_t)d.nsamp4));
    }
    RTOD_TRY(dalloc(c, &d.cand, W));
    RTOD_TRY(dalloc(c, &d.units, W));
    RTOD_TRY(dalloc(c, &d.cand_cnt, W));
    d.bm_words = (((W + 31) / 32) + 63) / 64 * 64;
    // one epilogue part per chunk of the bitmap
    d.fin_parts = std::min(32, std::max(1, (int)(((W + 31) / 32 + kChunkWords - 1) / kChunkWords)));
    RTOD_TRY(dalloc(c, &d.bitmap, (size_t)kSlideSets * d.bm_words));
    if (!opt.batch_member) // a batch's streams never run one launch per slide
        RTOD_TRY(dalloc(c, &d.ready, (size_t)kSlideSets * kReadyFlags * kReadyStride));
    RTOD_TRY(dalloc(c, &d.ctl, 1));
    RTOD_TRY(dalloc(c, &d.tstamp, (size_t)kTsCount));
    RTOD_TRY(dalloc(c, &c->d_prime, dim * W));
    RTOD_HIP(hipMemsetAsync(d.bitmap, 0, sizeof(uint32_t) * kSlideSets * d.bm_words, c->stream));
    RTOD_HIP(hipMemsetAsync(d.ctl, 0, sizeof(Ctl), c->stream));
    RTOD_HIP(hipMemsetAsync(d.tstamp, 0, sizeof(uint64_t) * kTsCount, c->stream));
    if (d.ready)
        RTOD_HIP(hipMemsetAsync(d.ready, 0,
                                sizeof(int32_t) * kSlideSets * kReadyFlags * kReadyStride,
                                c->stream));

    c->lbvh = cfg->index_kind == RTOD_INDEX_LBVH;
    if (c->lbvh) {
        Lbvh &L = c->L;
        for (int p = 0; p < 2; ++p) {
            RTOD_TRY(dalloc(c, &L.code[p], W));
            RTOD_TRY(dalloc(c, &L.item[p], W));
        }
        RTOD_TRY(dalloc(c, &L.left, W));
        RTOD_TRY(dalloc(c, &L.right, W));
        RTOD_TRY(dalloc(c, &L.parent, 2 * W));
        RTOD_TRY(dalloc(c, &L.box, 2 * W * 6));
        RTOD_TRY(dalloc(c, &L.flag, W));
        RTOD_TRY(dalloc(c, &L.count, W));
        RTOD_TRY(dalloc(c, &d.cmask, (W + 63) / 64));
        int passes = 0;
        for (size_t Lr = 4096; Lr < W; Lr *= 2)
            ++passes;
        c->lbvh_passes = passes;
        // the ablations' tree over the whole window (leaf items: ring slots)
        L.whole = cfg->ablation != RTOD_ABL_NONE ? 1 : 0;
        RTOD_TRY(dalloc(c, &L.n_whole, 1));
        const int32_t w32 = (int32_t)W;
        RTOD_HIP(hipMemcpy(L.n_whole, &w32, sizeof(w32), hipMemcpyHostToDevice));
    }
    c->refit = cfg->refit != 0;
    c->ray_num = cfg->launch_ray_num;
    c->list_cap = std::max(c->W, c->ray_num);
    d.no_filter = cfg->ablation == RTOD_ABL_NO_FILTER ? 1 : 0;
    d.ray_lim = (!c->lbvh && d.no_filter && c->ray_num > 0 && c->ray_num < c->W) ? c->ray_num : 0;
    c->learned = cfg->sched_kind == RTOD_SCHED_LEARNED;
    d.lr.on = c->learned ? 1 : 0;
    {
        // the detect's resident wavefronts: at least 6 per SIMD, 4 SIMDs per CU
        int cus = 0;
        RTOD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
        c->learn_min_units = std::max(1, cus) * 4 * 6;
    }
    {
        int khz = 0;
        RTOD_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device));
        c->wall_khz = khz > 0 ? (double)khz : 100000.0;
    }
    if (c->learned) {
        Learn &lr = d.lr;
        lr.miss_thr = cfg->miss_threshold > 0.f ? cfg->miss_threshold : 0.25f;
        RTOD_TRY(dalloc(c, &lr.params, (size_t)kMlpNP));
        RTOD_TRY(dalloc(c, &lr.pub, (size_t)2 * kMlpNP));
        RTOD_TRY(dalloc(c, &lr.adam_m, (size_t)kMlpNP));
        RTOD_TRY(dalloc(c, &lr.adam_v, (size_t)kMlpNP));
        RTOD_TRY(dalloc(c, &lr.rep_x, (size_t)kReplaySlides * W * kFeatStride));
        RTOD_TRY(dalloc(c, &lr.rep_c, (size_t)kReplaySlides * W));
        RTOD_TRY(dalloc(c, &lr.rep_n, (size_t)2 * kReplaySlides));
        RTOD_TRY(dalloc(c, &lr.lc, 1));
        RTOD_TRY(dalloc(c, &lr.ls, 1));
        RTOD_HIP(hipMemsetAsync(lr.rep_n, 0, sizeof(int32_t) * 2 * kReplaySlides, c->stream));
        RTOD_HIP(hipMemsetAsync(lr.lc, 0, sizeof(LearnCtl), c->stream));
        RTOD_HIP(hipMemsetAsync(lr.ls, 0, sizeof(LearnShared), c->stream));
        std::vector<float> p0 = initial_predictor();
        RTOD_HIP(hipMemcpyAsync(lr.params, p0.data(), sizeof(float) * kMlpNP,
                                hipMemcpyHostToDevice, c->stream));
        for (int h = 0; h < 2; ++h)
            RTOD_HIP(hipMemcpyAsync(lr.pub + (size_t)h * kMlpNP, p0.data(),
                                    sizeof(float) * kMlpNP, hipMemcpyHostToDevice, c->stream));
        if (opt.trainer) {
            RTOD_HIP(hipStreamCreateWithFlags(&c->train_stream, hipStreamNonBlocking));
            RTOD_HIP(hipEventCreateWithFlags(&c->train_ev, hipEventDisableTiming));
        }
        RTOD_HIP(hipMemsetAsync(lr.adam_m, 0, sizeof(float) * kMlpNP, c->stream));
        RTOD_HIP(hipMemsetAsync(lr.adam_v, 0, sizeof(float) * kMlpNP, c->stream));
    }
    // staging: a batch's streams use two slots (slide t in t & 1); a single
    // stream kStageSlots, so that pipelined slides from host memory can be
    // staged three ahead of the slide being detected (rtod_slides)
    c->stage_slots =
        (opt.batch_member || S * dim * sizeof(float) > kStageSlotCap) ? 2 : kStageSlots;
    const size_t in_bytes = (size_t)c->stage_slots * S * dim * sizeof(float);
    // kHostOuts host-out buffers (slide t: t % kHostOuts), each aligned to
    // 256 bytes
    const size_t out_bytes = (sizeof(HostOut) + W * sizeof(int32_t) + 255) & ~(size_t)255;
    RTOD_HIP(hipHostMalloc((void **)&c->h_in, in_bytes, hipHostMallocMapped | hipHostMallocCoherent));
    RTOD_HIP(hipHostMalloc((void **)&c->h_out, kHostOuts * out_bytes,
                           hipHostMallocMapped | hipHostMallocCoherent));
    c->pinned_bytes = (int64_t)(in_bytes + kHostOuts * out_bytes);
    std::memset(c->h_in, 0, in_bytes);
    std::memset((void *)c->h_out, 0, kHostOuts * out_bytes);
    d.h_in = c->h_in;
    d.h_out = c->h_out;
    d.h_out_bytes = (int64_t)out_bytes;
    RTOD_HIP(hipStreamSynchronize(c->stream));
    return RTOD_OK;
}

rtod_status runs_prime(rtod_ctx *c, const float *pts, hipMemcpyKind kind);
rtod_status runs_run_slide(rtod_ctx *c, const float *src, double t_entry, double t_h2d,
                           int32_t *outs, int32_t *n_out, rtod_times *times);

rtod_status prime_impl(rtod_ctx *c, const float *pts, int64_t n,
                       hipMemcpyKind kind = hipMemcpyHostToDevice)
{
    if (!pts || n != c->W) {
        set_error("prime needs exactly W points");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    RTOD_HIP(hipSetDevice(c->device));
    if (c->runs)
        return runs_prime(c, pts, kind);
    if (c->train_stream) { // the replay is about to be cleared
        RTOD_HIP(hipStreamSynchronize(c->train_stream));
        c->train_pending = false;
    }
    hipStream_t s = c->stream;
    Step st{};
    RTOD_HIP(hipMemcpyAsync(c->d_prime, pts, sizeof(float) * (size_t)c->W * c->dim, kind, s));
    // sort each initial batch (generation b -> batch-ring entry b)
    launch_ingest_dim(c, s, st, kModePrime, c->d_prime, c->m * c->d.cpb, c->d.cpb == 1 ? 1 : 0);
    if (c->d.cpb > 1)
        launch_batch_merges(c, s, st, kModePrime, c->W);
    // merge the m sorted batches (entries 0..m-1 are contiguous) into the
    // sorted window A[0]; batch positions ride along: the first pass derives
    // them from the index inside the batch ring, later passes copy them
    {
        const size_t bxs = (size_t)c->d.nent * c->S;
        const uint64_t *sk = c->d.bkey;
        const int32_t *ss = c->d.bslot;
        const float *sx = c->d.bxyz;
        size_t sxs = bxs;
        const int32_t *sp = nullptr;
        int src = 0;
        int L = c->S;
        for (;;) {
            const bool last = 2 * (int64_t)L >= c->W;
            uint64_t *dk = last ? c->d.akey[0] : c->d.tkey[src ^ 1];
            int32_t *ds = last ? c->d.aslot[0] : c->d.tslot[src ^ 1];
            float *dx = last ? c->d.axyz[0] : c->d.txyz[src ^ 1];
            const size_t dxs = last ? (size_t)c->W : (size_t)c->d.tmpn;
            int32_t *dp = last ? c->d.apos[0] : ((src ^ 1) ? c->d.apos[1] : c->d.cand_cnt);
            launch_merge_pass(c, s, st, sk, ss, sx, sxs, sp, dk, ds, dx, dxs, dp, c->W, c->W, L, 0,
                              (int)kModePrime);
            src ^= 1;
            sk = c->d.tkey[src];
            ss = c->d.tslot[src];
            sx = c->d.txyz[src];
            sxs = (size_t)c->d.tmpn;
            sp = dp;
            if (last)
                break;
            L *= 2;
        }
    }
    hipLaunchKernelGGL(k_window_samples,
                       dim3((std::max(c->d.nsamp, c->d.nsamp4) + 255) / 256), dim3(256), 0, s, c->d,
                       0);
    // refit: the whole-window tree is built once, over the initial window
    // (the ring the ingest just filled), and only refit by the slides
    if (c->refit)
        launch_lbvh_build(c, s, st);
    RTOD_HIP(hipMemsetAsync(c->d.bitmap, 0, sizeof(uint32_t) * kSlideSets * c->d.bm_words, s));
    RTOD_HIP(hipMemsetAsync(c->d.ctl, 0, sizeof(Ctl), s));
    if (c->learned) { // replay and per-slide counters restart; weights stay
        RTOD_HIP(hipMemsetAsync(c->d.lr.rep_n, 0, sizeof(int32_t) * 2 * kReplaySlides, s));
        RTOD_HIP(hipMemsetAsync(c->d.lr.lc, 0, sizeof(LearnCtl), s));
    }
    RTOD_HIP(hipGetLastError());
    RTOD_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < kHostOuts; ++k)
        host_out_of(c, k)->slide = 0;
    for (auto &q : c->learn_t)
        q = -1;
    c->host_t = 0;
    c->primed = true;
    c->staged = false;
    c->phase = 0;
    return RTOD_OK;
}

// the messages of the device error flags set in `err`
std::string err_text(int32_t err)
{
    static const struct {
        int32_t bit;
        const char *what;
    } kMsgs[] = {{kErrMergeSurvivor, "a window-merge survivor ranked out of range"},
                 {kErrMergeNew, "a window-merge new point ranked out of range"},
                 {kErrStencilRows, "a detect unit's stencil exceeded 32 rows"},
                 {kErrClassifyWait, "a detect timed out waiting for its classify"},
                 {kErrLbvhWalk, "an LBVH walk overflowed its node stack (malformed tree)"},
                 {kErrRunBucket, "a slide-run point's bucket fell outside its run's range"}};
    std::string s;
    for (const auto &m : kMsgs)
        if (err & m.bit)
            s += std::string(" ") + m.what + ";";
    return s.empty() ? std::string(" unknown flag") : s;
}

rtod_status collect(rtod_ctx *c, int32_t *outs, int32_t *n_out)
{
    HostOut *ho = host_out_of(c, c->host_t);
    if (c->host_t == c->fail_at) {
        set_error("failure injected at slide " + std::to_string(c->fail_at) +
                  " (RTOD_FAIL_AT_SLIDE)");
        c->host_t++;
        return RTOD_ERR_INTERNAL;
    }
    if (ho->slide != c->host_t + 1) {
        set_error("device slide counter out of step with the host");
        return RTOD_ERR_INTERNAL;
    }
    if (ho->err != 0) {
        set_error("device consistency check failed (code " + std::to_string(ho->err) + ":" +
                  err_text(ho->err) + ")");
        c->host_t++;
        return RTOD_ERR_INTERNAL;
    }
    const int n = ho->n_out;
    if (n < 0 || n > c->W) {
        set_error("outlier count out of range");
        return RTOD_ERR_INTERNAL;
    }
    int nl = n;
    if (c->ray_num > c->W) {
        // launch_ray_num > W: ray r starts from slot r mod W (as the
        // reference's raygen, optixScan.cu:29-32), so the ray indices whose
        // count is <= K are the listed slots repeated every W, below ray_num
        nl = 0;
        for (int64_t base = 0; base < c->ray_num; base += c->W)
            for (int i = 0; i < n && base + ho->list[i] < c->ray_num; ++i)
                if (outs)
                    outs[nl++] = (int32_t)(base + ho->list[i]);
                else
                    ++nl;
    } else if (outs && n) {
        std::memcpy(outs, ho->list, sizeof(int32_t) * (size_t)n);
    }
    if (n_out)
        *n_out = nl;
    c->host_t++;
    c->units_last = ho->n_units;
    c->cand_last = ho->n_cand;
    c->cand_total += ho->n_cand;
    c->tests_last = (int64_t)ho->pair_tests;
    c->tests_total += (int64_t)ho->pair_tests;
    c->hits_last = (int64_t)ho->hits;
    c->hits_total += (int64_t)ho->hits;
    c->out_last = n;
    c->out_total += n;
    if (c->learned) {
        c->learn_new = ho->learn_kernels != 0;
        if (c->learn_new) {
            c->misses_total += ho->misses;
            c->pred_points += ho->n_units;
        } else {
            c->plain_slides++;
        }
        c->used_pred_last = ho->used_pred != 0;
        if (!ho->used_pred)
            c->fallbacks++;
        if (ho->train_steps > c->train_steps)
            c->train_loss_last = ho->train_loss;
        c->train_steps = ho->train_steps;
    }
    return RTOD_OK;
}

// Core-stage kernel durations of a pipelined slide (events 2..6): the window
// merge and the detect carry the next slide's ingest work when `fused`.
void accumulate_core_times(rtod_ctx *c, bool fused)
{
    float ms[4];
    for (int i = 0; i < 4; ++i) {
        ms[i] = 0.f;
        (void)hipEventElapsedTime(&ms[i], c->ev[2 + i], c->ev[3 + i]);
    }
    const int mw = fused ? kKMergeWindowIngest : kKMergeWindow;
    c->ksum[mw] += ms[0];
    c->kcnt[mw]++;
    c->ksum[kKClassify] += ms[1];
    c->kcnt[kKClassify]++;
    if (c->lbvh) {
        c->ksum[kKLbvhBuild] += ms[2];
        c->kcnt[kKLbvhBuild]++;
        c->ksum[kKLbvhQuery] += ms[3];
        c->kcnt[kKLbvhQuery]++;
    } else {
        const int dk = (fused && c->d.cpb > 1) ? kKDetectRuns : kKDetect;
        c->ksum[dk] += ms[2];
        c->kcnt[dk]++;
    }
}

// event pair [2, 3]: the classify stage; [3, 4]: the detect stage
void accumulate_stage_times(rtod_ctx *c, bool fused)
{
    if (fused) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, c->ev[2], c->ev[3]);
        c->ksum[kKSlideStage] += ms;
        c->kcnt[kKSlideStage]++;
        return;
    }
    float ms[2];
    for (int i = 0; i < 2; ++i) {
        ms[i] = 0.f;
        (void)hipEventElapsedTime(&ms[i], c->ev[2 + i], c->ev[3 + i]);
    }
    c->ksum[kKClassifyStage] += ms[0];
    c->kcnt[kKClassifyStage]++;
    c->ksum[kKDetectStage] += ms[1];
    c->kcnt[kKDetectStage]++;
}

void accumulate_kernel_times(rtod_ctx *c, double *phase)
{
    float ms[6];
    for (int i = 0; i < 6; ++i) {
        ms[i] = 0.f;
        (void)hipEventElapsedTime(&ms[i], c->ev[i], c->ev[i + 1]);
    }
    c->ksum[kKIngest] += ms[0];
    c->kcnt[kKIngest]++;
    if (c->d.cpb > 1) {
        c->ksum[kKMergePass] += ms[1];
        c->kcnt[kKMergePass]++;
    }
    c->ksum[kKMergeWindow] += ms[2];
    c->kcnt[kKMergeWindow]++;
    c->ksum[kKClassify] += ms[3];
    c->kcnt[kKClassify]++;
    if (c->lbvh) {
        c->ksum[kKLbvhBuild] += ms[4];
        c->kcnt[kKLbvhBuild]++;
        c->ksum[kKLbvhQuery] += ms[5];
        c->kcnt[kKLbvhQuery]++;
    } else {
        c->ksum[kKDetect] += ms[4];
        c->kcnt[kKDetect]++;
    }
    if (phase) {
        // the cell-sorted window (merge) is the cell-stencil index; with the
        // LBVH it only feeds the grid filter and the index is the LBVH
        phase[0] = (double)ms[0] + ms[1] + ms[3] + (c->lbvh ? ms[2] : 0.f); // prepare cell
        phase[1] = c->lbvh ? (double)ms[4] : (double)ms[2];                 // build index
        phase[2] = c->lbvh ? (double)ms[5] : (double)ms[4];                 // detect
    }
}

// Learned schedule: one training launch on the trainer's stream after a
// slide, unless the previous one is still running.  It overlaps the next
// slide's kernels; the weights it publishes are picked up by whichever
// classify starts after it finishes.
rtod_status maybe_train(rtod_ctx *c)
{
    // a slide on the plain kernels added nothing to the replay
    if (!c->learned || !c->train_stream || !c->learn_new)
        return RTOD_OK;
    if (c->train_pending) {
        const hipError_t q = hipEventQuery(c->train_ev);
        if (q == hipErrorNotReady) {
            (void)hipGetLastError(); // not an error: still running
            return RTOD_OK;
        }
        RTOD_HIP(q);
        c->train_pending = false;
    }
    hipLaunchKernelGGL(k_train, dim3(1), dim3(kTrainNT), 0, c->train_stream, c->d,
                       (uint64_t)c->host_t, kTrainSteps);
    RTOD_HIP(hipGetLastError());
    RTOD_HIP(hipEventRecord(c->train_ev, c->train_stream));
    c->train_pending = true;
    c->train_t_next = c->host_t;
    return RTOD_OK;
}

// Slide t rewrites its replay slot (classify, detect on stream s): past the
// pending training launch's skipped slides (kReplaySkip), s first waits for
// it.
rtod_status guard_trainer(rtod_ctx *c, int64_t t, hipStream_t s)
{
    if (!c->train_pending || t < c->train_t_next + kReplaySkip)
        return RTOD_OK;
    const hipError_t q = hipEventQuery(c->train_ev);
    if (q == hipSuccess) {
        c->train_pending = false;
        return RTOD_OK;
    }
    if (q != hipErrorNotReady)
        RTOD_HIP(q);
    (void)hipGetLastError(); // still running: not an error
    RTOD_HIP(hipStreamWaitEvent(s, c->train_ev, 0));
    return RTOD_OK;
}

rtod_status wait_slide(rtod_ctx *c, int64_t t);

// the device stamps of a slide run one per call (its kernels' starts, in
// its host-out buffer) as the reference's buckets, in ms: prepare cell (the
// ingest, the batch merges and the classify; with the LBVH also the window
// merge), build index (the cell-sorted window merge, or the LBVH build) and
// detect; out[3] the epilogue (list into host memory)
void stamp_phases(const rtod_ctx *c, const HostOut *ho, double *phase)
{
    const double k = 1.0 / c->wall_khz;
    auto span = [&](int a, int b) {
        return ho->ts[b] >= ho->ts[a] ? (double)(ho->ts[b] - ho->ts[a]) * k : 0.0;
    };
    if (c->lbvh) {
        phase[0] = span(kTsIngest, kTsIndex);
        phase[1] = span(kTsIndex, kTsDetect);
    } else {
        phase[0] = span(kTsIngest, kTsMerge) + span(kTsClassify, kTsDetect);
        phase[1] = span(kTsMerge, kTsClassify);
    }
    phase[2] = span(kTsDetect, kTsFinish);
    phase[3] = span(kTsFinish, kTsPublished);
}

// One slide whose new points are already where the ingest kernel reads them
// (pinned staging, or c->dev_src in HBM).  The host waits for the slide's
// stamp in its host-out buffer (polling, not a stream synchronisation: that
// costs ~5 us more per slide); the breakdown comes from the kernels' device
// stamps.  RTOD_FLAG_PHASE_TIMES also records HIP events between the
// kernels (per-kernel averages, rtod_kernel_times) and waits for them.
rtod_status run_slide(rtod_ctx *c, double t_entry, double t_h2d, int32_t *outs, int32_t *n_out,
                      rtod_times *times)
{
    double phase[4] = {0, 0, 0, 0};
    const bool timed = (c->cfg.flags & RTOD_FLAG_PHASE_TIMES) != 0;
    RTOD_TRY(guard_trainer(c, c->host_t, c->stream));
    launch_slide(c, c->stream, timed ? c->ev : nullptr);
    RTOD_HIP(hipGetLastError());
    const rtod_status ws = wait_slide(c, c->host_t);
    if (ws != RTOD_OK) { // the launches ended without publishing the slide
        (void)hipStreamSynchronize(c->stream);
        c->primed = false;
        return ws;
    }
    if (timed) {
        RTOD_HIP(hipStreamSynchronize(c->stream));
        double unused[3];
        accumulate_kernel_times(c, unused);
    }
    stamp_phases(c, host_out_of(c, c->host_t), phase);
    const double t_col = now_ms();
    RTOD_TRY(collect(c, outs, n_out));
    const double t_end = now_ms();
    RTOD_TRY(maybe_train(c));
    if (times) {
        times->copy_h2d = t_h2d - t_entry;
        times->prepare_cell = phase[0];
        times->build_index = phase[1];
        times->detect_outlier = phase[2];
        times->copy_d2h = phase[3] + (t_end - t_col);
        times->total = t_end - t_entry;
        times->launch = times->total - times->copy_h2d - times->prepare_cell -
                        times->build_index - times->detect_outlier - times->copy_d2h;
    }
    return RTOD_OK;
}

rtod_status slide_checks(rtod_ctx *c, const void *new_pts, int32_t *outs, int32_t *n_out)
{
    if (!c->primed) {
        set_error("rtod_slide before rtod_prime");
        return RTOD_ERR_STATE;
    }
    if (!new_pts || !outs || !n_out) {
        set_error("NULL argument");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (c->staged) {
        set_error("a phase-API slide is in progress");
        return RTOD_ERR_STATE;
    }
    return RTOD_OK;
}

rtod_status slide_impl(rtod_ctx *c, const float *new_pts, int32_t *outs, int32_t *n_out,
                       rtod_times *times)
{
    const double t_entry = now_ms();
    RTOD_TRY(slide_checks(c, new_pts, outs, n_out));
    RTOD_HIP(hipSetDevice(c->device));
    const size_t nf = (size_t)c->S * c->dim;
    // a slide in a buffer from rtod_host_alloc is read in place over PCIe
    if (const float *dp = pinned_device_ptr(new_pts, nf * sizeof(float))) {
        c->dev_src = dp;
        const rtod_status s = c->runs
                                  ? runs_run_slide(c, dp, t_entry, t_entry, outs, n_out, times)
                                  : run_slide(c, t_entry, t_entry, outs, n_out, times);
        c->dev_src = nullptr;
        return s;
    }
    // a slide run one per call is staged in slot 0 on the slide-run index:
    // every launch that read a staging slot has completed by now, and the
    // device's translations of that slot's pages stay warm
    float *slot = c->h_in + (c->runs ? 0 : (size_t)(c->host_t % c->stage_slots) * nf);
    std::memcpy(slot, new_pts, nf * sizeof(float));
    c->dev_src = nullptr;
    if (c->runs)
        return runs_run_slide(c, slot, t_entry, now_ms(), outs, n_out, times);
    return run_slide(c, t_entry, now_ms(), outs, n_out, times);
}

rtod_status slide_device_impl(rtod_ctx *c, const float *dev_pts, int32_t *outs, int32_t *n_out,
                              rtod_times *times)
{
    const double t_entry = now_ms();
    RTOD_TRY(slide_checks(c, dev_pts, outs, n_out));
    RTOD_HIP(hipSetDevice(c->device));
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, dev_pts) != hipSuccess || attr.type != hipMemoryTypeDevice) {
        (void)hipGetLastError();
        set_error("rtod_slide_device needs a device (HBM) pointer");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    c->dev_src = dev_pts;
    rtod_status s = c->runs ? runs_run_slide(c, dev_pts, t_entry, t_entry, outs, n_out, times)
                            : run_slide(c, t_entry, t_entry, outs, n_out, times);
    c->dev_src = nullptr;
    return s;
}

// Wait for slide t's stamp in its host-out buffer (written last, after a
// system-scope release, by the slide's detect).  The stream is queried now
// and then, so a failed launch surfaces as an error instead of a hang.
rtod_status wait_slide(rtod_ctx *c, int64_t t)
{
    const int64_t *stamp = &host_out_of(c, t)->slide;
    const int64_t want = t + 1;
    // the stream is queried only once the wait has lasted kQueryMs (and then
    // every kQueryMs): a query takes microseconds, and one that overlaps the
    // stamp's arrival delays the slide by as much
    constexpr double kQueryMs = 0.2;
    double next_query = -1.0;
    for (uint64_t spin = 1;; ++spin) {
        if (__atomic_load_n(stamp, __ATOMIC_ACQUIRE) == want)
            return RTOD_OK;
        if ((spin & 63) == 0) {
            const double now = now_ms();
            if (next_query < 0.0)
                next_query = now + kQueryMs;
            if (now < next_query)
                continue;
            next_query = now + kQueryMs;
            const hipError_t q = hipStreamQuery(c->stream);
            if (q == hipSuccess) {
                if (__atomic_load_n(stamp, __ATOMIC_ACQUIRE) == want)
                    return RTOD_OK;
                set_error("the stream went idle without completing the slide");
                return RTOD_ERR_INTERNAL;
            }
            if (q != hipErrorNotReady) {
                set_error(std::string("HIP error: ") + hipGetErrorString(q));
                return RTOD_ERR_HIP;
            }
            (void)hipGetLastError(); // not ready: not an error
        }
        __builtin_ia32_pause();
    }
}

// Where the slides of a pipelined run come from: HBM (dev, slide i at
// dev + i*S*dim) or host memory (host, same layout).  Host slides are copied
// into pinned staging slot (t0 + i) % stage_slots just before the launch
// whose ingest reads them (zero-copy over PCIe): the copy is inside the
// run, as the reference's per-slide H2D is inside its loop
// (optixScan.cpp:1051-1058).
struct SlideSrc {
    const float *dev = nullptr;
    const float *host = nullptr;
    bool pinned = false; // dev: a rtod_host_alloc buffer's device address
};

const float *slide_src(const rtod_ctx *c, const SlideSrc &src, int64_t t0, int32_t i)
{
    const size_t per = (size_t)c->S * c->dim;
    if (src.dev)
        return src.dev + (size_t)i * per;
    return c->h_in + (size_t)((t0 + i) % c->stage_slots) * per;
}

void stage_slide(rtod_ctx *c, const SlideSrc &src, int64_t t0, int32_t i)
{
    if (!src.host)
        return;
    const size_t per = (size_t)c->S * c->dim;
    stream_copy(c->h_in + (size_t)((t0 + i) % c->stage_slots) * per, src.host + (size_t)i * per,
                per);
}

// n slides from HBM or host memory on the context's stream, up to kHostOuts
// in flight: the host enqueues slide i once slide i - kHostOuts is collected (polling its
// stamp), and slide i+1's ingest stage rides in slide i's core launches
// (ride_along; else it follows them).  The results are those of n
// rtod_slide_device calls.  On an error the stream is drained first.
// RTOD_FLAG_PHASE_TIMES: HIP events around the core kernels, and each slide
// is waited for before the next is enqueued (a measuring mode).
// Pipelined slides in stage launches (cell-stencil index, the ingest
// riding; see k_classify_stage): after a prologue (slide t0's ingest and
// window merge, the ingest of slides t0 + 1 and t0 + 2), slide t0 + i is one
// launch, k_slide_stage (plain schedule), or two, the classify stage and the
// detect stage (learned schedule), and the last slide's epilogue follows in
// a launch of its own.  Up to kHostOuts slides are in flight.
rtod_status slides_two_stage(rtod_ctx *c, const SlideSrc &src, int32_t n, int32_t *offs,
                             int32_t *slots)
{
    const int64_t t0 = c->host_t;
    const bool timed = (c->cfg.flags & RTOD_FLAG_PHASE_TIMES) != 0;
    const bool runs = c->d.cpb > 1;
    hipStream_t s = c->stream;
    auto step = [&](int32_t i) {
        return make_step_at(c, (uint64_t)(t0 + i), slide_src(c, src, t0, i));
    };
    int64_t used = 0;
    int32_t done = 0;
    offs[0] = 0;
    auto collect_next = [&]() -> rtod_status {
        RTOD_TRY(wait_slide(c, c->host_t));
        int32_t k = 0;
        RTOD_TRY(collect(c, slots + used, &k));
        used += k;
        offs[++done] = (int32_t)used;
        return maybe_train(c);
    };
    // prologue: slide 0's ingest; slide 0's window merge with slide 1's
    // tile sorts; slide 1's run merge; slide 2's tile sorts
    for (int32_t i = 0; i < 3 && i < n; ++i)
        stage_slide(c, src, t0, i);
    const Step s0 = step(0);
    launch_ingest_stage(c, s, s0, nullptr);
    if (n > 1) {
        const Step s1 = step(1);
        launch_merge_window(c, s, s0, &s1);
        if (runs)
            launch_batch_merges(c, s, s1, kModeSlide, c->S);
        if (n > 2)
            launch_ingest_dim(c, s, step(2), kModeSlide, nullptr, c->d.cpb,
                              c->d.cpb == 1 ? 1 : 0);
    } else {
        launch_merge_window(c, s, s0);
    }
    rtod_status st = RTOD_OK;
    int32_t launched = 0; // slides whose two stages are queued
    for (int32_t i = 0; i < n && st == RTOD_OK; ++i) {
        if (i >= kHostOuts && (st = collect_next()) != RTOD_OK)
            break;
        const int64_t t = t0 + i;
        if ((st = guard_trainer(c, t, s)) != RTOD_OK)
            break;
        decide_learn(c, t);
        const Step cur = step(i);
        // plain kernels: one launch per slide; learned: classify, then detect
        const bool fused = !cur.learn && c->d.ready;
        const Step prev = i > 0 ? step(i - 1) : Step{};
        const Step nx1 = i + 1 < n ? step(i + 1) : Step{};
        const Step nx2 = i + 2 < n ? step(i + 2) : Step{};
        const Step nx3 = i + 3 < n ? step(i + 3) : Step{};
        const StageParts parts{i > 0 ? &prev : nullptr, i + 1 < n ? &nx1 : nullptr,
                               (i + 2 < n && runs) ? &nx2 : nullptr, i + 3 < n ? &nx3 : nullptr};
        if (i + 3 < n) // its tile sorts ride in this launch
            stage_slide(c, src, t0, i + 3);
        if (timed)
            (void)hipEventRecord(c->ev[2], s);
        if (fused) {
            launch_slide_stage(c, s, cur, parts);
        } else {
            launch_classify_stage(c, s, cur);
            if (timed)
                (void)hipEventRecord(c->ev[3], s);
            launch_detect_stage(c, s, cur, parts);
        }
        if (timed)
            (void)hipEventRecord(c->ev[fused ? 3 : 4], s);
        if (i + 1 == n)
            launch_finish(c, s, cur);
        launched = i + 1;
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            set_error(std::string("launch: ") + hipGetErrorString(e));
            st = RTOD_ERR_HIP;
        } else if (timed) {
            const hipError_t q = hipStreamSynchronize(s);
            if (q != hipSuccess) {
                set_error(std::string("HIP error: ") + hipGetErrorString(q));
                st = RTOD_ERR_HIP;
            } else {
                accumulate_stage_times(c, fused);
            }
        }
    }
    if (st != RTOD_OK) {
        // the window may have been merged for a slide that was not detected:
        // the detector has to be primed again
        if (launched > 0 && launched < n)
            launch_finish(c, s, step(launched - 1));
        (void)hipStreamSynchronize(s);
        c->primed = false;
        return st;
    }
    while (st == RTOD_OK && done < n)
        st = collect_next();
    if (st != RTOD_OK) {
        (void)hipStreamSynchronize(s);
        c->primed = false;
    }
    return st;
}

rtod_status runs_slides(rtod_ctx *c, const SlideSrc &src, int32_t n, int32_t *offs,
                        int32_t *slots);

rtod_status slides_impl(rtod_ctx *c, const SlideSrc &src, int32_t n, int32_t *offs,
                        int32_t *slots)
{
    if (c->runs)
        return runs_slides(c, src, n, offs, slots);
    if (!c->lbvh && c->ride)
        return slides_two_stage(c, src, n, offs, slots);
    const int64_t t0 = c->host_t;
    const bool timed = (c->cfg.flags & RTOD_FLAG_PHASE_TIMES) != 0;
    int64_t used = 0;
    int32_t done = 0;
    offs[0] = 0;
    auto collect_next = [&]() -> rtod_status {
        RTOD_TRY(wait_slide(c, c->host_t));
        int32_t k = 0;
        RTOD_TRY(collect(c, slots + used, &k));
        used += k;
        offs[++done] = (int32_t)used;
        return maybe_train(c);
    };
    rtod_status s = RTOD_OK;
    stage_slide(c, src, t0, 0);
    launch_ingest_stage(c, c->stream, make_step_at(c, (uint64_t)t0, slide_src(c, src, t0, 0)),
                        nullptr);
    // cell-stencil index: every slide but the last leaves its epilogue to
    // the next slide's window-merge launch; pending is that slide
    Step pending{};
    bool has_pending = false;
    for (int32_t i = 0; i < n && s == RTOD_OK; ++i) {
        if (i >= kHostOuts && (s = collect_next()) != RTOD_OK)
            break;
        const uint64_t t = (uint64_t)(t0 + i);
        decide_learn(c, (int64_t)t);
        const Step st = make_step_at(c, t, slide_src(c, src, t0, i));
        const bool more = i + 1 < n;
        if (more) // its ingest rides in (or follows) this slide's launches
            stage_slide(c, src, t0, i + 1);
        const Step nx = more ? make_step_at(c, t + 1, slide_src(c, src, t0, i + 1)) : Step{};
        if ((s = guard_trainer(c, (int64_t)t, c->stream)) != RTOD_OK)
            break;
        // the next slide's ingest writes its points into the ring: it rides in
        // this slide's launches unless they read the ring (a whole-window
        // LBVH, whose leaves are ring slots), then it follows them
        const bool ride = c->ride && !c->L.whole;
        const bool fused = more && ride;
        const bool defer = more && !c->lbvh;
        launch_core_stage(c, c->stream, st, timed ? c->ev : nullptr, fused ? &nx : nullptr,
                          has_pending ? &pending : nullptr, defer);
        pending = st;
        has_pending = defer;
        if (more && !ride)
            launch_ingest_stage(c, c->stream, nx, nullptr);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            set_error(std::string("launch: ") + hipGetErrorString(e));
            s = RTOD_ERR_HIP;
        } else if (timed) {
            const hipError_t q = hipStreamSynchronize(c->stream);
            if (q != hipSuccess) {
                set_error(std::string("HIP error: ") + hipGetErrorString(q));
                s = RTOD_ERR_HIP;
            } else {
                accumulate_core_times(c, fused);
            }
        }
    }
    if (has_pending) // stopped early: the last launched slide's epilogue
        launch_finish(c, c->stream, pending);
    while (s == RTOD_OK && done < n)
        s = collect_next();
    if (s != RTOD_OK) {
        // the window may have been merged for slides that were not
        // collected: the detector has to be primed again
        (void)hipStreamSynchronize(c->stream);
        c->primed = false;
    }
    return s;
}

// ---- slide-run index (1-D): host side of rtod_runs.h ----------------------

void runs_launch(rtod_ctx *c, hipStream_t s, const RunJob &job)
{
    const int grid = job.fused ? 1 : job.nsort * c->rd.rps + job.ndet;
    const size_t lds = job.fused ? run_fused_lds_bytes(c->W)
                                 : run_lds_bytes(c->W, job.nsort > 0, job.ndet > 0);
    auto kern = job.fused ? k_run_slide
                          : (job.nsort && job.ndet ? k_run_stage
                                                   : (job.nsort ? k_run_sort : k_run_detect));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kRunNT), lds, s, c->rd, job);
}

// The initial window: batch b (slots b*S + [0, S)) sorted into run b.
rtod_status runs_prime(rtod_ctx *c, const float *pts, hipMemcpyKind kind)
{
    hipStream_t s = c->stream;
    RTOD_HIP(hipMemcpyAsync(c->d_prime, pts, sizeof(float) * (size_t)c->W, kind, s));
    RTOD_HIP(hipMemsetAsync(c->rd.err, 0, sizeof(int32_t), s));
    RunJob job{};
    job.nsort = c->m;
    job.sort_q0 = 0;
    job.src = c->d_prime; // batch k at d_prime + k*S
    runs_launch(c, s, job);
    RTOD_HIP(hipGetLastError());
    RTOD_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < c->n_hout; ++k)
        host_out_of(c, k)->slide = 0;
    c->host_t = 0;
    c->primed = true;
    c->staged = false;
    c->phase = 0;
    return RTOD_OK;
}

// Slide host_t one per call: with one run per slide a single launch (one
// workgroup sorts the run, then detects the slide), else the sort and the
// detect in two launches.  src: the new points (pinned staging or HBM).
void runs_launch_slide(rtod_ctx *c, hipStream_t s, const float *src, bool timed)
{
    const int64_t t = c->host_t;
    RunJob job{};
    job.stamps = 1;
    job.sort_q0 = c->m + t;
    job.det_i0 = t;
    job.src = src;
    if (timed)
        (void)hipEventRecord(c->ev[0], s);
    if (c->rd.rps == 1) {
        job.fused = 1;
        runs_launch(c, s, job);
        if (timed)
            (void)hipEventRecord(c->ev[1], s);
        return;
    }
    RunJob a = job, b = job;
    a.nsort = 1;
    b.ndet = 1;
    runs_launch(c, s, a);
    if (timed)
        (void)hipEventRecord(c->ev[1], s);
    runs_launch(c, s, b);
    if (timed)
        (void)hipEventRecord(c->ev[2], s);
}

void runs_accumulate_call_times(rtod_ctx *c)
{
    float ms[2] = {0.f, 0.f};
    (void)hipEventElapsedTime(&ms[0], c->ev[0], c->ev[1]);
    if (c->rd.rps == 1) {
        c->ksum[kKRunSlide] += ms[0];
        c->kcnt[kKRunSlide]++;
        return;
    }
    (void)hipEventElapsedTime(&ms[1], c->ev[1], c->ev[2]);
    c->ksum[kKRunSort] += ms[0];
    c->kcnt[kKRunSort]++;
    c->ksum[kKRunDetect] += ms[1];
    c->kcnt[kKRunDetect]++;
}

// run_slide on the slide-run index
rtod_status runs_run_slide(rtod_ctx *c, const float *src, double t_entry, double t_h2d,
                           int32_t *outs, int32_t *n_out, rtod_times *times)
{
    const bool timed = (c->cfg.flags & RTOD_FLAG_PHASE_TIMES) != 0;
    runs_launch_slide(c, c->stream, src, timed);
    RTOD_HIP(hipGetLastError());
    const rtod_status ws = wait_slide(c, c->host_t);
    if (ws != RTOD_OK) {
        (void)hipStreamSynchronize(c->stream);
        c->primed = false;
        return ws;
    }
    if (timed) {
        RTOD_HIP(hipStreamSynchronize(c->stream));
        runs_accumulate_call_times(c);
    }
    double phase[4] = {0, 0, 0, 0};
    stamp_phases(c, host_out_of(c, c->host_t), phase);
    const double t_col = now_ms();
    RTOD_TRY(collect(c, outs, n_out));
    const double t_end = now_ms();
    if (times) {
        times->copy_h2d = t_h2d - t_entry;
        times->prepare_cell = phase[0];
        times->build_index = phase[1];
        times->detect_outlier = phase[2];
        times->copy_d2h = phase[3] + (t_end - t_col);
        times->total = t_end - t_entry;
        times->launch = times->total - times->copy_h2d - times->prepare_cell -
                        times->build_index - times->detect_outlier - times->copy_d2h;
    }
    return RTOD_OK;
}

// n consecutive slides, pipelined on the context's stream, in batches (the
// first kRunFirstBatch slides, then kRunBatch each from HBM, kRunPinnedBatch
// from a rtod_host_alloc buffer, kRunHostBatch from pageable host memory):
// after a prologue launch
// sorting batch 0, launch k detects batch k and sorts batch k + 1, each
// slide's points staged into pinned memory just before the launch that sorts
// it.  At most two launches are in flight: launch k is enqueued once every
// batch before k - 1 is collected, so the staging slots, host-out buffers
// and runs that launch k rewrites are no longer read.  Results equal those
// of n rtod_slide calls.
rtod_status runs_slides(rtod_ctx *c, const SlideSrc &src, int32_t n, int32_t *offs,
                        int32_t *slots)
{
    // batch k: slides [first(k), first(k + 1)) of the run
    const int64_t B = src.host ? kRunHostBatch : src.pinned ? kRunPinnedBatch : kRunBatch;
    auto first = [n, B](int32_t k) -> int32_t {
        const int64_t f = k == 0 ? 0 : kRunFirstBatch + (int64_t)(k - 1) * B;
        return (int32_t)std::min<int64_t>(f, n);
    };
    const int64_t t0 = c->host_t;
    const bool timed = (c->cfg.flags & RTOD_FLAG_PHASE_TIMES) != 0;
    hipStream_t s = c->stream;
    int64_t used = 0;
    int32_t done = 0;
    offs[0] = 0;
    auto collect_next = [&]() -> rtod_status {
        RTOD_TRY(wait_slide(c, c->host_t));
        int32_t k = 0;
        RTOD_TRY(collect(c, slots + used, &k));
        used += k;
        offs[++done] = (int32_t)used;
        return RTOD_OK;
    };
    auto sort_job = [&](RunJob &job, int32_t first, int32_t count) {
        job.nsort = count;
        job.sort_q0 = c->m + t0 + first;
        for (int32_t k = 0; k < count; ++k)
            stage_slide(c, src, t0, first + k);
        if (src.dev) { // slide i of the run at dev + i*S (slide_src)
            job.src = src.dev;
            job.src_i0 = first;
        } else {       // staged in slot (t0 + i) mod stage_slots
            job.src = c->h_in;
            job.src_i0 = t0 + first;
            job.src_ring = c->stage_slots;
        }
    };
    rtod_status st = RTOD_OK;
    {
        RunJob pj{};
        sort_job(pj, 0, first(1));
        runs_launch(c, s, pj);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            set_error(std::string("launch: ") + hipGetErrorString(e));
            st = RTOD_ERR_HIP;
        }
    }
    for (int32_t k = 0; first(k) < n && st == RTOD_OK; ++k) {
        while (st == RTOD_OK && k >= 1 && done < first(k - 1))
            st = collect_next();
        if (st != RTOD_OK)
            break;
        const int32_t j0 = first(k), j1 = first(k + 1);
        RunJob job{};
        job.ndet = j1 - j0;
        job.det_i0 = t0 + j0;
        if (j1 < n)
            sort_job(job, j1, first(k + 2) - j1);
        if (timed)
            (void)hipEventRecord(c->ev[2], s);
        runs_launch(c, s, job);
        if (timed)
            (void)hipEventRecord(c->ev[3], s);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            set_error(std::string("launch: ") + hipGetErrorString(e));
            st = RTOD_ERR_HIP;
        } else if (timed) {
            const hipError_t q = hipStreamSynchronize(s);
            if (q != hipSuccess) {
                set_error(std::string("HIP error: ") + hipGetErrorString(q));
                st = RTOD_ERR_HIP;
            } else {
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, c->ev[2], c->ev[3]);
                // a full launch (k_run_stage), or the run's last, which
                // only detects (k_run_detect)
                const int kid = job.nsort ? kKRunStage : kKRunDetect;
                c->ksum[kid] += ms;
                c->kcnt[kid]++;
            }
        }
    }
    while (st == RTOD_OK && done < n)
        st = collect_next();
    if (st != RTOD_OK) {
        // runs of slides that were not collected may have been sorted: the
        // detector has to be primed again
        (void)hipStreamSynchronize(s);
        c->primed = false;
    }
    return st;
}

// phase API on the slide-run index: prepare_cell sorts the new run,
// detect_outlier detects the slide (build_index has nothing to do: the
// sorted runs are the index)
rtod_status runs_phase(rtod_ctx *c, bool detect)
{
    RunJob job{};
    job.stamps = 1;
    if (detect) {
        job.ndet = 1;
        job.det_i0 = c->host_t;
    } else {
        job.nsort = 1;
        job.sort_q0 = c->m + c->host_t;
        job.src = c->h_in;
        job.src_i0 = c->host_t;
        job.src_ring = c->stage_slots;
    }
    runs_launch(c, c->stream, job);
    RTOD_HIP(hipGetLastError());
    RTOD_HIP(hipStreamSynchronize(c->stream));
    return RTOD_OK;
}

} // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

RTOD_API rtod_status rtod_create(const rtod_config *cfg, rtod_ctx **out)
{
    if (!out) {
        set_error("out is NULL");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    rtod_ctx *c = nullptr;
    rtod_status s;
    try {
        s = create_impl(cfg, &c);
    } catch (...) {
        s = RTOD_ERR_INTERNAL;
    }
    if (s != RTOD_OK) {
        if (c)
            rtod_destroy(c);
        return s;
    }
    *out = c;
    return RTOD_OK;
}

RTOD_API rtod_status rtod_prime(rtod_ctx *ctx, const float *pts, int64_t n_pts)
{
    if (!ctx) {
        set_error("ctx is NULL");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    try {
        return prime_impl(ctx, pts, n_pts);
    } catch (...) {
        return RTOD_ERR_INTERNAL;
    }
}

RTOD_API rtod_status rtod_slide(rtod_ctx *ctx, const float *new_pts, int32_t *outlier_slots,
                                int32_t *n_outliers, rtod_times *times)
{
    if (!ctx) {
        set_error("ctx is NULL");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    try {
        return slide_impl(ctx, new_pts, outlier_slots, n_outliers, times);
    } catch (...) {
        return RTOD_ERR_INTERNAL;
    }
}

RTOD_API rtod_status rtod_slide_device(rtod_ctx *ctx, const float *dev_new_pts,
                                       int32_t *outlier_slots, int32_t *n_outliers,
                                       rtod_times *times)
{
    if (!ctx) {
        set_error("ctx is NULL");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    try {
        return slide_device_impl(ctx, dev_new_pts, outlier_slots, n_outliers, times);
    } catch (...) {
        return RTOD_ERR_INTERNAL;
    }
}

RTOD_API rtod_status rtod_slides_device(rtod_ctx *ctx, const float *dev_pts, int32_t n_slides,
                                        int32_t *out_offsets, int32_t *out_slots,
                                        int64_t slots_cap)
{
    if (!ctx || !dev_pts || !out_offsets || !out_slots) {
        set_error("NULL argument");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (n_slides < 1 || slots_cap < (int64_t)n_slides * ctx->list_cap) {
        set_error("n_slides must be >= 1 and out_slots hold n_slides * max(W, "
                  "launch_ray_num) entries");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (!ctx->primed || ctx->staged) {
        set_error(!ctx->primed ? "rtod_slides_device before rtod_prime"
                               : "a phase-API slide is in progress");
        return RTOD_ERR_STATE;
    }
    if (!ctx->own_stream) {
        set_error("a batch member's slides run through its batch");
        return RTOD_ERR_UNSUPPORTED;
    }
    try {
        RTOD_HIP(hipSetDevice(ctx->device));
        hipPointerAttribute_t attr{};
        if (hipPointerGetAttributes(&attr, dev_pts) != hipSuccess ||
            attr.type != hipMemoryTypeDevice) {
            (void)hipGetLastError();
            set_error("rtod_slides_device needs a device (HBM) pointer");
            return RTOD_ERR_INVALID_ARGUMENT;
        }
        SlideSrc src;
        src.dev = dev_pts;
        return slides_impl(ctx, src, n_slides, out_offsets, out_slots);
    } catch (...) {
        return RTOD_ERR_INTERNAL;
    }
}

RTOD_API rtod_status rtod_slides(rtod_ctx *ctx, const float *pts, int32_t n_slides,
                                 int32_t *out_offsets, int32_t *out_slots, int64_t slots_cap)
{
    if (!ctx || !pts || !out_offsets || !out_slots) {
        set_error("NULL argument");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (n_slides < 1 || slots_cap < (int64_t)n_slides * ctx->list_cap) {
        set_error("n_slides must be >= 1 and out_slots hold n_slides * max(W, "
                  "launch_ray_num) entries");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (!ctx->primed || ctx->staged) {
        set_error(!ctx->primed ? "rtod_slides before rtod_prime"
                               : "a phase-API slide is in progress");
        return RTOD_ERR_STATE;
    }
    if (!ctx->own_stream) {
        set_error("a batch member's slides run through its batch");
        return RTOD_ERR_UNSUPPORTED;
    }
    // slides wholly inside a buffer from rtod_host_alloc are read in place
    const float *dp = pinned_device_ptr(
        pts, (size_t)n_slides * (size_t)ctx->S * (size_t)ctx->dim * sizeof(float));
    if (!dp && ctx->stage_slots < kStageSlots) {
        set_error("slides over 64 MB are staged one at a time: use rtod_slide");
        return RTOD_ERR_UNSUPPORTED;
    }
    try {
        RTOD_HIP(hipSetDevice(ctx->device));
        SlideSrc src;
        if (dp) {
            src.dev = dp;
            src.pinned = true;
        } else {
            src.host = pts;
        }
        return slides_impl(ctx, src, n_slides, out_offsets, out_slots);
    } catch (...) {
        return RTOD_ERR_INTERNAL;
    }
}

RTOD_API rtod_status rtod_copy_new_points(rtod_ctx *c, const float *new_pts)
{
    if (!c || !new_pts) {
        set_error("NULL argument");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (!c->primed || c->phase != 0) {
        set_error("copy_new_points out of order");
        return RTOD_ERR_STATE;
    }
    const size_t nf = (size_t)c->S * c->dim;
    std::memcpy(c->h_in + (size_t)(c->host_t % c->stage_slots) * nf, new_pts, nf * sizeof(float));
    c->staged = true;
    c->phase = 1;
    return RTOD_OK;
}

RTOD_API rtod_status rtod_prepare_cell(rtod_ctx *c)
{
    if (!c) {
        set_error("ctx is NULL");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (c->phase != 1) {
        set_error("prepare_cell out of order");
        return RTOD_ERR_STATE;
    }
    RTOD_HIP(hipSetDevice(c->device));
    if (c->runs) {
        RTOD_TRY(runs_phase(c, false));
        c->phase = 2;
        return RTOD_OK;
    }
    RTOD_TRY(guard_trainer(c, c->host_t, c->stream));
    decide_learn(c, c->host_t);
    const Step st = make_step(c);
    launch_ingest_dim(c, c->stream, st, kModeSlide, nullptr, c->d.cpb, c->d.cpb == 1 ? 1 : 0);
    if (c->d.cpb > 1)
        launch_batch_merges(c, c->stream, st, kModeSlide, c->S);
    launch_merge_window(c, c->stream, st);
    launch_classify(c, c->stream, st);
    RTOD_HIP(hipGetLastError());
    RTOD_HIP(hipStreamSynchronize(c->stream));
    c->phase = 2;
    return RTOD_OK;
}

RTOD_API rtod_status rtod_build_index(rtod_ctx *c)
{
    if (!c) {
        set_error("ctx is NULL");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (c->phase != 2) {
        set_error("build_index out of order");
        return RTOD_ERR_STATE;
    }
    // cell stencil: the cell-sorted window and its sample table were built by
    // the merge in prepare_cell; LBVH: build the tree over the candidates
    if (c->lbvh) {
        RTOD_HIP(hipSetDevice(c->device));
        launch_lbvh_build(c, c->stream, make_step(c));
        RTOD_HIP(hipGetLastError());
        RTOD_HIP(hipStreamSynchronize(c->stream));
    }
    c->phase = 3;
    return RTOD_OK;
}

RTOD_API rtod_status rtod_detect_outlier(rtod_ctx *c)
{
    if (!c) {
        set_error("ctx is NULL");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (c->phase != 3) {
        set_error("detect_outlier out of order");
        return RTOD_ERR_STATE;
    }
    RTOD_HIP(hipSetDevice(c->device));
    if (c->runs) {
        RTOD_TRY(runs_phase(c, true));
        c->phase = 4;
        return RTOD_OK;
    }
    if (c->lbvh)
        launch_lbvh_detect(c, c->stream, make_step(c));
    else
        launch_detect(c, c->stream, make_step(c));
    RTOD_HIP(hipGetLastError());
    RTOD_HIP(hipStreamSynchronize(c->stream));
    c->phase = 4;
    return RTOD_OK;
}

RTOD_API rtod_status rtod_copy_outlier(rtod_ctx *c, int32_t *outs, int32_t *n_out)
{
    if (!c || !outs || !n_out) {
        set_error("NULL argument");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (c->phase != 4) {
        set_error("copy_outlier out of order");
        return RTOD_ERR_STATE;
    }
    rtod_status s = collect(c, outs, n_out);
    c->phase = 0;
    c->staged = false;
    if (s == RTOD_OK)
        s = maybe_train(c);
    return s;
}

RTOD_API rtod_status rtod_get_stats(const rtod_ctx *c, rtod_stats *o)
{
    if (!c || !o) {
        set_error("NULL argument");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    std::memset(o, 0, sizeof(*o));
    o->slides = c->host_t;
    o->candidates_last = c->cand_last;
    o->candidates_total = c->cand_total;
    o->pair_tests_last = c->tests_last;
    o->pair_tests_total = c->tests_total;
    o->hits_last = c->hits_last;
    o->hits_total = c->hits_total;
    o->outliers_last = c->out_last;
    o->outliers_total = c->out_total;
    o->device_bytes = c->dev_bytes;
    o->host_pinned_bytes = c->pinned_bytes;
    o->index_kind = c->runs ? RTOD_INDEX_SLIDE_RUNS
                            : (c->lbvh ? RTOD_INDEX_LBVH : RTOD_INDEX_CELL_STENCIL);
    o->sched_kind = c->used_pred_last ? RTOD_SCHED_LEARNED : RTOD_SCHED_PLAIN;
    o->predictor_misses = c->misses_total;
    o->predictor_points = c->pred_points;
    o->fallbacks = c->fallbacks;
    o->plain_kernel_slides = c->plain_slides;
    o->train_steps = c->train_steps;
    o->train_loss_last = c->train_loss_last;
    return RTOD_OK;
}

RTOD_API rtod_status rtod_get_predictor(rtod_ctx *c, float *params, int32_t n)
{
    if (!c || !params) {
        set_error("NULL argument");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (!c->learned) {
        set_error("the context has no predictor (sched_kind is not LEARNED)");
        return RTOD_ERR_STATE;
    }
    if (n != kMlpNP) {
        set_error("predictor has RTOD_MLP_NPARAMS parameters");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    RTOD_HIP(hipSetDevice(c->device));
    if (c->train_stream)
        RTOD_HIP(hipStreamSynchronize(c->train_stream));
    c->train_pending = false;
    RTOD_HIP(hipMemcpyAsync(params, c->d.lr.params, sizeof(float) * kMlpNP,
                            hipMemcpyDeviceToHost, c->stream));
    RTOD_HIP(hipStreamSynchronize(c->stream));
    return RTOD_OK;
}

RTOD_API rtod_status rtod_set_predictor(rtod_ctx *c, const float *params, int32_t n,
                                        int32_t frozen)
{
    if (!c || !params) {
        set_error("NULL argument");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (!c->learned) {
        set_error("the context has no predictor (sched_kind is not LEARNED)");
        return RTOD_ERR_STATE;
    }
    if (n != kMlpNP) {
        set_error("predictor has RTOD_MLP_NPARAMS parameters");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (c->staged) {
        set_error("a phase-API slide is in progress");
        return RTOD_ERR_STATE;
    }
    RTOD_HIP(hipSetDevice(c->device));
    if (c->train_stream)
        RTOD_HIP(hipStreamSynchronize(c->train_stream));
    c->train_pending = false;
    const Learn &lr = c->d.lr;
    RTOD_HIP(hipMemcpyAsync(lr.params, params, sizeof(float) * kMlpNP, hipMemcpyHostToDevice,
                            c->stream));
    for (int h = 0; h < 2; ++h)
        RTOD_HIP(hipMemcpyAsync(lr.pub + (size_t)h * kMlpNP, params, sizeof(float) * kMlpNP,
                                hipMemcpyHostToDevice, c->stream));
    RTOD_HIP(hipMemsetAsync(lr.adam_m, 0, sizeof(float) * kMlpNP, c->stream));
    RTOD_HIP(hipMemsetAsync(lr.adam_v, 0, sizeof(float) * kMlpNP, c->stream));
    const int32_t kept[2] = {0, frozen ? 1 : 0}; // adam_t, frozen
    RTOD_HIP(hipMemcpyAsync(&lr.ls->adam_t, kept, sizeof(kept), hipMemcpyHostToDevice,
                            c->stream));
    RTOD_HIP(hipStreamSynchronize(c->stream));
    return RTOD_OK;
}

namespace {
struct DevBuf {
    void *p = nullptr;
    ~DevBuf()
    {
        if (p)
            (void)hipFree(p);
    }
};
} // namespace

RTOD_API rtod_status rtod_mlp_forward(int32_t device, const float *params, const float *X,
                                      int32_t n, float *out)
{
    if (!params || (n > 0 && (!X || !out))) {
        set_error("NULL argument");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (n < 0 || n > (1 << 24)) {
        set_error("n must be in [0, 2^24]");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (n == 0)
        return RTOD_OK;
    RTOD_HIP(hipSetDevice(device));
    DevBuf bp, bx, bo;
    RTOD_HIP(hipMalloc(&bp.p, sizeof(float) * kMlpNP));
    RTOD_HIP(hipMalloc(&bx.p, sizeof(float) * kMlpIn * (size_t)n));
    RTOD_HIP(hipMalloc(&bo.p, sizeof(float) * (size_t)n));
    RTOD_HIP(hipMemcpy(bp.p, params, sizeof(float) * kMlpNP, hipMemcpyHostToDevice));
    RTOD_HIP(hipMemcpy(bx.p, X, sizeof(float) * kMlpIn * (size_t)n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mlp_forward, dim3((n + 255) / 256), dim3(256), 0, nullptr,
                       (const float *)bp.p, (const float *)bx.p, n, (float *)bo.p);
    RTOD_HIP(hipGetLastError());
    RTOD_HIP(hipMemcpy(out, bo.p, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost));
    return RTOD_OK;
}

RTOD_API rtod_status rtod_mlp_grad(int32_t device, const float *params, const float *X,
                                   const float *y, int32_t n, float *grad, float *loss)
{
    if (!params || !X || !y || !grad || !loss) {
        set_error("NULL argument");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (n < 1 || n > kTrainRows) {
        set_error("n must be in [1, 128]");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    RTOD_HIP(hipSetDevice(device));
    DevBuf bp, bx, by, bg;
    RTOD_HIP(hipMalloc(&bp.p, sizeof(float) * kMlpNP));
    RTOD_HIP(hipMalloc(&bx.p, sizeof(float) * kMlpIn * (size_t)n));
    RTOD_HIP(hipMalloc(&by.p, sizeof(float) * (size_t)n));
    RTOD_HIP(hipMalloc(&bg.p, sizeof(float) * (kMlpNP + 1)));
    RTOD_HIP(hipMemcpy(bp.p, params, sizeof(float) * kMlpNP, hipMemcpyHostToDevice));
    RTOD_HIP(hipMemcpy(bx.p, X, sizeof(float) * kMlpIn * (size_t)n, hipMemcpyHostToDevice));
    RTOD_HIP(hipMemcpy(by.p, y, sizeof(float) * (size_t)n, hipMemcpyHostToDevice));
    float *g = (float *)bg.p;
    hipLaunchKernelGGL(k_mlp_grad, dim3(1), dim3(kTrainNT), 0, nullptr, (const float *)bp.p,
                       (const float *)bx.p, (const float *)by.p, n, g, g + kMlpNP);
    RTOD_HIP(hipGetLastError());
    std::vector<float> h(kMlpNP + 1);
    RTOD_HIP(hipMemcpy(h.data(), g, sizeof(float) * (kMlpNP + 1), hipMemcpyDeviceToHost));
    std::memcpy(grad, h.data(), sizeof(float) * kMlpNP);
    *loss = h[kMlpNP];
    return RTOD_OK;
}

// ---- rtod_mlp_grad_static: the gradient in a fixed device scratch area ----
// Regions of `scratch` (floats, each starting on a 256-byte boundary): the
// parameters, kTrainRows x kMlpIn features, kTrainRows targets, the
// gradient and the loss.  k_mlp_grad reads only the first n rows and
// overwrites the whole gradient and the loss, so nothing in the area has to
// start at zero; a call leaves its inputs and results there.
namespace {
constexpr size_t kScrAlign = 64; // floats: 256 bytes
constexpr size_t scr_round(size_t n) { return (n + kScrAlign - 1) / kScrAlign * kScrAlign; }
constexpr size_t kScrParams = 0;
constexpr size_t kScrX = kScrParams + scr_round(kMlpNP);
constexpr size_t kScrY = kScrX + scr_round((size_t)kTrainRows * kMlpIn);
constexpr size_t kScrGrad = kScrY + scr_round(kTrainRows);
constexpr size_t kScrLoss = kScrGrad + scr_round(kMlpNP);
constexpr size_t kScrFloats = kScrLoss + kScrAlign; // 3648 floats, 14592 bytes
constexpr int kScrDevices = 64;

std::mutex g_scr_mu;
float *g_scr_ptr[kScrDevices] = {}; // device address of `scratch`, per device
} // namespace

static __device__ __align__(256) float scratch[kScrFloats];

namespace {
// this device's copy of `scratch` (the current device must be `device`)
rtod_status scratch_of(int device, float **out)
{
    std::lock_guard<std::mutex> lk(g_scr_mu);
    if (!g_scr_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error_hip("hipGetSymbolAddress(scratch)", e);
            return RTOD_ERR_HIP;
        }
        g_scr_ptr[device] = static_cast<float *>(p);
    }
    *out = g_scr_ptr[device];
    return RTOD_OK;
}
} // namespace

// HIP check whose error path allocates nothing either
#define RTOD_HIP_NA(call)                                                                         \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            set_error_hip(#call, e_);                                                             \
            return RTOD_ERR_HIP;                                                                  \
        }                                                                                         \
    } while (0)

RTOD_API rtod_status rtod_mlp_grad_static(int32_t device, const float *params, const float *X,
                                          const float *y, int32_t n, float *grad, float *loss)
{
    if (!params || !X || !y || !grad || !loss) {
        set_error("NULL argument");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (n < 1 || n > kTrainRows) {
        set_error("n must be in [1, 128]: the static scratch area holds 128 samples");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    int ndev = 0;
    RTOD_HIP_NA(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev || device >= kScrDevices) {
        set_error("no such device");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    RTOD_HIP_NA(hipSetDevice(device));
    float *scr = nullptr;
    RTOD_TRY(scratch_of(device, &scr));
    RTOD_HIP_NA(hipMemcpy(scr + kScrParams, params, sizeof(float) * kMlpNP, hipMemcpyHostToDevice));
    RTOD_HIP_NA(hipMemcpy(scr + kScrX, X, sizeof(float) * kMlpIn * (size_t)n,
                       hipMemcpyHostToDevice));
    RTOD_HIP_NA(hipMemcpy(scr + kScrY, y, sizeof(float) * (size_t)n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mlp_grad, dim3(1), dim3(kTrainNT), 0, nullptr,
                       (const float *)(scr + kScrParams), (const float *)(scr + kScrX),
                       (const float *)(scr + kScrY), n, scr + kScrGrad, scr + kScrLoss);
    RTOD_HIP_NA(hipGetLastError());
    RTOD_HIP_NA(hipMemcpy(grad, scr + kScrGrad, sizeof(float) * kMlpNP, hipMemcpyDeviceToHost));
    RTOD_HIP_NA(hipMemcpy(loss, scr + kScrLoss, sizeof(float), hipMemcpyDeviceToHost));
    return RTOD_OK;
}

RTOD_API rtod_status rtod_kernel_times(const rtod_ctx *c, const char **names, double *avg_ms,
                                       int64_t *launches, int32_t *n)
{
    if (!c || !n) {
        set_error("NULL argument");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    const int cap = *n;
    int k = 0;
    for (int i = 0; i < kKCount && k < cap; ++i) {
        if (names)
            names[k] = kKernelNames[i];
        if (avg_ms)
            avg_ms[k] = c->kcnt[i] ? c->ksum[i] / (double)c->kcnt[i] : 0.0;
        if (launches)
            launches[k] = c->kcnt[i];
        ++k;
    }
    *n = k;
    return RTOD_OK;
}

RTOD_API rtod_status rtod_stage_times(const rtod_ctx *c, double *out, int32_t n)
{
    if (!c || !out || n < 6) {
        set_error("rtod_stage_times needs a context and room for 6 values");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (!c->primed || c->host_t < 1) {
        set_error("no slide has run since the context was primed");
        return RTOD_ERR_STATE;
    }
    const HostOut *ho = host_out_of(c, c->host_t - 1);
    const double k = 1.0 / c->wall_khz;
    auto span = [&](int a, int b) {
        return ho->ts[b] >= ho->ts[a] ? (double)(ho->ts[b] - ho->ts[a]) * k : 0.0;
    };
    out[0] = span(kTsIngest, kTsMerge);
    out[1] = span(kTsMerge, kTsClassify);
    out[2] = span(kTsClassify, c->lbvh ? kTsIndex : kTsDetect);
    out[3] = c->lbvh ? span(kTsIndex, kTsDetect) : 0.0;
    out[4] = span(kTsDetect, kTsFinish);
    out[5] = span(kTsFinish, kTsPublished);
    return RTOD_OK;
}

RTOD_API rtod_status rtod_reset_kernel_times(rtod_ctx *c)
{
    if (!c) {
        set_error("ctx is NULL");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    for (int i = 0; i < kKCount; ++i) {
        c->ksum[i] = 0.0;
        c->kcnt[i] = 0;
    }
    return RTOD_OK;
}

RTOD_API void rtod_destroy(rtod_ctx *c)
{
    if (!c)
        return;
    (void)hipSetDevice(c->device);
    if (c->stream)
        (void)hipStreamSynchronize(c->stream);
    if (c->train_stream) {
        (void)hipStreamSynchronize(c->train_stream);
        (void)hipStreamDestroy(c->train_stream);
    }
    if (c->train_ev)
        (void)hipEventDestroy(c->train_ev);
    for (auto &e : c->ev)
        if (e)
            (void)hipEventDestroy(e);
    for (void *p : c->dev_allocs)
        (void)hipFree(p);
    if (c->h_in)
        (void)hipHostFree(c->h_in);
    if (c->h_out)
        (void)hipHostFree(c->h_out);
    if (c->stream && c->own_stream)
        (void)hipStreamDestroy(c->stream);
    delete c;
}

RTOD_API const char *rtod_status_str(rtod_status s)
{
    switch (s) {
    case RTOD_OK:
        return "ok";
    case RTOD_ERR_INVALID_ARGUMENT:
        return "invalid argument";
    case RTOD_ERR_UNSUPPORTED:
        return "unsupported";
    case RTOD_ERR_STATE:
        return "call out of order";
    case RTOD_ERR_HIP:
        return "HIP runtime error";
    case RTOD_ERR_NOMEM:
        return "out of memory";
    case RTOD_ERR_INTERNAL:
        return "internal error";
    }
    return "unknown status";
}

RTOD_API const char *rtod_last_error(void) { return g_last_error; }

RTOD_API const char *rtod_version(void) { return "rtod-mi355x 0.2 (gfx950, sources " RTOD_SRC_HASH ")"; }

RTOD_API rtod_status rtod_host_alloc(int64_t bytes, void **out)
{
    if (!out || bytes <= 0) {
        set_error("rtod_host_alloc needs bytes > 0 and an output pointer");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    void *p = nullptr;
    hipError_t e = hipHostMalloc(&p, (size_t)bytes,
                                 hipHostMallocMapped | hipHostMallocPortable |
                                     hipHostMallocCoherent);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error(std::string("hipHostMalloc: ") + hipGetErrorString(e));
        return RTOD_ERR_NOMEM;
    }
    void *dev = nullptr;
    e = hipHostGetDevicePointer(&dev, p, 0);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        (void)hipHostFree(p);
        set_error(std::string("hipHostGetDevicePointer: ") + hipGetErrorString(e));
        return RTOD_ERR_HIP;
    }
    {
        std::lock_guard<std::mutex> lk(g_pinned_mu);
        g_pinned.push_back(PinnedBuf{static_cast<const char *>(p), (size_t)bytes,
                                     static_cast<const char *>(dev)});
        g_pinned_n.fetch_add(1, std::memory_order_release);
    }
    *out = p;
    return RTOD_OK;
}

RTOD_API rtod_status rtod_host_free(void *p)
{
    if (!p)
        return RTOD_OK;
    {
        std::lock_guard<std::mutex> lk(g_pinned_mu);
        auto it = std::find_if(g_pinned.begin(), g_pinned.end(),
                               [&](const PinnedBuf &b) { return b.base == p; });
        if (it == g_pinned.end()) {
            set_error("rtod_host_free: not a buffer from rtod_host_alloc");
            return RTOD_ERR_INVALID_ARGUMENT;
        }
        g_pinned.erase(it);
        g_pinned_n.fetch_sub(1, std::memory_order_release);
    }
    RTOD_HIP(hipHostFree(p));
    return RTOD_OK;
}

} // extern "C"

// ===========================================================================
// Batched streams: many contexts advanced together, each phase one launch
// over all streams of a dimensionality (rtod_batch.h), the step one hipGraph
// replay.
// ===========================================================================
struct rtod_batch {
    int n = 0, device = 0;
    hipStream_t stream = nullptr;
    std::vector<rtod_ctx *> ctx;
    std::vector<rtod_config> cfg;
    rtod_batch_tuning tuning{};
    BatchStream *d_streams = nullptr;
    BatchItem *d_items = nullptr;
    struct Launch {
        int kind, dim, first, count;
    };
    std::vector<Launch> launches;
    int n_items = 0;
    std::vector<const float *> src; // device-resident streams, empty: host input
    std::vector<int64_t> npts;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool plan_ok = false, primed = false;
    int64_t t = 0;
    // learned batches: one predictor shared by every stream
    bool learned = false;
    bool learn_any = false; // the plan runs learned kernels for some streams
    Learn shared{};
    std::vector<void *> shared_allocs;
    hipStream_t train_stream = nullptr;
    hipEvent_t train_ev = nullptr;
    bool train_pending = false;
    int64_t train_t_next = 0; // the pending launch skips kReplaySkip steps from here
};

namespace {

void batch_free_plan(rtod_batch *b)
{
    if (b->exec)
        (void)hipGraphExecDestroy(b->exec);
    if (b->graph)
        (void)hipGraphDestroy(b->graph);
    if (b->d_items)
        (void)hipFree(b->d_items);
    if (b->d_streams)
        (void)hipFree(b->d_streams);
    b->exec = nullptr;
    b->graph = nullptr;
    b->d_items = nullptr;
    b->d_streams = nullptr;
    b->plan_ok = false;
}

void batch_free_members(rtod_batch *b)
{
    for (auto *c : b->ctx)
        rtod_destroy(c);
    b->ctx.clear();
}

// Learned batches: one launch of the shared predictor's trainer after a
// step, unless the previous one is still running.
rtod_status batch_maybe_train(rtod_batch *b)
{
    if (!b->learned || !b->learn_any) // no stream adds to the replay
        return RTOD_OK;
    if (b->train_pending) {
        const hipError_t q = hipEventQuery(b->train_ev);
        if (q == hipErrorNotReady) {
            (void)hipGetLastError(); // not an error: still running
            return RTOD_OK;
        }
        RTOD_HIP(q);
        b->train_pending = false;
    }
    hipLaunchKernelGGL(k_train_batch, dim3(1), dim3(kTrainNT), 0, b->train_stream, b->d_streams,
                       b->n, b->shared, (uint64_t)b->t, kTrainStepsBatch);
    RTOD_HIP(hipGetLastError());
    RTOD_HIP(hipEventRecord(b->train_ev, b->train_stream));
    b->train_pending = true;
    b->train_t_next = b->t;
    return RTOD_OK;
}

rtod_status batch_sync_trainer(rtod_batch *b)
{
    if (b->train_stream) {
        RTOD_HIP(hipStreamSynchronize(b->train_stream));
        b->train_pending = false;
    }
    return RTOD_OK;
}

rtod_status batch_create_members(rtod_batch *b)
{
    batch_free_members(b);
    for (int i = 0; i < b->n; ++i) {
        CreateOpts opt;
        opt.sort_shape = kBatchSortShape;
        opt.cell_scale = b->tuning.cell_scale[b->cfg[i].dim - 1];
        opt.stream = b->stream;
        opt.trainer = false;
        opt.batch_member = true;
        rtod_ctx *c = nullptr;
        rtod_status st = create_impl(&b->cfg[i], &c, opt);
        if (st != RTOD_OK) {
            if (c)
                rtod_destroy(c);
            return st;
        }
        if (b->learned) { // the batch's predictor replaces the member's own
            Learn &lr = c->d.lr;
            lr.params = b->shared.params;
            lr.pub = b->shared.pub;
            lr.adam_m = b->shared.adam_m;
            lr.adam_v = b->shared.adam_v;
            lr.ls = b->shared.ls;
        }
        b->ctx.push_back(c);
    }
    b->primed = false;
    return RTOD_OK;
}

template <int DIM> void *batch_kernel(int kind, bool learned)
{
    switch (kind) {
    case 0:
        return reinterpret_cast<void *>(&k_batch_ingest<DIM>);
    case 1:
        return reinterpret_cast<void *>(&k_batch_merge_pass<DIM>);
    case 2:
        return reinterpret_cast<void *>(&k_batch_merge_window<DIM>);
    case 3:
        return learned ? reinterpret_cast<void *>(&k_batch_classify<DIM, kClsLearned>)
                       : reinterpret_cast<void *>(&k_batch_classify<DIM, kClsUnits>);
    default:
        return learned ? reinterpret_cast<void *>(&k_batch_detect<DIM, true>)
                       : reinterpret_cast<void *>(&k_batch_detect<DIM, false>);
    }
}

// A step: per dimensionality present, a chain of launches -- ingest, merge
// pass 0..P-1 (only streams whose slide spans several sort tiles), window
// merge, classify, detect -- each over all of the group's streams, larger
// windows first.  Items of one launch are contiguous in one device table;
// the chains are independent branches of one hipGraph.
rtod_status batch_build_plan(rtod_batch *b)
{
    batch_free_plan(b);
    std::vector<BatchStream> hs(b->n);
    std::vector<int> passes(b->n, 0);
    for (int i = 0; i < b->n; ++i) {
        rtod_ctx *c = b->ctx[i];
        for (int L = kBatchCH; L < c->S; L *= 2)
            ++passes[i];
        if (passes[i] > kBatchMaxPasses) {
            set_error("batch stream needs too many merge passes (S too large)");
            return RTOD_ERR_UNSUPPORTED;
        }
        std::memset(&hs[i], 0, sizeof(BatchStream));
        hs[i].d = c->d;
        hs[i].src_base = b->src.empty() ? nullptr : b->src[i];
    }
    const int q = b->tuning.detect_points_per_wg;
    std::vector<BatchItem> items;
    b->launches.clear();
    for (int dim = 1; dim <= 3; ++dim) {
        std::vector<int> order;
        int pmax = 0;
        for (int i = 0; i < b->n; ++i)
            if (b->ctx[i]->dim == dim) {
                order.push_back(i);
                pmax = std::max(pmax, passes[i]);
            }
        if (order.empty())
            continue;
        std::stable_sort(order.begin(), order.end(),
                         [&](int x, int y) { return b->ctx[x]->W > b->ctx[y]->W; });
        auto add_launch = [&](int kind, auto blocks_of, auto arg_of, int pass) {
            const int first = (int)items.size();
            for (int i : order) {
                const int nb = blocks_of(i);
                for (int k = 0; k < nb; ++k)
                    items.push_back(BatchItem{i, k, arg_of(i), pass});
            }
            const int cnt = (int)items.size() - first;
            if (cnt > 0)
                b->launches.push_back({kind, dim, first, cnt});
        };
        add_launch(0, [&](int i) { return b->ctx[i]->d.cpb; }, [](int) { return 0; }, 0);
        for (int k = 0; k < pmax; ++k)
            add_launch(
                1, [&](int i) { return passes[i] > k ? (b->ctx[i]->S + 255) / 256 : 0; },
                [&](int) { return kBatchCH << k; }, k);
        add_launch(2, [&](int i) { return (b->ctx[i]->W + kMergeT - 1) / kMergeT; },
                   [](int) { return 0; }, 0);
        add_launch(3, [&](int i) { return (b->ctx[i]->W + kClassifyNT - 1) / kClassifyNT; },
                   [](int) { return 0; }, 0);
        auto ndb = [&](int i) {
            return std::max(1, std::min(4096, (b->ctx[i]->W + q - 1) / q));
        };
        add_launch(4, ndb, ndb, 0);
    }
    b->n_items = (int)items.size();
    RTOD_HIP(hipMalloc(&b->d_streams, sizeof(BatchStream) * hs.size()));
    RTOD_HIP(hipMalloc(&b->d_items, sizeof(BatchItem) * items.size()));
    RTOD_HIP(hipMemcpy(b->d_streams, hs.data(), sizeof(BatchStream) * hs.size(),
                       hipMemcpyHostToDevice));
    RTOD_HIP(hipMemcpy(b->d_items, items.data(), sizeof(BatchItem) * items.size(),
                       hipMemcpyHostToDevice));
    RTOD_HIP(hipGraphCreate(&b->graph, 0));
    hipGraphNode_t prev[4] = {nullptr, nullptr, nullptr, nullptr};
    b->learn_any = false;
    for (const auto &l : b->launches) {
        const BatchItem *it = b->d_items + l.first;
        const BatchStream *st = b->d_streams;
        void *args[2] = {(void *)&it, (void *)&st};
        hipKernelNodeParams kp{};
        // learned batches: the learned kernels for the dimensionalities
        // the tuning names (never 1-D, whose near stencil is the whole one)
        const bool lk = b->learned && l.dim > 1 && ((b->tuning.learned_dims >> (l.dim - 1)) & 1);
        b->learn_any = b->learn_any || lk;
        kp.func = l.dim == 1 ? batch_kernel<1>(l.kind, false)
                  : l.dim == 2 ? batch_kernel<2>(l.kind, lk)
                               : batch_kernel<3>(l.kind, lk);
        kp.gridDim = dim3(l.count);
        kp.blockDim = dim3(l.kind == 0 ? kBatchSortNT : l.kind == 4 ? kDetectNT : kBatchNT);
        kp.sharedMemBytes = 0;
        kp.kernelParams = args;
        kp.extra = nullptr;
        hipGraphNode_t node;
        RTOD_HIP(hipGraphAddKernelNode(&node, b->graph, prev[l.dim] ? &prev[l.dim] : nullptr,
                                       prev[l.dim] ? 1 : 0, &kp));
        prev[l.dim] = node;
    }
    RTOD_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->plan_ok = true;
    return RTOD_OK;
}

rtod_status batch_check_tuning(const rtod_batch_tuning *t)
{
    if (t->learned_dims & ~6) {
        set_error("tuning: learned_dims holds bits 1 (2-D) and 2 (3-D) only");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (t->detect_points_per_wg < 1 || t->detect_points_per_wg > (1 << 20)) {
        set_error("tuning: detect_points_per_wg in [1, 2^20]");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    // smaller cells keep the dense rule exact; the stencil must stay within
    // 32 rows: 2-D down to 1/4, 3-D fixed at 1
    if (!(t->cell_scale[0] >= 0.0625f && t->cell_scale[0] <= 1.f) ||
        !(t->cell_scale[1] >= 0.25f && t->cell_scale[1] <= 1.f) || t->cell_scale[2] != 1.f) {
        set_error("tuning: cell_scale 1-D in [1/16, 1], 2-D in [1/4, 1], 3-D 1");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    return RTOD_OK;
}

rtod_status batch_apply_tuning(rtod_batch *b, const rtod_batch_tuning &t)
{
    RTOD_TRY(batch_check_tuning(&t));
    RTOD_TRY(batch_sync_trainer(b));
    bool cells = false;
    for (int a = 0; a < 3; ++a)
        cells = cells || t.cell_scale[a] != b->tuning.cell_scale[a];
    b->tuning = t;
    if (cells)
        RTOD_TRY(batch_create_members(b)); // keys change: state must be primed again
    return batch_build_plan(b);
}

rtod_status batch_prime_device(rtod_batch *b)
{
    RTOD_TRY(batch_sync_trainer(b));
    for (int i = 0; i < b->n; ++i)
        RTOD_TRY(prime_impl(b->ctx[i], b->src[i], b->cfg[i].window, hipMemcpyDeviceToDevice));
    b->t = 0;
    b->primed = true;
    return RTOD_OK;
}

rtod_status batch_launch_collect(rtod_batch *b, int32_t *offsets, int32_t *slots)
{
    if (!b->plan_ok)
        RTOD_TRY(batch_build_plan(b));
    // a step past the pending training launch's skipped replay slots waits
    // for it before rewriting one
    if (b->train_pending && b->t >= b->train_t_next + kReplaySkip)
        RTOD_HIP(hipStreamWaitEvent(b->stream, b->train_ev, 0));
    RTOD_HIP(hipGraphLaunch(b->exec, b->stream));
    RTOD_HIP(hipStreamSynchronize(b->stream));
    int64_t off = 0;
    if (offsets)
        offsets[0] = 0;
    rtod_status first_err = RTOD_OK;
    for (int i = 0; i < b->n; ++i) {
        int32_t k = 0;
        rtod_status st = collect(b->ctx[i], slots ? slots + off : nullptr, &k);
        if (st != RTOD_OK && first_err == RTOD_OK)
            first_err = st;
        off += k;
        if (offsets)
            offsets[i + 1] = (int32_t)off;
    }
    b->t++;
    if (first_err == RTOD_OK)
        first_err = batch_maybe_train(b);
    return first_err;
}

rtod_status batch_step_device(rtod_batch *b, int32_t *offsets, int32_t *slots)
{
    for (int i = 0; i < b->n; ++i)
        if ((int64_t)b->cfg[i].window + (b->t + 1) * (int64_t)b->cfg[i].slide > b->npts[i]) {
            set_error("a bound device stream has no points left for this step");
            return RTOD_ERR_INVALID_ARGUMENT;
        }
    return batch_launch_collect(b, offsets, slots);
}

} // namespace

extern "C" {

RTOD_API rtod_status rtod_batch_create(const rtod_config *cfgs, int32_t n, int32_t device,
                                       rtod_batch **out)
{
    if (!out) {
        set_error("out is NULL");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    if (!cfgs || n < 1 || n > 65536) {
        set_error("need 1..65536 stream configurations");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    for (int i = 0; i < n; ++i) {
        RTOD_TRY(validate(&cfgs[i]));
        if (cfgs[i].index_kind == RTOD_INDEX_LBVH || cfgs[i].index_kind == RTOD_INDEX_SLIDE_RUNS) {
            set_error("batched streams run on the cell-stencil index");
            return RTOD_ERR_UNSUPPORTED;
        }
        if (cfgs[i].ablation != RTOD_ABL_NONE) {
            set_error("batched streams run the default path (no ablation)");
            return RTOD_ERR_UNSUPPORTED;
        }
        if (cfgs[i].sched_kind != cfgs[0].sched_kind) {
            set_error("the streams of a batch share one schedule (all PLAIN or all LEARNED)");
            return RTOD_ERR_UNSUPPORTED;
        }
    }
    if (cfgs[0].sched_kind == RTOD_SCHED_LEARNED && n > kTrainMaxStreams) {
        set_error("a learned batch holds at most 1024 streams");
        return RTOD_ERR_UNSUPPORTED;
    }
    rtod_batch *b = new (std::nothrow) rtod_batch();
    if (!b)
        return RTOD_ERR_NOMEM;
    rtod_status st = RTOD_OK;
    try {
        b->n = n;
        b->device = device;
        b->cfg.assign(cfgs, cfgs + n);
        for (auto &c : b->cfg) {
            c.device = device;
            c.flags = 0;
        }
        b->tuning.detect_points_per_wg = 64;
        b->tuning.learned_dims = 6;
        b->tuning.cell_scale[0] = b->tuning.cell_scale[1] = b->tuning.cell_scale[2] = 1.f;
        b->tuning.us_per_step = 0.0;
        b->learned = cfgs[0].sched_kind == RTOD_SCHED_LEARNED;
        st = [&]() -> rtod_status {
            RTOD_HIP(hipSetDevice(device));
            RTOD_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
            if (b->learned) {
                auto alloc = [&](void **p, size_t bytes) -> rtod_status {
                    RTOD_HIP(hipMalloc(p, bytes));
                    b->shared_allocs.push_back(*p);
                    return RTOD_OK;
                };
                Learn &sh = b->shared;
                RTOD_TRY(alloc((void **)&sh.params, sizeof(float) * kMlpNP));
                RTOD_TRY(alloc((void **)&sh.pub, sizeof(float) * 2 * kMlpNP));
                RTOD_TRY(alloc((void **)&sh.adam_m, sizeof(float) * kMlpNP));
                RTOD_TRY(alloc((void **)&sh.adam_v, sizeof(float) * kMlpNP));
                RTOD_TRY(alloc((void **)&sh.ls, sizeof(LearnShared)));
                sh.on = 1;
                const std::vector<float> p0 = initial_predictor();
                RTOD_HIP(hipMemcpy(sh.params, p0.data(), sizeof(float) * kMlpNP,
                                   hipMemcpyHostToDevice));
                for (int h = 0; h < 2; ++h)
                    RTOD_HIP(hipMemcpy(sh.pub + (size_t)h * kMlpNP, p0.data(),
                                       sizeof(float) * kMlpNP, hipMemcpyHostToDevice));
                RTOD_HIP(hipMemset(sh.adam_m, 0, sizeof(float) * kMlpNP));
                RTOD_HIP(hipMemset(sh.adam_v, 0, sizeof(float) * kMlpNP));
                RTOD_HIP(hipMemset(sh.ls, 0, sizeof(LearnShared)));
                RTOD_HIP(hipStreamCreateWithFlags(&b->train_stream, hipStreamNonBlocking));
                RTOD_HIP(hipEventCreateWithFlags(&b->train_ev, hipEventDisableTiming));
            }
            return batch_create_members(b);
        }();
    } catch (...) {
        st = RTOD_ERR_INTERNAL;
    }
    if (st != RTOD_OK) {
        rtod_batch_destroy(b);
        return st;
    }
    *out = b;
    return RTOD_OK;
}

RTOD_API rtod_status rtod_batch_prime(rtod_batch *b, const float *const *pts)
{
    if (!b || !pts) {
        set_error("NULL argument");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    for (int i = 0; i < b->n; ++i)
        if (!pts[i]) {
            set_error("NULL stream pointer");
            return RTOD_ERR_INVALID_ARGUMENT;
        }
    try {
        RTOD_HIP(hipSetDevice(b->device));
        RTOD_TRY(batch_sync_trainer(b));
        for (int i = 0; i < b->n; ++i)
            RTOD_TRY(prime_impl(b->ctx[i], pts[i], b->cfg[i].window));
        b->t = 0;
        b->primed = true;
        return RTOD_OK;
    } catch (...) {
        return RTOD_ERR_INTERNAL;
    }
}

RTOD_API rtod_status rtod_batch_step(rtod_batch *b, const float *const *new_pts,
                                     int32_t *out_offsets, int32_t *out_slots)
{
    if (!b || !new_pts || !out_offsets || !out_slots) {
        set_error("NULL argument");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (!b->primed) {
        set_error("rtod_batch_step before rtod_batch_prime");
        return RTOD_ERR_STATE;
    }
    if (!b->src.empty()) {
        set_error("the batch reads bound device streams: use rtod_batch_step_device");
        return RTOD_ERR_STATE;
    }
    for (int i = 0; i < b->n; ++i)
        if (!new_pts[i]) {
            set_error("NULL stream pointer");
            return RTOD_ERR_INVALID_ARGUMENT;
        }
    try {
        RTOD_HIP(hipSetDevice(b->device));
        for (int i = 0; i < b->n; ++i) {
            rtod_ctx *c = b->ctx[i];
            const size_t nf = (size_t)c->S * c->dim;
            std::memcpy(c->h_in + (size_t)(c->host_t & 1) * nf, new_pts[i], nf * sizeof(float));
        }
        return batch_launch_collect(b, out_offsets, out_slots);
    } catch (...) {
        return RTOD_ERR_INTERNAL;
    }
}

RTOD_API rtod_status rtod_batch_bind_device(rtod_batch *b, const float *const *dev_streams,
                                            const int64_t *n_points)
{
    if (!b) {
        set_error("NULL argument");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    try {
        RTOD_HIP(hipSetDevice(b->device));
        if (!dev_streams) { // back to host input
            b->src.clear();
            b->npts.clear();
            b->plan_ok = false;
            return RTOD_OK;
        }
        if (!n_points) {
            set_error("NULL argument");
            return RTOD_ERR_INVALID_ARGUMENT;
        }
        for (int i = 0; i < b->n; ++i) {
            hipPointerAttribute_t attr{};
            if (!dev_streams[i] || hipPointerGetAttributes(&attr, dev_streams[i]) != hipSuccess ||
                attr.type != hipMemoryTypeDevice) {
                (void)hipGetLastError();
                set_error("rtod_batch_bind_device needs device (HBM) pointers");
                return RTOD_ERR_INVALID_ARGUMENT;
            }
            if (n_points[i] < b->cfg[i].window) {
                set_error("a bound stream holds fewer than W points");
                return RTOD_ERR_INVALID_ARGUMENT;
            }
        }
        b->src.assign(dev_streams, dev_streams + b->n);
        b->npts.assign(n_points, n_points + b->n);
        b->plan_ok = false;
        return RTOD_OK;
    } catch (...) {
        return RTOD_ERR_INTERNAL;
    }
}

RTOD_API rtod_status rtod_batch_prime_device(rtod_batch *b)
{
    if (!b) {
        set_error("NULL argument");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (b->src.empty()) {
        set_error("no device streams bound");
        return RTOD_ERR_STATE;
    }
    try {
        RTOD_HIP(hipSetDevice(b->device));
        return batch_prime_device(b);
    } catch (...) {
        return RTOD_ERR_INTERNAL;
    }
}

RTOD_API rtod_status rtod_batch_step_device(rtod_batch *b, int32_t *out_offsets,
                                            int32_t *out_slots)
{
    if (!b || !out_offsets || !out_slots) {
        set_error("NULL argument");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (b->src.empty()) {
        set_error("no device streams bound");
        return RTOD_ERR_STATE;
    }
    if (!b->primed) {
        set_error("rtod_batch_step_device before priming");
        return RTOD_ERR_STATE;
    }
    try {
        RTOD_HIP(hipSetDevice(b->device));
        return batch_step_device(b, out_offsets, out_slots);
    } catch (...) {
        return RTOD_ERR_INTERNAL;
    }
}

RTOD_API rtod_status rtod_batch_get_tuning(const rtod_batch *b, rtod_batch_tuning *out)
{
    if (!b || !out) {
        set_error("NULL argument");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    *out = b->tuning;
    return RTOD_OK;
}

RTOD_API rtod_status rtod_batch_set_tuning(rtod_batch *b, const rtod_batch_tuning *t)
{
    if (!b || !t) {
        set_error("NULL argument");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    try {
        RTOD_HIP(hipSetDevice(b->device));
        RTOD_TRY(batch_check_tuning(t));
        RTOD_TRY(batch_apply_tuning(b, *t));
        b->primed = false; // every stream restarts from rtod_batch_prime*
        return RTOD_OK;
    } catch (...) {
        return RTOD_ERR_INTERNAL;
    }
}

RTOD_API rtod_status rtod_batch_autotune(rtod_batch *b, int32_t slides, rtod_batch_tuning *out)
{
    if (!b) {
        set_error("NULL argument");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    if (b->src.empty()) {
        set_error("autotuning replays the bound device streams: bind them first");
        return RTOD_ERR_STATE;
    }
    if (slides < 2 || slides > 1000) {
        set_error("slides must be in [2, 1000]");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    for (int i = 0; i < b->n; ++i)
        if ((int64_t)b->cfg[i].window + (int64_t)slides * b->cfg[i].slide > b->npts[i]) {
            set_error("a bound stream is shorter than W + slides*S");
            return RTOD_ERR_INVALID_ARGUMENT;
        }
    try {
        RTOD_HIP(hipSetDevice(b->device));
        const int qs[4] = {16, 64, 256, 1024};
        const float c1[3] = {1.f, 0.5f, 0.25f};
        rtod_batch_tuning best{};
        double best_us = 1e300;
        std::vector<int32_t> offs(b->n + 1);
        int64_t cap = 0;
        for (auto &c : b->cfg)
            cap += c.window;
        std::vector<int32_t> slots((size_t)cap);
        auto measure = [&](const rtod_batch_tuning &t, double *us) -> rtod_status {
            RTOD_TRY(batch_apply_tuning(b, t));
            RTOD_TRY(batch_prime_device(b));
            double sum = 0.0;
            int cnt = 0;
            for (int k = 0; k < slides; ++k) {
                const double t0 = now_ms();
                RTOD_TRY(batch_step_device(b, offs.data(), slots.data()));
                if (k > 0) { // the first step warms caches and code
                    sum += now_ms() - t0;
                    ++cnt;
                }
            }
            *us = sum / cnt * 1e3;
            return RTOD_OK;
        };
        for (float cs : c1)
            for (int q : qs) {
                rtod_batch_tuning t{};
                t.detect_points_per_wg = q;
                t.learned_dims = b->tuning.learned_dims;
                t.cell_scale[0] = cs;
                t.cell_scale[1] = cs;
                t.cell_scale[2] = 1.f;
                double us = 0.0;
                RTOD_TRY(measure(t, &us));
                if (us < best_us) {
                    best_us = us;
                    best = t;
                }
            }
        // learned batches: then which dimensionalities run the learned
        // kernels (the others the plain ones), at the best shape
        if (b->learned) {
            int present = 0;
            for (auto &c : b->cfg)
                present |= c.dim > 1 ? 1 << (c.dim - 1) : 0;
            const int first = best.learned_dims & present;
            for (int ld : {6, 4, 2, 0}) {
                if ((ld & present) != ld || ld == first)
                    continue;
                rtod_batch_tuning t = best;
                t.learned_dims = ld;
                double us = 0.0;
                RTOD_TRY(measure(t, &us));
                if (us < best_us) {
                    best_us = us;
                    best = t;
                }
            }
        }
        best.us_per_step = best_us;
        RTOD_TRY(batch_apply_tuning(b, best));
        RTOD_TRY(batch_prime_device(b));
        if (out)
            *out = best;
        return RTOD_OK;
    } catch (...) {
        return RTOD_ERR_INTERNAL;
    }
}

RTOD_API rtod_status rtod_batch_stream_stats(const rtod_batch *b, int32_t i, rtod_stats *out)
{
    if (!b || !out || i < 0 || i >= b->n) {
        set_error("NULL argument or stream index out of range");
        return RTOD_ERR_INVALID_ARGUMENT;
    }
    return rtod_get_stats(b->ctx[i], out);
}

RTOD_API void rtod_batch_destroy(rtod_batch *b)
{
    if (!b)
        return;
    (void)hipSetDevice(b->device);
    if (b->stream)
        (void)hipStreamSynchronize(b->stream);
    if (b->train_stream) {
        (void)hipStreamSynchronize(b->train_stream);
        (void)hipStreamDestroy(b->train_stream);
    }
    if (b->train_ev)
        (void)hipEventDestroy(b->train_ev);
    batch_free_plan(b);
    batch_free_members(b);
    for (void *p : b->shared_allocs)
        (void)hipFree(p);
    if (b->stream)
        (void)hipStreamDestroy(b->stream);
    delete b;
}

} // extern "C"
