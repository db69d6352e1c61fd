// rtod_state.h -- device-resident state of one detector context.
//
// HBM layout (W window, S slide, d dims, m = W/S):
//   ring      f32 [d][W]        window coordinates by ring slot (SoA), updated
//                               in place each slide (reference: params.points,
//                               optixScan.cpp:232-238, :1053-1057)
//   akey/aslot/axyz [2]         the window sorted by (cell key, arrival): the
//                               cell-stencil index; ping-pong by slide parity
//   apos [2][W]                 position of each sorted-window point inside
//                               its own sorted batch
//   bkey/bslot [(m+3)*S]        each batch of S arrivals sorted by cell key,
//                               kept until it expires m slides later (three
//                               extra entries: later slides' batches are
//                               sorted while this slide's merge reads two)
//   bxyz f32 [d][(m+3)*S]       coordinates of each sorted batch (SoA)
//   bsamp [(m+3)*nbs]           every 2^bs_log2-th key of each sorted batch
//   tkey/tslot/txyz [2]         scratch runs for chunked sorts (max(W,2S):
//                               two slides' tiles at (t & 1) * S)
//   samples  u64 [2][W/stride]  every stride-th key of each sorted window
//   samples4 u64 [2][W/4]       every 4th key (when the stride exceeds 4)
//   cand     i32 [W]            candidate positions (sorted-window index; the
//                               LBVH index's per-point list)
//   cmask    u64 [W/64]         LBVH index: candidate bits of each 64 sorted
//                               positions
//   units    i32x2 [W]          cell units (start, len <= 4 in 1-D, 8 else): runs of one
//                               sparse cell's candidates in the sorted window
//   bitmap   u32 [3][W/32]      outlier flags by ring slot, one set per
//                               slide modulo 3
//   ready    i32 [3][64 x 1024] one launch per slide: classify-done flags of
//                               slide t at set t % 3, 4 KB apart
//   ctl                         slide counters, tickets, pair tests, error
//                               flags, classify counters (three sets)
// Learned scheduler (only when configured):
//   params/adam_m/adam_v f32 [1313]   predictor MLP, FP32 master + Adam moments
//   pub f32 [2][1313]           published weights (classify reads pub[lc.pub&1])
//   rep_x f32 [20][W][8]        replay: features of each cell unit, by slide
//                               (slot t mod 20) and unit-list index
//   rep_c i32 [20][W]           replay: the unit's smallest exact neighbour
//                               count, capped at K+1
//   rep_n i32 [20][2]           replay: bucket sizes of the stored slide
// Pinned host memory, device-mapped:
//   h_in  f32 [8][S*d]          new slide staging (row-major, as given), slide
//                               t in slot t % 8 (a batch's streams, and slides
//                               over 64 MB: [2], t & 1)
//   h_out [4]                   outlier count + ascending slot list +
//                               counters + stamp; slide t uses t mod 4
#pragma once

#include <stdint.h>

namespace rtod {

// A slide's classify counters.  There are three sets, by slide modulo 3
// (kSlideSets): while slide t is classified and detected, the epilogue of
// slide t - 1 may still read its set, and the window merge of slide t + 1
// (resetting the next set) may run beside both.
// Each counter sits on a cache line of its own: every classify workgroup
// adds to each with one device-scope atomic, and atomics on one line are
// performed one after another.
struct SlideCnt {
    alignas(128) int32_t n_cand; // candidates
    alignas(128) int32_t n_units; // cell units (learned: the probable-outlier bucket)
    alignas(128) int32_t n_units_i; // learned: the probable-inlier bucket
    alignas(128) int32_t n_done; // one launch per slide: classify workgroups done
    alignas(128) unsigned long long pair_tests; // predicate evaluations of the slide
    alignas(128) unsigned long long hits; // sum over candidates of min(count, K + 1)
    alignas(128) int32_t misses; // learned: mispredicted units of the slide
    // LBVH index: per-axis minimum key field over the window (classify), the
    // origin of the slide's Morton codes (0 before any classify: raw fields)
    alignas(128) unsigned long long fmin[3];
};

struct Ctl {
    uint64_t t;            // slides completed since prime
    uint64_t t_det;        // the slide last classified: a batch's detect takes its
                           // slide from here, since t advances while it runs
    uint32_t ticket;       // last-workgroup ticket of the detect kernel
    uint32_t fin_ticket;   // parts done of a deferred slide epilogue
    int32_t err;
    int32_t pad;
    SlideCnt cnt[3];       // slide t's at cnt[t % 3]; reset by its window merge
};

struct HostOut {
    int32_t n_out;
    int32_t n_cand;
    int64_t slide;         // t + 1 after the slide's detect kernel
    unsigned long long pair_tests;
    int32_t err;
    int32_t misses;        // learned: mispredicted candidates of the slide
    int32_t used_pred;     // learned: the slide was scheduled by prediction
    int32_t train_steps;   // learned: optimiser steps since creation
    float train_loss;      // learned: mean loss of those steps
    int32_t n_units;       // cell units of the slide
    int32_t learn_kernels; // learned: the slide ran the learned kernels
    unsigned long long hits; // sum over the slide's candidates of min(count, K + 1)
    uint64_t ts[8];        // one slide per call: device stamps (kTs*)
    int32_t list[1];       // W entries follow
};

// Device wall-clock stamps of a slide run one per call, in ticks of the
// device's constant wall clock: the start of its ingest, window merge,
// classify, LBVH build (LBVH index only), detect and epilogue kernels, and
// the slide being published.
enum { kTsIngest = 0, kTsMerge, kTsClassify, kTsIndex, kTsDetect, kTsFinish, kTsPublished,
       kTsCount };

// Device-side state of the learned scheduler: per stream ...
// The fallback decision for slide t is made by the epilogue of slide t - 2
// (from that slide's miss rate) and kept in dec[t % kDecSlots]: that
// epilogue is complete before slide t's detect starts both when slides run
// one per call and when they are pipelined (where the epilogue of slide
// t - 1 runs beside the detect of slide t), so the schedule of every slide
// is the same either way.
constexpr int kDecSlots = 4;
struct LearnCtl {
    int32_t dec[kDecSlots]; // 1: slide t is scheduled by prediction (dec[t % 4])
};

// ... and of the predictor, which several streams may share (a batch).
struct LearnShared {
    float loss_last;       // mean loss of the last training launch
    int32_t adam_t;        // optimiser steps since the weights were set
    int32_t frozen;        // 1: no training
    int32_t steps_total;   // optimiser steps since creation
    int32_t pub;           // the current half of the published weights: pub & 1
    int32_t pad[3];
};

struct Learn {
    int on;                // learned schedule configured
    float miss_thr;        // fall back to the plain schedule above this rate
    float *params;         // [kMlpNP] FP32 master weights (the trainer's own)
    float *pub;            // [2][kMlpNP] published weights, double-buffered
    float *adam_m, *adam_v;
    float *rep_x;          // [kReplaySlides][W][kFeatStride] features by unit
    int32_t *rep_c;        // [kReplaySlides][W] the unit's minimum count, capped
    int32_t *rep_n;        // [kReplaySlides][2] bucket sizes
    LearnCtl *lc;          // this stream's
    LearnShared *ls;       // the predictor's
};

// Per-slide quantities, passed by value to every kernel of the slide.
struct Step {
    uint64_t t; // slides completed before this one
    int u;      // ring batch index t mod m: slots [u*S, u*S+S) expire and refill
    int par;    // sorted-window parity: read [par], write [par^1]
    int eold;   // batch-ring entry of the expiring batch (generation t)
    int enew;   // batch-ring entry of the new batch (generation m+t)
    int buf;    // pinned staging buffer holding the new slide
    int learn;  // learned schedule: this slide runs the learned classify and
                // detect (else the plain ones)
    const float *src; // the new slide ([S][d] row-major): pinned staging or HBM
};

struct Dev {
    int dim, W, S, m, K;
    int nent;         // batch-ring entries, m + 3: the batches of the next three
                      // slides can be ingested while this slide's merge reads its two
    int unit_max;     // candidates per cell unit (a power of two, <= kUnitCap)
    int tile_par;     // slide batches sorted as several tiles go to the tile
                      // scratch at offset (t & 1) * S (two slides' tiles live)
    int cpb;          // sort chunks per batch
    int run_len;      // points per sorted run k_merge_runs merges (the tile)
    int nsamp;        // samples in the sorted-window sample table
    int stride_log2;  // sample stride
    int nbs;          // samples per sorted batch
    int bs_log2;      // batch sample stride
    float r2;
    double inv_c;     // 1 / cell side
    double Rp;        // R (1 + 1e-5): stencil half-width
    float *ring;
    uint64_t *akey[2];
    int32_t *aslot[2];
    float *axyz[2];
    int32_t *apos[2];
    uint64_t *bkey;
    int32_t *bslot;
    uint64_t *bsamp;
    float *bxyz;
    uint64_t *tkey[2];
    int32_t *tslot[2];
    float *txyz[2];
    int tmpn;         // plane stride of txyz
    uint64_t *samples[2];
    uint64_t *samples4[2]; // every 4th key of the sorted window (when the
                           // sample stride exceeds 4; else null)
    int nsamp4;            // entries of samples4 (0: none)
    int32_t *cand;
    uint64_t *cmask;    // LBVH index: candidate bits of each 64 sorted-window
                        // points from a multiple of 64 (classify), or nullptr
    int2 *units;        // cell units: {start, len}
    int32_t *cand_cnt;  // per-unit minimum count (capped)
    uint32_t *bitmap;   // [3][bm_words]: outlier bits of slide t at (t % 3)
    int bm_words;       // ceil(W / 32), rounded up to 64
    int fin_parts;      // parts of a deferred slide epilogue
    int32_t *ready;     // [3][kReadyFlags * kReadyStride]: classify-done flags
    int wait_spins;     // bound of a detect's wait for its classify (one launch per slide)
    int no_filter;      // ablation "no grid filter": every point is a candidate
    int ray_lim;        // no grid filter with launch_ray_num < W: only ring slots
                        // below it are queried (0: all)
    uint64_t *tstamp;   // [kTsCount] device stamps of the slide run one per call
    Ctl *ctl;
    const float *h_in;
    HostOut *h_out;      // kHostOuts host-out buffers, slide t uses t % kHostOuts
    int64_t h_out_bytes; // their stride
    Learn lr;
};

// Error flags a slide's kernels raise in ctl->err (one bit each); the slide
// epilogue reports them in HostOut::err and the host maps each to a message.
constexpr int kErrMergeSurvivor = 1; // window merge: a survivor's rank out of range
constexpr int kErrMergeNew = 2;      // window merge: a new point's rank out of range
constexpr int kErrStencilRows = 4;   // detect: a unit's stencil wider than 32 rows
constexpr int kErrClassifyWait = 8;  // detect: the bounded wait for its classify ran out
constexpr int kErrLbvhWalk = 16;     // LBVH: a walk exceeded its step bound

// the host-out buffer of slide t: consecutive slides rotate through
// kHostOuts buffers, so up to that many slides can be in flight while the
// host reads the oldest one's list
constexpr int kHostOuts = 4;
// Pinned staging slots of a single stream.  A pipelined run from host
// memory stages slide i + 3 just before the launch carrying its ingest (one
// launch per slide); once slide i - kHostOuts is collected, every launch
// reading a slide before i has completed, so with eight slots the slot
// rewritten (that of slide i - 5) is no longer read; before the first
// collection the slots rewritten were last read by an earlier, completed
// call.
constexpr int kStageSlots = 8;
constexpr size_t kStageSlotCap = (size_t)64 << 20; // bytes per slot beyond which
                                                   // a stream keeps two slots

// the classify counters of slide t
constexpr int kSlideSets = 3;
__device__ __forceinline__ SlideCnt *cnt_of(const Dev &d, uint64_t t)
{
    return &d.ctl->cnt[t % kSlideSets];
}

// the outlier bitmap of slide t (kSlideSets of them: the epilogue of slide t
// may still read its bits while slide t + 1 is detected and the bitmap of
// slide t + 2 is cleared by its window merge)
__device__ __forceinline__ uint32_t *bitmap_of(const Dev &d, uint64_t t)
{
    return d.bitmap + (size_t)(t % kSlideSets) * (size_t)d.bm_words;
}
// the classify-done flags of slide t: kReadyFlags copies, kReadyStride
// words (4 KB) apart, so that the detect workgroups waiting on them spread
// their polling over the memory channels (polling one address saturates its
// channel and stalls every other access that maps to it)
constexpr int kReadyFlags = 64, kReadyStride = 1024;
__device__ __forceinline__ int32_t *ready_of(const Dev &d, uint64_t t, int j)
{
    return d.ready + ((size_t)(t % kSlideSets) * kReadyFlags + j) * kReadyStride;
}
__device__ __forceinline__ HostOut *host_out(const Dev &d, uint64_t t)
{
    return reinterpret_cast<HostOut *>(reinterpret_cast<char *>(d.h_out) +
                                       (size_t)(t % kHostOuts) * (size_t)d.h_out_bytes);
}

} // namespace rtod
