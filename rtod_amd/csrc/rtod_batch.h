// rtod_batch.h -- many independent window streams (each with its own W, S,
// R, K and dim) advanced by one slide per step.
//
// Streams are grouped by dimensionality.  For each group a step is a chain
// of launches -- ingest, the merge passes of slides larger than one sort
// tile, window merge, classify, detect -- each covering that phase of ALL
// the group's streams: workgroup b of a launch looks up its (stream,
// workgroup-of-phase) in a table built once per tuning, so a 2k-point and a
// 100k-point stream share a launch in proportion to their work.  The (up to
// three) chains are independent branches of one hipGraph, replayed once per
// step.  The bodies are the single-stream kernels' (rtod_plain.h); a
// stream's Dev comes from a device array and the slide's Step from the
// stream's own device slide counter, so no launch takes per-step arguments.
#pragma once

#include "rtod_plain.h"

namespace rtod {

constexpr int kBatchNT = 256;
constexpr int kBatchSortNT = 1024, kBatchSortIT = 4; // ingest tile: 4096 points
constexpr int kBatchCH = kBatchSortNT * kBatchSortIT;
constexpr int kBatchMaxPasses = 12;

struct BatchItem {
    int32_t stream;
    int32_t bid;  // workgroup within the stream's phase
    int32_t arg;  // merge pass: run length L; detect: the phase's workgroups
    int32_t pass; // merge pass index
};

struct BatchStream {
    Dev d;
    const float *src_base; // device-resident stream (slide t at row W + t*S), or
                           // nullptr: the pinned staging buffers
};

// det: for the detect phase, whose last workgroup advances ctl->t while
// other workgroups of the stream may still be starting (they take the slide
// that classify recorded)
__device__ __forceinline__ Step batch_step(const BatchStream &bs, bool det = false)
{
    const Dev &d = bs.d;
    const uint64_t t = det ? d.ctl->t_det : d.ctl->t;
    const uint64_t m = (uint64_t)d.m;
    Step st;
    st.t = t;
    st.u = (int)(t % m);
    st.par = (int)(t & 1);
    const uint64_t ne = (uint64_t)d.nent;
    st.eold = (int)(t % ne);
    st.enew = (int)((m + t) % ne);
    st.buf = (int)(t & 1);
    st.learn = 0; // a batch's kernels are chosen per dimensionality at plan time
    st.src = bs.src_base ? bs.src_base + ((size_t)d.W + (size_t)t * d.S) * d.dim
                         : d.h_in + (size_t)st.buf * d.S * d.dim;
    return st;
}

// Kernels of one phase for the streams of one dimensionality.
template <int DIM>
__global__ __launch_bounds__(kBatchSortNT) void k_batch_ingest(const BatchItem *items,
                                                              const BatchStream *streams)
{
    __shared__ IngestLds<DIM, kBatchSortNT, kBatchSortIT> lds;
    const BatchItem it = items[blockIdx.x];
    const BatchStream &bs = streams[it.stream];
    const Dev &d = bs.d;
    const Step st = batch_step(bs);
    ingest_sort_body<DIM, kBatchSortNT, kBatchSortIT>(d, st, kModeSlide, nullptr,
                                                      d.cpb == 1 ? 1 : 0, it.bid, lds);
}

// merge pass `pass` of the new batches' sorted chunks: reads scratch run
// pass&1, writes (pass+1)&1, the last pass writes the batch ring entry
template <int DIM>
__global__ __launch_bounds__(kBatchNT) void k_batch_merge_pass(const BatchItem *items,
                                                              const BatchStream *streams)
{
    const BatchItem it = items[blockIdx.x];
    const BatchStream &bs = streams[it.stream];
    const Dev &d = bs.d;
    const Step st = batch_step(bs);
    const int L = it.arg, src = it.pass & 1;
    const bool last = 2 * L >= d.S;
    const size_t bxs = (size_t)d.nent * d.S;
    merge_pass_body<DIM>(d, st, d.tkey[src], d.tslot[src], d.txyz[src], (size_t)d.tmpn, nullptr,
                         last ? d.bkey : d.tkey[src ^ 1], last ? d.bslot : d.tslot[src ^ 1],
                         last ? d.bxyz : d.txyz[src ^ 1], last ? bxs : (size_t)d.tmpn, nullptr,
                         d.S, d.S, L, last ? 1 : 0, (int)kModeSlide, it.bid);
}

template <int DIM>
__global__ __launch_bounds__(kBatchNT) void k_batch_merge_window(const BatchItem *items,
                                                                const BatchStream *streams)
{
    __shared__ MergeWindowLds<256> lds;
    const BatchItem it = items[blockIdx.x];
    const BatchStream &bs = streams[it.stream];
    merge_window_body<DIM, 256>(bs.d, batch_step(bs), it.bid, lds);
}

template <int DIM, int MODE>
__global__ __launch_bounds__(kBatchNT) void k_batch_classify(const BatchItem *items,
                                                            const BatchStream *streams)
{
    __shared__ ClassifyLds<MODE == kClsLearned> lds;
    const BatchItem it = items[blockIdx.x];
    const BatchStream &bs = streams[it.stream];
    classify_body<DIM, MODE>(bs.d, batch_step(bs), it.bid, lds);
}

template <int DIM, bool LEARN>
__global__ __launch_bounds__(kDetectNT) RTOD_DETECT_OCCUPANCY void k_batch_detect(const BatchItem *items,
                                                          const BatchStream *streams)
{
    __shared__ DetectLds lds;
    const BatchItem it = items[blockIdx.x];
    const BatchStream &bs = streams[it.stream];
    detect_body<DIM, LEARN>(bs.d, batch_step(bs, true), it.bid, it.arg, true, lds);
}

// The batch's shared predictor (learned batches), trained on every stream's
// replay: a stream is drawn with probability proportional to its stored
// units, then one of them uniformly; the slots of the steps from t_next on
// are skipped (replay_skipped).  n <= kTrainMaxStreams.
__global__ __launch_bounds__(kTrainNT) void k_train_batch(const BatchStream *streams, int n,
                                                         Learn shared, uint64_t t_next,
                                                         int steps)
{
    constexpr int PER = kTrainMaxStreams / kTrainNT;
    __shared__ TrainWaveLds lds_wave[kTrainNT / kWave];
    __shared__ TrainLds s;
    __shared__ int s_pref[kTrainMaxStreams + 1];
    __shared__ int s_scan[kTrainNT / kWave];
    if (shared.ls->frozen)
        return;
    const int tid = threadIdx.x;
    int tot[PER], mine = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int i = tid * PER + q;
        tot[q] = 0;
        if (i < n) {
            const int32_t *rn = streams[i].d.lr.rep_n;
            for (int j = 0; j < kReplaySlides; ++j)
                if (!replay_skipped(j, t_next))
                    tot[q] += rn[2 * j] + rn[2 * j + 1];
        }
        mine += tot[q];
    }
    int total = 0;
    int run = block_excl_scan<kTrainNT>(mine, s_scan, total);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int i = tid * PER + q;
        if (i < n)
            s_pref[i] = run;
        run += tot[q];
    }
    if (tid == 0)
        s_pref[n] = total;
    __syncthreads();
    if (total <= 0)
        return;
    train_core(shared, steps, lds_wave, s, [&](uint32_t h, float *xs, float &y) {
        const int u = (int)(h % (uint32_t)total);
        int si = 0, hi = n - 1; // last stream with s_pref[si] <= u
        while (si < hi) {
            const int mid = (si + hi + 1) >> 1;
            if (s_pref[mid] <= u)
                si = mid;
            else
                hi = mid - 1;
        }
        const Dev &d = streams[si].d;
        int off = u - s_pref[si];
        int4 rn4[kReplaySlides / 2];
        const int4 *src = reinterpret_cast<const int4 *>(d.lr.rep_n);
#pragma unroll
        for (int k = 0; k < kReplaySlides / 2; ++k)
            rn4[k] = src[k];
        int slot = 0, n_o = 0;
#pragma unroll
        for (int k = 0; k < kReplaySlides / 2; ++k) {
#pragma unroll
            for (int h2 = 0; h2 < 2; ++h2) {
                const int j = 2 * k + h2;
                const int o = h2 ? rn4[k].z : rn4[k].x, in = h2 ? rn4[k].w : rn4[k].y;
                const int c = replay_skipped(j, t_next) ? 0 : o + in;
                if (off >= 0 && off < c) {
                    slot = j;
                    n_o = o;
                    off = -1 - off; // found: remember the in-slot index
                } else if (off >= 0) {
                    off -= c;
                }
            }
        }
        if (off >= 0) { // counts changed underneath: an empty sample
#pragma unroll
            for (int j = 0; j < kFeatStride; ++j)
                xs[j] = 0.f;
            y = 0.f;
            return;
        }
        const int in_slot = -1 - off;
        load_replay_row(d.lr, d.W, d.K, slot,
                        in_slot < n_o ? in_slot : d.W - 1 - (in_slot - n_o), xs, y);
    });
}

} // namespace rtod
