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

__device__ __forceinline__ Step batch_step(const BatchStream &bs)
{
    const Dev &d = bs.d;
    const uint64_t t = d.ctl->t;
    const uint64_t m = (uint64_t)d.m;
    Step st;
    st.t = t;
    st.u = (int)(t % m);
    st.par = (int)(t & 1);
    st.eold = (int)(t % (m + 1));
    st.enew = (int)((m + t) % (m + 1));
    st.buf = (int)(t & 1);
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
    const size_t bxs = (size_t)(d.m + 1) * d.S;
    merge_pass_body<DIM>(d, st, d.tkey[src], d.tslot[src], d.txyz[src], (size_t)d.tmpn, nullptr,
                         last ? d.bkey : d.tkey[src ^ 1], last ? d.bslot : d.tslot[src ^ 1],
                         last ? d.bxyz : d.txyz[src ^ 1], last ? bxs : (size_t)d.tmpn, nullptr,
                         d.S, d.S, L, last ? 1 : 0, (int)kModeSlide, it.bid);
}

template <int DIM>
__global__ __launch_bounds__(kBatchNT) void k_batch_merge_window(const BatchItem *items,
                                                                const BatchStream *streams)
{
    __shared__ MergeWindowLds lds;
    const BatchItem it = items[blockIdx.x];
    const BatchStream &bs = streams[it.stream];
    merge_window_body<DIM>(bs.d, batch_step(bs), it.bid, lds);
}

template <int DIM>
__global__ __launch_bounds__(kBatchNT) void k_batch_classify(const BatchItem *items,
                                                            const BatchStream *streams)
{
    __shared__ ClassifyLds<false> lds;
    const BatchItem it = items[blockIdx.x];
    const BatchStream &bs = streams[it.stream];
    classify_body<DIM, kClsUnits>(bs.d, batch_step(bs), it.bid, lds);
}

template <int DIM>
__global__ __launch_bounds__(kBatchNT) void k_batch_detect(const BatchItem *items,
                                                          const BatchStream *streams)
{
    __shared__ DetectLds lds;
    const BatchItem it = items[blockIdx.x];
    const BatchStream &bs = streams[it.stream];
    detect_body<DIM, false>(bs.d, batch_step(bs), it.bid, it.arg, lds);
}

} // namespace rtod
