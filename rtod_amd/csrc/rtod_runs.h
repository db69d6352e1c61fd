// rtod_runs.h -- the slide-run index of 1-D streams.
//
// Each slide's S points are bucketed once, on arrival, into a "run": a
// counting sort by sub-cell key (the cell of side c = R (1 - 1e-5) split
// into 2^s sub-cells, s chosen per run so the run's key range fits
// kRunBMax buckets).  The window is the m most recent runs; nothing is
// merged, so a slide costs O(S) work plus the range counts of its
// candidates, where the cell-sorted window of rtod_plain.h rewrites all W
// points every slide.
//
// A run holds, in HBM: its points in bucket order (coordinates and ring
// slots), the start of every bucket (an exclusive prefix over the run's
// buckets), and its cell table (each cell with points, ascending, and the
// index of its first point).  A run whose cell range exceeds kRunBMax (data
// far apart, or clamped/non-finite values) is sorted by cell with the LDS
// radix sort instead and searched by cell ("search" mode, s = 0).
//
// A slide's detect (one workgroup):
//   1. the window's cell counts: the runs' cell tables added into an LDS
//      histogram over the window's cell range (or, when the range exceeds
//      kRunHCap or a run is in search mode, summed per cell by searching
//      every run's cell table);
//   2. candidates: the points of every cell with at most K points in the
//      window, or whose cell is clamped (the reference's rule `num > K`
//      fast-accepts a cell, optixScan.cpp:987-993);
//   3. each candidate's exact count over each run: with x the candidate and
//      f(v) = floor(v 2^s / c) its run's sub-cell key, every point whose key
//      lies in [f(x - R(1-1e-5)) + 1, f(x + R(1-1e-5)) - 1] passes the FP32
//      test d2 <= fl(R*R) and every point below f(x - R(1+1e-5)) or above
//      f(x + R(1+1e-5)) fails it (the margins exceed the rounding of the
//      double keys and of the FP32 test for every unclamped cell), so the
//      count is a difference of two bucket starts plus the exact FP32 test
//      of the points in the few boundary buckets;
//   4. counts <= K set bits by ring slot in an LDS bitmap, listed ascending
//      straight into pinned host memory with the slide's counters, then the
//      slide's stamp after a system-scope release.
//
// Reference counterparts: prepare_c_non_points_queue (optixScan.cpp:
// 967-1017: cell map, dense cells, the sparse cells' points), rebuild_gas /
// kGenAABB + launch + __intersection__cube (optixScan.cpp:378-426, 727-750,
// optixScan.cu:158-199: the candidates' neighbour counts), D2H + select
// (optixScan.cpp:1083-1101).
#pragma once

#include <climits>

#include "rtod_plain.h"

namespace rtod {

constexpr int kRunNT = 1024;                 // threads of every runs workgroup (four
                                             // waves per SIMD hide the LDS and memory
                                             // latency of its chain; 256 threads: 2x slower)
constexpr int kRunIT = 8;                    // points per thread in a run's sort
constexpr int kRunMaxLen = kRunNT * kRunIT;  // points per run (a slide: 1 or 2 runs)
constexpr int kRunBMax = 4096;               // buckets of a direct-mode run
constexpr int kRunSMax = 8;                  // at most 256 sub-cells per cell
constexpr int kRunMaxRuns = 256;             // runs per window (m x runs per slide)
constexpr int kRunHCap = 8192;               // cells of the window histogram in LDS
constexpr int kRunCandCap = 2048;            // candidates counted per round
constexpr int kRunECap = 2048;               // cell-table entries staged in LDS per round
// A pipelined run's slides go in batches: the first kRunFirstBatch slides
// (sorted by the prologue), then kRunBatch each from HBM, kRunPinnedBatch
// from a rtod_host_alloc buffer and kRunHostBatch staged from pageable host
// memory; launch k detects batch k and sorts batch k + 1.  The small first
// batch keeps the prologue, whose sorts nothing overlaps, short.  Over
// PCIe smaller batches pay: staged, the host's copy of a batch bounds each
// launch and a run's fill and drain cost less (64 measured 0.08 us per slide
// slower than 32 over 160-slide calls); from a pinned buffer 128 measured
// 0.78 us per slide against 0.73-0.75 for 64.
constexpr int kRunBatch = 128;               // slides per full pipelined launch (from HBM)
constexpr int kRunPinnedBatch = 64;          // the same, from a rtod_host_alloc buffer
constexpr int kRunHostBatch = 32;            // the same, staged from pageable memory
constexpr int kRunFirstBatch = 32;           // a pipelined run's first batch
// host-out buffers of a runs context: a pipelined run holds at most two
// batches' slides uncollected
constexpr int kRunHostOuts = 2 * kRunBatch;
// Pinned staging slots of a runs context: a pipelined run stages batch k + 1
// (the slides launch k sorts) once the batches before k - 1 are collected.
// Batch k's slots may still be read by the sorts of launch k - 1, and, for
// k = 1, batch 0's by the prologue's, which no collection covers: three
// batches' worth of slots rotate
constexpr int kRunStageSlots = 3 * kRunBatch;
constexpr int kRunMaxW = 1 << 18;            // the outlier bitmap of W bits lives in LDS

constexpr int kErrRunBucket = 32;
constexpr int kRunRankCap = 256; // outliers listed by rank (beyond: from the bitmap)            // a point's bucket fell outside its run's range

enum { kRunDirect = 0, kRunSearch = 1 };

// One sorted run (written by its sort, read by the detects of the m slides
// whose windows hold it).
struct RunMeta {
    int32_t n;          // points
    int32_t mode;       // kRunDirect / kRunSearch
    int32_t s;          // direct: 2^s sub-cells per cell (search: 0)
    int32_t nb;         // direct: buckets, (c_hi - c_lo + 1) << s
    int32_t c_lo, c_hi; // cell range (clamped cell coordinates)
    int32_t ne;         // cells holding points (cell-table entries)
    int32_t pad;
};

// Device state of a runs context.  Run (q, j): part j of run id q, where
// the initial window's batch b is q = b and 0-based slide i is q = m + i;
// it lives at ring entry (q mod NE) * rps + j and holds ring slots
// (q mod m) * S + j * RL + [0, n).
struct RunsDev {
    int W, S, m, K;
    int rps;            // runs per slide (ceil(S / kRunMaxLen))
    int RL;             // points per run (ceil(S / rps))
    int NE;             // run-ring entries (each rps runs)
    int n_hout;         // host-out buffers (slide i uses i mod n_hout)
    float r2;           // fl(R * R)
    int pad0;
    double inv_c;       // 1 / cell side
    double R_in;        // R (1 - 1e-5): keys strictly inside are neighbours
    double R_out;       // R (1 + 1e-5): keys strictly outside are not
    float *rx;          // [NE*rps][RL] coordinates in bucket order
    int32_t *rslot;     // [NE*rps][RL] their ring slots
    int32_t *rpre;      // [NE*rps][kRunBMax + 1] bucket starts (direct mode)
    int32_t *rcell;     // [NE*rps][RL] cells holding points, ascending
    int32_t *rcs;       // [NE*rps][RL + 1] index of each cell's first point (then n)
    RunMeta *meta;      // [NE*rps]
    int32_t *err;       // error flags (kErrRunBucket); reported and cleared by a detect
    uint64_t *tstamp;   // [kTsCount] stamps of a slide run one per call
    HostOut *h_out;     // n_hout buffers of h_out_bytes each (pinned host memory)
    int64_t h_out_bytes;
};

// The work of one runs launch: the sorts of nsort slides (rps workgroups
// each) then the detects of ndet slides (one workgroup each); fused: one
// workgroup sorts slide det_i0's run and then detects the slide.
struct RunJob {
    int nsort;
    int ndet;
    int fused;
    int stamps;             // one slide per call: stamp its phases
    int src_ring;           // the sorted slides' points: run_job_src
    int pad;
    int64_t sort_q0;        // run id of the first slide sorted
    int64_t det_i0;         // first slide detected (0-based)
    int64_t src_i0;         // the sorted slides' points: run_job_src
    const float *src;
};

// The points of the job's k-th sorted slide: src + ((src_i0 + k) mod
// src_ring) * S in a staging ring (src_ring > 0), else src + (src_i0 + k) * S
// (a slide run from HBM, the initial window's batches, one slide per call).
// A job's description stays a few words however many slides it sorts (the
// launch's arguments: a larger block made each launch call slower).
__device__ __forceinline__ const float *run_job_src(const RunsDev &d, const RunJob &job, int k)
{
    const int64_t i = job.src_i0 + k;
    return job.src + (size_t)(job.src_ring > 0 ? i % job.src_ring : i) * d.S;
}

__device__ __forceinline__ HostOut *run_host_out(const RunsDev &d, int64_t i)
{
    return reinterpret_cast<HostOut *>(reinterpret_cast<char *>(d.h_out) +
                                       (size_t)(i % d.n_hout) * (size_t)d.h_out_bytes);
}

// Sub-cell key: floor(v * 2^s / c) in double, clamped to [-2^(30+s),
// 2^(30+s) - 1].  inv_s = 2^s / c exactly (a power-of-two multiple of 1/c),
// so key >> s equals cell_coord<1> (the cell key of every other kernel and
// of the oracle), clamping included, and NaN takes the lowest key.
__device__ __forceinline__ int64_t sub_coord(double v, double inv_s, int s)
{
    double f = floor(v * inv_s);
    const double lo = -(double)(1ll << (30 + s));
    const double hi = (double)((1ll << (30 + s)) - 1);
    if (!(f > lo))
        f = lo;
    else if (!(f < hi))
        f = hi;
    return (int64_t)f;
}

// sub_coord's key as a double (an integer below 2^53 in magnitude, exact)
__device__ __forceinline__ double sub_key_d(double v, double inv_s, int s)
{
    double f = floor(v * inv_s);
    const double lo = -(double)(1ll << (30 + s));
    const double hi = (double)((1ll << (30 + s)) - 1);
    if (!(f > lo))
        f = lo;
    else if (!(f < hi))
        f = hi;
    return f;
}

// a cell whose diameter bound holds (strictly inside the clamp range)
__device__ __forceinline__ bool run_cell_guaranteed(int32_t c)
{
    return c > -(1 << 30) && c < (1 << 30) - 1;
}

// Exclusive scan of one int per thread over the workgroup: the waves'
// totals go through LDS once, and each wave scans them in registers (DPP),
// so no thread loops over the waves.  s_tmp: kRunNT / kWave ints.  Two
// barriers; every thread gets the total.
__device__ __forceinline__ int run_scan(int x, int32_t *s_tmp, int &total)
{
    constexpr int NW = kRunNT / kWave;
    const int lane = lane_id(), wid = threadIdx.x / kWave;
    const int inc = wave_incl_scan(x);
    if (lane == kWave - 1)
        s_tmp[wid] = inc;
    __syncthreads();
    const int v = lane < NW ? s_tmp[lane] : 0;
    const int sc = wave_incl_scan(v);
    total = __builtin_amdgcn_readlane(sc, NW - 1);
    const int before = __builtin_amdgcn_readlane(sc - v, wid);
    __syncthreads(); // s_tmp is free for the next scan
    return before + inc - x;
}

// In-place exclusive scan of a[0, n) in LDS (n <= 32 * kRunNT), a[n] =
// total.  a is 16-byte aligned with room for n rounded up to 4, plus one;
// each thread scans a contiguous run of 16-byte groups.
__device__ __forceinline__ void run_lds_scan(int32_t *a, int n, int32_t *s_tmp)
{
    if (n <= 4 * kWave) { // n uniform: up to 256 values scanned by wave 0 alone (one barrier)
        if (threadIdx.x < kWave) {
            const int b = 4 * (int)threadIdx.x;
            int x[4];
#pragma unroll
            for (int j = 0; j < 4; ++j)
                x[j] = b + j < n ? a[b + j] : 0;
            const int loc = x[0] + x[1] + x[2] + x[3];
            const int inc = wave_incl_scan(loc);
            int run = inc - loc;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (b + j < n)
                    a[b + j] = run;
                run += x[j];
            }
            if (threadIdx.x == kWave - 1)
                a[n] = inc;
        }
        __syncthreads();
        return;
    }
    constexpr int MAXG = 8;
    int4 *a4 = reinterpret_cast<int4 *>(a);
    const int n4 = (n + 3) >> 2;
    const int per = (n4 + kRunNT - 1) / kRunNT;
    const int g0 = threadIdx.x * per;
    int4 v[MAXG];
    int loc = 0;
#pragma unroll
    for (int j = 0; j < MAXG; ++j) {
        const int g = g0 + j;
        v[j] = make_int4(0, 0, 0, 0);
        if (j < per && g < n4) {
            v[j] = a4[g];
            const int b = g * 4;
            if (b + 1 >= n)
                v[j].y = 0;
            if (b + 2 >= n)
                v[j].z = 0;
            if (b + 3 >= n)
                v[j].w = 0;
        }
        loc += v[j].x + v[j].y + v[j].z + v[j].w;
    }
    int total;
    int run = run_scan(loc, s_tmp, total);
#pragma unroll
    for (int j = 0; j < MAXG; ++j) {
        const int g = g0 + j;
        if (j < per && g < n4) {
            const int b = g * 4;
            const int4 o = make_int4(run, run + v[j].x, run + v[j].x + v[j].y,
                                     run + v[j].x + v[j].y + v[j].z);
            if (b + 3 < n) {
                a4[g] = o;
            } else {
                a[b] = o.x;
                if (b + 1 < n)
                    a[b + 1] = o.y;
                if (b + 2 < n)
                    a[b + 2] = o.z;
            }
            run += v[j].x + v[j].y + v[j].z + v[j].w;
        }
    }
    if (threadIdx.x == 0)
        a[n] = total;
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Sort of one run: S points (or RL of them) from the slide's source (pinned
// staging, or HBM) into ring entry e.
// ---------------------------------------------------------------------------
struct RunSortLds {
    float x[kRunMaxLen]; // arrival order
    union {
        struct {
            alignas(16) int32_t hist[kRunBMax + 4]; // bucket counts, then starts
            float sx[kRunMaxLen];       // coordinates in bucket order
            uint16_t sidx[kRunMaxLen];  // their arrival index
        } dir;
        struct {
            uint64_t key[kRunMaxLen];   // (cell - c_lo) << 13 | arrival index
            int32_t hist[16 * (kRunNT / kWave)];
        } sea;
    } u;
    float red[3][kRunNT / kWave];
    int32_t scan[kRunNT / kWave];
};

// ---- pieces of a run's sort -------------------------------------------------

// n points from g into LDS x: one round of loads (16-byte loads when g is
// 16-byte aligned)
__device__ __forceinline__ void run_load_points(float *x, const float *g, int n)
{
    if ((((uintptr_t)g) & 15) == 0) {
        const int n4 = n >> 2;
        copy_to_lds<(kRunMaxLen / 4 + kRunNT - 1) / kRunNT>(reinterpret_cast<float4 *>(x),
                                                           reinterpret_cast<const float4 *>(g), n4);
        for (int k = (n4 << 2) + (int)threadIdx.x; k < n; k += kRunNT)
            x[k] = g[k];
    } else {
        copy_to_lds<kRunIT>(x, g, n);
    }
}

// The cell range of x[0, n): cell_coord is monotone in x and puts NaN in the
// lowest cell, so it is [cell(min x), cell(max x)] over the non-NaN points,
// from the lowest cell if a NaN is present.  Block-wide (one barrier).
__device__ __forceinline__ void run_cell_range(const RunsDev &d, const float *x, int n,
                                               float (*red)[kRunNT / kWave], int &c_lo, int &c_hi)
{
    constexpr int NW = kRunNT / kWave;
    const int tid = threadIdx.x, lane = lane_id(), wid = tid / kWave;
    float xmn = __builtin_inff(), xmx = -__builtin_inff(), nan = 0.f;
#pragma unroll
    for (int i = 0; i < kRunIT; ++i) {
        const int l = tid + i * kRunNT;
        if (l < n) {
            const float v = x[l];
            if (v != v) {
                nan = 1.f;
            } else {
                xmn = fminf(xmn, v);
                xmx = fmaxf(xmx, v);
            }
        }
    }
    xmn = wave_min_f32(xmn);
    xmx = wave_max_f32(xmx);
    nan = wave_max_f32(nan);
    if (lane == 0) {
        red[0][wid] = xmn;
        red[1][wid] = xmx;
        red[2][wid] = nan;
    }
    __syncthreads();
    xmn = wave_min_f32(lane < NW ? red[0][lane] : __builtin_inff());
    xmx = wave_max_f32(lane < NW ? red[1][lane] : -__builtin_inff());
    nan = wave_max_f32(lane < NW ? red[2][lane] : 0.f);
    c_lo = cell_coord<1>((double)xmn, d.inv_c);
    c_hi = cell_coord<1>((double)xmx, d.inv_c);
    if (nan != 0.f)
        c_lo = -(1 << 30);
    if (c_hi < c_lo) // every point NaN (min x = +inf)
        c_hi = c_lo;
}

// direct mode: the largest s <= kRunSMax with crange << s <= kRunBMax buckets
__device__ __forceinline__ int run_shift(int64_t crange)
{
    int s = 0;
    while (s < kRunSMax && (crange << (s + 1)) <= kRunBMax)
        ++s;
    return s;
}

// Counting sort of x[0, n) by sub-cell bucket (nb buckets from cell c_lo,
// 2^s per cell): sx/sidx hold the points in bucket order with their arrival
// index, hist[0, nb] the bucket starts (hist[nb] = n).  zeroed: hist[0, nb]
// is already zero and a barrier has passed since.  Errors go to *err (HBM
// or LDS), the stamp to ts (none when null).  Ends with a barrier.
__device__ __forceinline__ void run_bucket(const RunsDev &d, const float *x, int n, int c_lo,
                                           int s, int nb, int32_t *hist, float *sx,
                                           uint16_t *sidx, int32_t *scan, int32_t *err,
                                           uint64_t *ts, bool zeroed)
{
    constexpr int NT = kRunNT, IT = kRunIT;
    const int tid = threadIdx.x;
    const double inv_s = d.inv_c * (double)(1 << s);
    const double base = (double)c_lo * (double)(1 << s);
    if (!zeroed) {
        for (int b = tid; b <= nb; b += NT)
            hist[b] = 0;
        __syncthreads();
    }
    int bk[IT], rk[IT];
    int bad = 0;
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int l = tid + i * NT;
        bk[i] = 0;
        rk[i] = 0;
        if (l < n) {
            // the key relative to the run's first bucket, in double (exact:
            // integers below 2^53), converted once to int32
            const double fr = sub_key_d((double)x[l], inv_s, s) - base;
            int f = (fr >= 0.0 && fr < (double)nb) ? (int)fr : -1;
            if (f < 0) { // cannot happen: key >> s lies in [c_lo, c_hi]
                bad = 1;
                f = fr < 0.0 ? 0 : nb - 1;
            }
            bk[i] = f;
            rk[i] = atomicAdd(&hist[f], 1);
        }
    }
    if (bad)
        atomicOr(err, kErrRunBucket);
    __syncthreads();
    run_lds_scan(hist, nb, scan); // hist[b]: first position of bucket b; hist[nb] = n
    if (ts && tid == 0)
        ts[kTsMerge] = wall_clock();
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int l = tid + i * NT;
        if (l < n) {
            const int pos = hist[bk[i]] + rk[i];
            sx[pos] = x[l];
            sidx[pos] = (uint16_t)l;
        }
    }
    __syncthreads();
}

// A direct-mode run from LDS (bucket order sx/sidx, bucket starts hist)
// into ring entry e: coordinates, slots, bucket starts, cell table and
// descriptor.  With ecell != nullptr the cell table is also listed as
// staged window entries of run `er` (first point | points << 16).
// Returns the cell count (every thread).
__device__ __forceinline__ int run_write_direct(const RunsDev &d, int e, int n, int slot0, int s,
                                                int nb, int c_lo, int c_hi, const int32_t *hist,
                                                const float *sx, const uint16_t *sidx,
                                                int32_t *scan, int32_t *ecell, uint32_t *epos,
                                                uint8_t *erun, int er)
{
    constexpr int NT = kRunNT;
    const int tid = threadIdx.x;
    float *ox = d.rx + (size_t)e * d.RL;
    int32_t *os = d.rslot + (size_t)e * d.RL;
    int32_t *oc = d.rcell + (size_t)e * d.RL;
    int32_t *ocs = d.rcs + (size_t)e * (d.RL + 1);
    for (int p = tid; p < n; p += NT) {
        ox[p] = sx[p];
        os[p] = slot0 + (int)sidx[p];
    }
    int32_t *op = d.rpre + (size_t)e * (kRunBMax + 1);
    for (int b = tid; b <= nb; b += NT)
        op[b] = hist[b];
    // the cell table: cells of [c_lo, c_hi] holding points, ascending
    const int nc = c_hi - c_lo + 1;
    const int per = (nc + NT - 1) / NT;
    const int k0 = tid * per;
    int mine = 0;
    for (int j = 0; j < per; ++j) {
        const int k = k0 + j;
        if (k < nc && hist[(k + 1) << s] > hist[k << s])
            ++mine;
    }
    int ne;
    int o = run_scan(mine, scan, ne);
    for (int j = 0; j < per; ++j) {
        const int k = k0 + j;
        if (k >= nc)
            break;
        const int a = hist[k << s];
        if (hist[(k + 1) << s] > a) {
            oc[o] = c_lo + k;
            ocs[o] = a;
            if (ecell) {
                ecell[o] = c_lo + k;
                epos[o] = (uint32_t)a | ((uint32_t)(hist[(k + 1) << s] - a) << 16);
                erun[o] = (uint8_t)er;
            }
            ++o;
        }
    }
    if (tid == 0) {
        ocs[ne] = n;
        d.meta[e] = RunMeta{n, kRunDirect, s, nb, c_lo, c_hi, ne, 0};
    }
    return ne;
}

__device__ __forceinline__ void run_sort_body(const RunsDev &d, int64_t q, int part,
                                              const float *src, uint64_t *ts, RunSortLds &L)
{
    constexpr int NT = kRunNT, IT = kRunIT;
    const int tid = threadIdx.x, lane = lane_id(), wid = tid / kWave;
    const int e = (int)(q % d.NE) * d.rps + part;
    const int j0 = part * d.RL;
    const int n = min(d.RL, d.S - j0);
    const int slot0 = (int)(q % d.m) * d.S + j0;

    run_load_points(L.x, src + j0, n);
    __syncthreads();
    int c_lo, c_hi;
    run_cell_range(d, L.x, n, L.red, c_lo, c_hi);
    const int64_t crange = (int64_t)c_hi - (int64_t)c_lo + 1;

    if (crange <= kRunBMax) {
        // ---- direct mode: counting sort by sub-cell bucket -----------------
        const int s = run_shift(crange);
        const int nb = (int)(crange << s);
        run_bucket(d, L.x, n, c_lo, s, nb, L.u.dir.hist, L.u.dir.sx, L.u.dir.sidx, L.scan,
                   d.err, ts, false);
        run_write_direct(d, e, n, slot0, s, nb, c_lo, c_hi, L.u.dir.hist, L.u.dir.sx,
                         L.u.dir.sidx, L.scan, nullptr, nullptr, nullptr, 0);
        return;
    }

    // ---- search mode: LDS radix sort by cell, cell table by run length ----
    float *ox = d.rx + (size_t)e * d.RL;
    int32_t *os = d.rslot + (size_t)e * d.RL;
    int32_t *oc = d.rcell + (size_t)e * d.RL;
    int32_t *ocs = d.rcs + (size_t)e * (d.RL + 1);
    uint64_t k[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int l = wid * kWave * IT + i * kWave + lane;
        k[i] = ~0ull;
        if (l < n) {
            const int c = cell_coord<1>((double)L.x[l], d.inv_c);
            k[i] = ((uint64_t)((int64_t)c - (int64_t)c_lo) << 13) | (uint64_t)l;
        }
        L.u.sea.key[l] = k[i];
    }
    __syncthreads();
    block_radix_sort<NT, IT>(k, 13, bit_length((uint64_t)(crange - 1)), L.u.sea.key,
                             L.u.sea.hist);
    if (ts && tid == 0)
        ts[kTsMerge] = wall_clock();
    for (int p = tid; p < n; p += NT) {
        const int l = (int)(L.u.sea.key[p] & 8191ull);
        ox[p] = L.x[l];
        os[p] = slot0 + l;
    }
    // a cell starts at p when its cell differs from p - 1's
    const int per = (n + NT - 1) / NT;
    const int p0 = tid * per;
    int mine = 0;
    for (int j = 0; j < per; ++j) {
        const int p = p0 + j;
        if (p < n && (p == 0 || (L.u.sea.key[p] >> 13) != (L.u.sea.key[p - 1] >> 13)))
            ++mine;
    }
    int ne;
    int o = run_scan(mine, L.scan, ne);
    for (int j = 0; j < per; ++j) {
        const int p = p0 + j;
        if (p < n && (p == 0 || (L.u.sea.key[p] >> 13) != (L.u.sea.key[p - 1] >> 13))) {
            oc[o] = (int32_t)((int64_t)c_lo + (int64_t)(L.u.sea.key[p] >> 13));
            ocs[o] = p;
            ++o;
        }
    }
    if (tid == 0) {
        ocs[ne] = n;
        d.meta[e] = RunMeta{n, kRunSearch, 0, 0, c_lo, c_hi, ne, 0};
    }
}

// ---------------------------------------------------------------------------
// Detect of one slide over the m runs of its window.
// ---------------------------------------------------------------------------
struct RunDetLds {
    RunMeta meta[kRunMaxRuns];
    int32_t gidx[kRunMaxRuns];         // ring entry of window run r
    int32_t ebase[kRunMaxRuns + 4];    // cell-table entries before run r
    alignas(16) int32_t hist[kRunHCap]; // window cell counts over [cmin, cmax]
    int32_t ecell[kRunECap];           // the round's cell-table entries: cell,
    uint32_t epos[kRunECap];           //   first point | points << 16,
    uint8_t erun[kRunECap];            //   window run
    int32_t rent[kRunECap];            // candidate ranges: staged entry,
    alignas(16) int32_t rpos[kRunECap + 4]; //   candidates before it
    float cx[kRunCandCap];             // the round's candidates
    int32_t cslot[kRunCandCap];
    int32_t ccnt[kRunCandCap];
    int32_t scan[kRunNT / kWave];
    int32_t red[3][kRunNT / kWave];
    int32_t olist[kRunRankCap];        // the first outliers' slots, unordered
    int32_t nrange;
    int32_t nout;                      // outliers found
    int32_t pad[2];
    unsigned long long tests, hits;
    uint32_t bitmap[4];                // ceil(W / 32) words follow (sized at launch)
};

__host__ __device__ constexpr size_t run_det_lds_bytes(int W)
{
    return sizeof(RunDetLds) + (size_t)((W + 31) / 32) * sizeof(uint32_t);
}

static_assert(sizeof(RunSortLds) <= 160 * 1024 && sizeof(RunDetLds) + kRunMaxW / 8 <= 160 * 1024,
              "a runs workgroup's LDS fits the 160 KB of a CU");

__host__ __device__ constexpr size_t run_lds_bytes(int W, bool sort, bool det)
{
    const size_t a = sort ? sizeof(RunSortLds) : 0;
    const size_t b = det ? run_det_lds_bytes(W) : 0;
    return a > b ? a : b;
}

// the window run holding cell-table entry e: the last r with ebase[r] <= e
__device__ __forceinline__ int run_of_entry(const int32_t *ebase, int nr, int e)
{
    int lo = 0, hi = nr - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (ebase[mid] <= e)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// first point of run g whose key (at the run's resolution) is >= f
__device__ __forceinline__ int run_lb(const RunsDev &d, const RunMeta &M, int g, int64_t f)
{
    if (M.mode == kRunDirect) {
        int64_t b = f - (int64_t)M.c_lo * (1ll << M.s);
        b = b < 0 ? 0 : (b > M.nb ? M.nb : b);
        return d.rpre[(size_t)g * (kRunBMax + 1) + (size_t)b];
    }
    const int32_t *cells = d.rcell + (size_t)g * d.RL;
    int lo = 0, hi = M.ne;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((int64_t)cells[mid] < f)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo == M.ne ? M.n : d.rcs[(size_t)g * (d.RL + 1) + lo];
}

// points of run g in cell c (search of its cell table)
__device__ __forceinline__ int run_cell_count(const RunsDev &d, const RunMeta &M, int g, int32_t c)
{
    const int32_t *cells = d.rcell + (size_t)g * d.RL;
    int lo = 0, hi = M.ne;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cells[mid] < c)
            lo = mid + 1;
        else
            hi = mid;
    }
    if (lo == M.ne || cells[lo] != c)
        return 0;
    const int32_t *cs = d.rcs + (size_t)g * (d.RL + 1);
    return cs[lo + 1] - cs[lo];
}

// Exact count of the points q of a run with FP32 (xp - xq)^2 <= r2 (xp
// finite): keys in [f1, f2) are neighbours, keys below f0 or from f3 are not,
// the points of the boundary buckets [f0, f1) and [f2, f3) are tested.
// lb(f): the run's first point whose key (at its resolution s) is >= f.
template <typename LB>
__device__ __forceinline__ int run_count_keys(const RunsDev &d, int s, const float *xs, float xp,
                                              unsigned long long &tests, LB lb)
{
    const double inv_s = d.inv_c * (double)(1 << s);
    const double xd = (double)xp;
    const int64_t f0 = sub_coord(xd - d.R_out, inv_s, s);     // below: out
    const int64_t f1 = sub_coord(xd - d.R_in, inv_s, s) + 1;  // [f1, f2): in
    const int64_t f2 = sub_coord(xd + d.R_in, inv_s, s);
    const int64_t f3 = sub_coord(xd + d.R_out, inv_s, s) + 1; // from f3: out
    const int i0 = lb(f0);
    int i1 = lb(f1);
    int i2 = lb(f2);
    int i3 = lb(f3);
    i1 = max(i1, i0);
    i2 = max(i2, i1);
    i3 = max(i3, i2);
    int cnt = i2 - i1;
    const float r2 = d.r2;
    for (int k = i0; k < i1; ++k) {
        const float dx = xp - xs[k];
        cnt += (dx * dx <= r2) ? 1 : 0;
    }
    for (int k = i2; k < i3; ++k) {
        const float dx = xp - xs[k];
        cnt += (dx * dx <= r2) ? 1 : 0;
    }
    tests += (unsigned long long)(i1 - i0 + i3 - i2);
    return cnt;
}

// a direct-mode run given by its bucket starts pre[0, nb] and coordinates
// xs (HBM, or LDS for the run a one-slide workgroup has just sorted): the
// four keys of run_count_keys as bucket indices clamped to [0, nb], in
// double (exact integers) and converted once to int32
__device__ __forceinline__ int run_count_direct(const RunsDev &d, const RunMeta &M,
                                                const int32_t *pre, const float *xs, float xp,
                                                unsigned long long &tests)
{
    const int s = M.s;
    const double inv_s = d.inv_c * (double)(1 << s);
    const double base = (double)M.c_lo * (double)(1 << s);
    const double nbd = (double)M.nb;
    const double xd = (double)xp;
    auto idx = [&](double v, double plus) {
        const double b = sub_key_d(v, inv_s, s) + plus - base;
        return (int)(b < 0.0 ? 0.0 : (b > nbd ? nbd : b));
    };
    const int i0 = pre[idx(xd - d.R_out, 0.0)];     // below: out
    int i1 = pre[idx(xd - d.R_in, 1.0)];            // [i1, i2): in
    int i2 = pre[idx(xd + d.R_in, 0.0)];
    int i3 = pre[idx(xd + d.R_out, 1.0)];           // from i3: out
    i1 = max(i1, i0);
    i2 = max(i2, i1);
    i3 = max(i3, i2);
    int cnt = i2 - i1;
    const float r2 = d.r2;
    for (int k = i0; k < i1; ++k) {
        const float dx = xp - xs[k];
        cnt += (dx * dx <= r2) ? 1 : 0;
    }
    for (int k = i2; k < i3; ++k) {
        const float dx = xp - xs[k];
        cnt += (dx * dx <= r2) ? 1 : 0;
    }
    tests += (unsigned)(i1 - i0 + i3 - i2);
    return cnt;
}

// exact count over run g of the window (HBM)
__device__ __forceinline__ int run_count(const RunsDev &d, const RunMeta &M, int g, float xp,
                                         unsigned long long &tests)
{
    const float *xs = d.rx + (size_t)g * d.RL;
    if (M.mode == kRunDirect)
        return run_count_direct(d, M, d.rpre + (size_t)g * (kRunBMax + 1), xs, xp, tests);
    return run_count_keys(d, 0, xs, xp, tests, [&](int64_t f) { return run_lb(d, M, g, f); });
}

// candidate slot sl is an outlier: its bit, and its slot in the list while
// the list has room
__device__ __forceinline__ void run_mark_outlier(uint32_t *bitmap, int32_t *olist, int32_t *nout,
                                                 int sl)
{
    atomicOr(&bitmap[sl >> 5], 1u << (sl & 31));
    const int k = atomicAdd(nout, 1);
    if (k < kRunRankCap)
        olist[k] = sl;
}

// the threads' pair-test and hit counts into the workgroup's totals in LDS:
// summed per wave first, one atomic per wave (called by every thread)
__device__ __forceinline__ void run_add_counters(unsigned long long tests,
                                                 unsigned long long hits,
                                                 unsigned long long *l_tests,
                                                 unsigned long long *l_hits)
{
    tests = wave_sum_u64(tests);
    hits = wave_sum_u64(hits);
    if (lane_id() == 0) {
        if (tests)
            atomicAdd(l_tests, tests);
        if (hits)
            atomicAdd(l_hits, hits);
    }
}

// The slide's outliers listed ascending into slide i's host-out buffer
// with the slide's counters, then its stamp after one system-scope release.
// Up to kRunRankCap outliers (olist/nout, final like the bitmap at the
// caller's last barrier) each go straight to their rank among the others;
// more are listed from the bitmap (ceil(W / 32) words in LDS), a popcount
// and a scan over all words.  tests/hits: the workgroup's totals in LDS,
// final at the first barrier inside.  err: the slide's error flags (thread
// 0's value); ts: the slide's stamps (LDS or HBM; none when null).
__device__ __forceinline__ void run_publish(const RunsDev &d, int64_t i, const uint32_t *bitmap,
                                            const int32_t *olist, const int32_t *nout,
                                            int32_t *scan, int ncand, int nranges,
                                            const unsigned long long *tests,
                                            const unsigned long long *hits, uint64_t *ts, int err)
{
    constexpr int NT = kRunNT;
    const int tid = threadIdx.x;
    if (ts && tid == 0)
        ts[kTsFinish] = wall_clock();
    HostOut *ho = run_host_out(d, i);
    int total = *nout;
    if (total <= kRunRankCap) {
        for (int k = tid; k < total; k += NT) {
            const int sl = olist[k];
            int rank = 0;
            for (int j = 0; j < total; ++j)
                rank += olist[j] < sl ? 1 : 0; // slots are distinct
            ho->list[rank] = sl;
        }
    } else {
        const int words = (d.W + 31) >> 5;
        const int wpt = (words + NT - 1) / NT;
        const int w0 = tid * wpt, w1 = min(words, w0 + wpt);
        int mine = 0;
        for (int w = w0; w < w1; ++w)
            mine += __popc(bitmap[w]);
        int o = run_scan(mine, scan, total);
        for (int w = w0; w < w1; ++w) {
            uint32_t b = bitmap[w];
            while (b) {
                const int bit = __ffs(b) - 1;
                b &= b - 1;
                ho->list[o++] = (w << 5) + bit;
            }
        }
    }
    // every wave's list writes acknowledged, then one system-scope release
    // before the stamp (acknowledged writes to host memory are not yet
    // visible to the host without it)
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (tid == 0) {
        ho->n_out = total;
        ho->n_cand = ncand;
        ho->n_units = nranges;
        ho->pair_tests = *tests;
        ho->hits = *hits;
        ho->misses = 0;
        ho->used_pred = 0;
        ho->learn_kernels = 0;
        ho->train_steps = 0;
        ho->train_loss = 0.f;
        ho->err = err;
        if (ts) {
#pragma unroll
            for (int k = 0; k < kTsPublished; ++k)
                ho->ts[k] = ts[k];
            ho->ts[kTsPublished] = wall_clock();
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
        __hip_atomic_store(&ho->slide, (int64_t)(i + 1), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// stage cell-table entries [e0, e1) of the window (e1 - e0 <= kRunECap)
__device__ __forceinline__ void run_stage_entries(const RunsDev &d, RunDetLds &L, int nr, int e0,
                                                  int e1)
{
    for (int e = e0 + (int)threadIdx.x; e < e1; e += kRunNT) {
        const int r = run_of_entry(L.ebase, nr, e);
        const int g = L.gidx[r];
        const int k = e - L.ebase[r];
        const int32_t c = d.rcell[(size_t)g * d.RL + k];
        const int32_t *cs = d.rcs + (size_t)g * (d.RL + 1) + k;
        const int a = cs[0], b = cs[1];
        L.ecell[e - e0] = c;
        L.epos[e - e0] = (uint32_t)a | ((uint32_t)(b - a) << 16);
        L.erun[e - e0] = (uint8_t)r;
    }
}

__device__ __forceinline__ void run_detect_body(const RunsDev &d, int64_t i, uint64_t *ts,
                                                RunDetLds &L)
{
    constexpr int NT = kRunNT, NW = NT / kWave;
    const int tid = threadIdx.x, lane = lane_id(), wid = tid / kWave;
    const int nr = d.m * d.rps;
    // the error flags raised before this slide's detect (its sort, earlier
    // launches), taken now so that the exchange's round trip overlaps the
    // detect (a detect raises none of its own)
    int err = 0;
    if (tid == 0)
        err = atomicExch(d.err, 0);
    if (ts && tid == 0)
        ts[kTsClassify] = wall_clock();
    const int words = (d.W + 31) >> 5;
    for (int w = tid; w < words; w += NT)
        L.bitmap[w] = 0;
    if (tid == 0) {
        L.tests = 0;
        L.hits = 0;
        L.nout = 0;
    }

    // the window's runs: q = i + 1 .. i + m
    int cl = INT_MAX, ch = INT_MIN, srch = 0, ne = 0;
    if (tid < nr) {
        const int64_t q = i + 1 + tid / d.rps;
        const int g = (int)(q % d.NE) * d.rps + tid % d.rps;
        const RunMeta M = d.meta[g];
        L.meta[tid] = M;
        L.gidx[tid] = g;
        ne = M.ne;
        cl = M.c_lo;
        ch = M.c_hi;
        srch = M.mode == kRunSearch ? 1 : 0;
    }
    int E;
    const int pre = run_scan(ne, L.scan, E);
    if (tid < nr)
        L.ebase[tid] = pre;
    cl = wave_min_i32(cl);
    ch = wave_max_i32(ch);
    srch = wave_or_i32(srch);
    if (lane == 0) {
        L.red[0][wid] = cl;
        L.red[1][wid] = ch;
        L.red[2][wid] = srch;
    }
    __syncthreads();
    const int cmin = wave_min_i32(lane < NW ? L.red[0][lane] : INT_MAX);
    const int cmax = wave_max_i32(lane < NW ? L.red[1][lane] : INT_MIN);
    const int any_search = wave_or_i32(lane < NW ? L.red[2][lane] : 0);
    const bool fast = !any_search && (int64_t)cmax - (int64_t)cmin + 1 <= kRunHCap;
    const bool one_round = E <= kRunECap;

    // 1. window cell counts (fast path: LDS histogram of the staged entries)
    if (fast) {
        const int cw = cmax - cmin + 1;
        for (int c = tid; c < cw; c += NT)
            L.hist[c] = 0;
        for (int e0 = 0; e0 < E; e0 += kRunECap) {
            const int e1 = min(E, e0 + kRunECap);
            run_stage_entries(d, L, nr, e0, e1);
            __syncthreads();
            for (int e = tid; e < e1 - e0; e += NT)
                atomicAdd(&L.hist[L.ecell[e] - cmin], (int)(L.epos[e] >> 16));
            __syncthreads();
        }
    }

    // 2-4. candidates by rounds of cell-table entries, counted by rounds
    unsigned long long tests = 0, hits = 0;
    int ncand = 0, nranges = 0;
    for (int e0 = 0; e0 < E; e0 += kRunECap) {
        const int e1 = min(E, e0 + kRunECap);
        if (!(fast && one_round)) { // else staged by the histogram pass
            run_stage_entries(d, L, nr, e0, e1);
        }
        if (tid == 0)
            L.nrange = 0;
        __syncthreads();
        for (int e = tid; e < e1 - e0; e += NT) {
            const int32_t c = L.ecell[e];
            bool sparse = !run_cell_guaranteed(c);
            if (!sparse) {
                int total = 0;
                if (fast) {
                    total = L.hist[c - cmin];
                } else {
                    for (int r2 = 0; r2 < nr && total <= d.K; ++r2)
                        total += run_cell_count(d, L.meta[r2], L.gidx[r2], c);
                }
                sparse = total <= d.K;
            }
            if (sparse) {
                const int slot = atomicAdd(&L.nrange, 1);
                L.rent[slot] = e;
                L.rpos[slot] = (int)(L.epos[e] >> 16);
            }
        }
        __syncthreads();
        const int nrg = L.nrange;
        nranges += nrg;
        run_lds_scan(L.rpos, nrg, L.scan);
        const int nc = L.rpos[nrg];
        ncand += nc;
        if (ts && tid == 0 && e0 == 0)
            ts[kTsDetect] = wall_clock();
        for (int c0 = 0; c0 < nc; c0 += kRunCandCap) {
            const int ncs = min(kRunCandCap, nc - c0);
            for (int c = tid; c < ncs; c += NT) {
                const int cc = c0 + c;
                int lo = 0, hi = nrg - 1; // the range holding candidate cc
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (L.rpos[mid] <= cc)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                const int e = L.rent[lo];
                const int g = L.gidx[L.erun[e]];
                const size_t idx =
                    (size_t)g * d.RL + (size_t)((int)(L.epos[e] & 0xffffu) + (cc - L.rpos[lo]));
                L.cx[c] = d.rx[idx];
                L.cslot[c] = d.rslot[idx];
                L.ccnt[c] = 0;
            }
            __syncthreads();
            const int P = ncs * nr;
            // run-major: a wave's lanes take consecutive candidates against one
            // run, whose bucket starts and boundary points then share cache
            // lines (candidate-major, every lane gathered from another run)
            for (int p = tid; p < P; p += NT) {
                const int r = p / ncs, c = p - r * ncs;
                const float xp = L.cx[c];
                if (!__builtin_isfinite(xp)) // NaN and +-inf have no neighbour, not even themselves
                    continue;
                const int cnt = run_count(d, L.meta[r], L.gidx[r], xp, tests);
                if (cnt)
                    atomicAdd(&L.ccnt[c], cnt);
            }
            __syncthreads();
            for (int c = tid; c < ncs; c += NT) {
                const int cnt = L.ccnt[c];
                hits += (unsigned long long)min(cnt, d.K + 1);
                if (cnt <= d.K)
                    run_mark_outlier(L.bitmap, L.olist, &L.nout, L.cslot[c]);
            }
            __syncthreads();
        }
    }
    run_add_counters(tests, hits, &L.tests, &L.hits);
    run_publish(d, i, L.bitmap, L.olist, &L.nout, L.scan, ncand, nranges, &L.tests, &L.hits, ts,
                err);
}

// ---------------------------------------------------------------------------
// One slide per call (one run per slide): one workgroup sorts the new run
// and detects the slide, keeping the new run in LDS.  Its copy in HBM is for
// later slides only, so nothing waits for those stores; the old runs'
// descriptors load beside the new points and their cell tables while the
// new run is bucketed.  Where the window does not fit the workgroup's
// tables (a run in search mode, more than kSlideECap cell-table entries,
// a cell range beyond kSlideHCap) it sorts the run into HBM and detects
// from there like the pipelined workgroups.
// ---------------------------------------------------------------------------
constexpr int kSlideHCap = 1024;   // window cells counted in LDS
constexpr int kSlideECap = 2048;   // window cell-table entries staged in LDS
constexpr int kSlideCand = 1024;   // candidates counted per round

struct RunSlideLds {
    RunMeta meta[kRunMaxRuns];
    int32_t gidx[kRunMaxRuns];
    int32_t ebase[kRunMaxRuns + 4];
    int32_t ecell[kSlideECap];             // window cell-table entries: cell,
    uint32_t epos[kSlideECap];             //   first point | points << 16,
    uint8_t erun[kSlideECap];              //   window run (the new run: m - 1)
    alignas(16) int32_t bpre[kRunBMax + 4]; // the new run: bucket starts,
    float sx[kRunMaxLen];                  //   coordinates in bucket order,
    uint16_t sidx[kRunMaxLen];             //   their arrival index
    union {
        float x[kRunMaxLen];               // the new points in arrival order
        struct {                           // (dead once bucketed)
            int32_t hist[kSlideHCap];      // window cell counts over [cmin, cmax]
            int32_t rent[kSlideECap];      // candidate ranges: entry,
            alignas(16) int32_t rpos[kSlideECap + 4]; // candidates before it
            float cx[kSlideCand];
            int32_t cslot[kSlideCand];
            int32_t ccnt[kSlideCand];
        } d;
    } u;
    int32_t scan[kRunNT / kWave];
    float redf[3][kRunNT / kWave];
    int32_t redi[3][kRunNT / kWave];
    int32_t olist[kRunRankCap];            // the first outliers' slots, unordered
    int32_t nrange;
    int32_t err;                           // error flags raised by this workgroup
    int32_t nout;                          // outliers found
    int32_t pad;
    uint64_t ts[kTsCount];                 // the slide's stamps
    unsigned long long tests, hits;
    uint32_t bitmap[4];                    // ceil(W / 32) words follow (sized at launch)
};

__host__ __device__ constexpr size_t run_slide_lds_bytes(int W)
{
    return sizeof(RunSlideLds) + (size_t)((W + 31) / 32) * sizeof(uint32_t);
}

// a fused launch: the one-slide layout, or the sort's and the detect's
__host__ __device__ constexpr size_t run_fused_lds_bytes(int W)
{
    const size_t a = run_slide_lds_bytes(W), b = run_lds_bytes(W, true, true);
    return a > b ? a : b;
}

static_assert(sizeof(RunSlideLds) + kRunMaxW / 8 <= 160 * 1024,
              "a one-slide workgroup's LDS fits the 160 KB of a CU");

// Returns false (uniformly) when the slide takes the HBM path instead, with
// the error flags as it found them.  t_in: the launch's start stamp (thread
// 0), kept with the slide's other stamps in LDS when `stamps`.
__device__ __forceinline__ bool run_slide_body(const RunsDev &d, int64_t i, const float *src,
                                               bool stamps, uint64_t t_in, RunSlideLds &L)
{
    constexpr int NT = kRunNT, NW = NT / kWave, IT = kRunIT;
    const int tid = threadIdx.x, lane = lane_id(), wid = tid / kWave;
    const int nr = d.m, newr = d.m - 1; // one run per slide: the window's last run is new
    const int64_t qn = i + d.m;
    const int gn = (int)(qn % d.NE);
    const int n = d.S;
    const int slot0 = (int)(qn % d.m) * d.S;

    // 0. the old runs' descriptors, then the new points into registers: the
    //    descriptors (issued first) arrive first, and the old runs' cell
    //    tables are requested while the points still cross PCIe
    RunMeta M{};
    int g_old = 0;
    if (tid < newr) {
        g_old = (int)((i + 1 + tid) % d.NE);
        M = d.meta[g_old];
    }
    // the flags raised before this slide (its round trip overlaps the loads)
    int err_in = 0;
    if (tid == 0)
        err_in = atomicExch(d.err, 0);
    // The new points in 16-byte loads (a quarter of the requests over PCIe;
    // 4-byte alignment suffices), thread t holding groups t and t + NT
    // (points 4g..4g+3) and the tail point 4*n4 + t.  Every load is
    // unconditional, its index clamped into range (with no full group the
    // group loads read 16 discarded bytes of the run ring): the compiler's
    // wait counts then know that three loads follow the descriptors' and
    // wait for those alone before the descriptors are used (with loads
    // under branches it waited for every point, i.e. for PCIe).
    const int n4 = n >> 2;
    const float *b4 = n4 > 0 ? src : d.rx;
    const int gmax = n4 > 0 ? n4 - 1 : 0;
    float v[IT];
    float4 va, vb;
    __builtin_memcpy(&va, b4 + 4 * (tid < gmax ? tid : gmax), sizeof(float4));
    __builtin_memcpy(&vb, b4 + 4 * (tid + NT < gmax ? tid + NT : gmax), sizeof(float4));
    const float vt = src[4 * n4 + tid < n ? 4 * n4 + tid : n - 1];
    v[0] = va.x;
    v[1] = va.y;
    v[2] = va.z;
    v[3] = va.w;
    v[4] = vb.x;
    v[5] = vb.y;
    v[6] = vb.z;
    v[7] = vb.w;
    int ne = 0, cl = INT_MAX, ch = INT_MIN, srch = 0;
    if (tid < newr) {
        L.meta[tid] = M;
        L.gidx[tid] = g_old;
        ne = M.ne;
        cl = M.c_lo;
        ch = M.c_hi;
        srch = M.mode == kRunSearch ? 1 : 0;
    }
    if (tid == 0) {
        L.err = 0;
        L.ts[kTsIngest] = t_in;
    }
    // the new run's bucket counts zeroed while the points are in flight
    for (int b = tid; b <= kRunBMax; b += NT)
        L.bpre[b] = 0;
    int E_old;
    const int pre = run_scan(ne, L.scan, E_old);
    if (tid < newr)
        L.ebase[tid] = pre;
    cl = wave_min_i32(cl);
    ch = wave_max_i32(ch);
    srch = wave_or_i32(srch);
    if (lane == 0) {
        L.redi[0][wid] = cl;
        L.redi[1][wid] = ch;
        L.redi[2][wid] = srch;
    }
    __syncthreads();
    const int ocmin = wave_min_i32(lane < NW ? L.redi[0][lane] : INT_MAX);
    const int ocmax = wave_max_i32(lane < NW ? L.redi[1][lane] : INT_MIN);
    const int any_search = wave_or_i32(lane < NW ? L.redi[2][lane] : 0);
    constexpr int EJ = kSlideECap / NT;
    int32_t sc[EJ], sr[EJ];
    uint32_t sp[EJ];
    const bool stage_old = !any_search && E_old <= kSlideECap;
#pragma unroll
    for (int j = 0; j < EJ; ++j) {
        const int e = tid + j * NT;
        sr[j] = -1;
        if (stage_old && e < E_old) {
            const int r = run_of_entry(L.ebase, newr, e);
            const int g = L.gidx[r];
            const int k = e - L.ebase[r];
            const int32_t *cs = d.rcs + (size_t)g * (d.RL + 1) + k;
            sc[j] = d.rcell[(size_t)g * d.RL + k];
            const int a = cs[0], b = cs[1];
            sp[j] = (uint32_t)a | ((uint32_t)(b - a) << 16);
            sr[j] = r;
        }
    }

    // 1. the new points into LDS and their cell range (from the registers)
    float xmn = __builtin_inff(), xmx = -__builtin_inff(), nan = 0.f;
    auto take = [&](int l, bool ok, float x) {
        if (ok) {
            L.u.x[l] = x;
            if (x != x) {
                nan = 1.f;
            } else {
                xmn = fminf(xmn, x);
                xmx = fmaxf(xmx, x);
            }
        }
    };
#pragma unroll
    for (int k = 0; k < IT; ++k) {
        const int g = k < 4 ? tid : tid + NT;
        take(4 * g + (k & 3), g < n4, v[k]);
    }
    take(4 * n4 + tid, 4 * n4 + tid < n, vt);
    xmn = wave_min_f32(xmn);
    xmx = wave_max_f32(xmx);
    nan = wave_max_f32(nan);
    if (lane == 0) {
        L.redf[0][wid] = xmn;
        L.redf[1][wid] = xmx;
        L.redf[2][wid] = nan;
    }
    __syncthreads();
    xmn = wave_min_f32(lane < NW ? L.redf[0][lane] : __builtin_inff());
    xmx = wave_max_f32(lane < NW ? L.redf[1][lane] : -__builtin_inff());
    nan = wave_max_f32(lane < NW ? L.redf[2][lane] : 0.f);
    int c_lo = cell_coord<1>((double)xmn, d.inv_c);
    int c_hi = cell_coord<1>((double)xmx, d.inv_c);
    if (nan != 0.f)
        c_lo = -(1 << 30);
    if (c_hi < c_lo) // every point NaN (min x = +inf)
        c_hi = c_lo;
    const int cmin = min(c_lo, ocmin), cmax = max(c_hi, ocmax);
    const int64_t crange = (int64_t)c_hi - (int64_t)c_lo + 1;
    const int64_t ne_max = crange < n ? crange : n; // the new run's cells
    if (!stage_old || crange > kRunBMax || (int64_t)cmax - (int64_t)cmin + 1 > kSlideHCap ||
        (int64_t)E_old + ne_max > kSlideECap) {
        if (tid == 0 && err_in)
            atomicOr(d.err, err_in); // left for the HBM path's detect to report
        __syncthreads(); // every thread past its reads of this layout
        return false;
    }
    uint64_t *ts = stamps ? L.ts : nullptr;

    // 2. the new run bucketed in LDS (its HBM copy is written after the
    //    slide is published: only later slides read it)
    const int s = run_shift(crange);
    const int nb = (int)(crange << s);
    const int nc_new = (int)crange;
    run_bucket(d, L.u.x, n, c_lo, s, nb, L.bpre, L.sx, L.sidx, L.scan, &L.err, ts, true);
#pragma unroll
    for (int j = 0; j < EJ; ++j) {
        const int e = tid + j * NT;
        if (sr[j] >= 0) {
            L.ecell[e] = sc[j];
            L.epos[e] = sp[j];
            L.erun[e] = (uint8_t)sr[j];
        }
    }
    if (tid == 0) {
        L.meta[newr] = RunMeta{n, kRunDirect, s, nb, c_lo, c_hi, 0, 0};
        L.tests = 0;
        L.hits = 0;
        L.nrange = 0;
        L.nout = 0;
    }
    const int words = (d.W + 31) >> 5;
    for (int w = tid; w < words; w += NT)
        L.bitmap[w] = 0;

    // 3. window cell counts in LDS: the old runs' staged tables and the new
    //    run's cells straight from its bucket starts
    if (ts && tid == 0)
        ts[kTsClassify] = wall_clock();
    const int cw = cmax - cmin + 1;
    for (int c = tid; c < cw; c += NT)
        L.u.d.hist[c] = 0;
    __syncthreads();
    for (int e = tid; e < E_old; e += NT)
        atomicAdd(&L.u.d.hist[L.ecell[e] - cmin], (int)(L.epos[e] >> 16));
    for (int k = tid; k < nc_new; k += NT) {
        const int cnt = L.bpre[(k + 1) << s] - L.bpre[k << s];
        if (cnt)
            atomicAdd(&L.u.d.hist[c_lo + k - cmin], cnt);
    }
    __syncthreads();

    // 4. candidates: the points of the sparse and the clamped cells; a range
    //    is an old run's staged entry e (>= 0) or the new run's cell k (-1 - k)
    for (int e = tid; e < E_old; e += NT) {
        const int32_t c = L.ecell[e];
        if (!run_cell_guaranteed(c) || L.u.d.hist[c - cmin] <= d.K) {
            const int slot = atomicAdd(&L.nrange, 1);
            L.u.d.rent[slot] = e;
            L.u.d.rpos[slot] = (int)(L.epos[e] >> 16);
        }
    }
    for (int k = tid; k < nc_new; k += NT) {
        const int cnt = L.bpre[(k + 1) << s] - L.bpre[k << s];
        const int32_t c = c_lo + k;
        if (cnt && (!run_cell_guaranteed(c) || L.u.d.hist[c - cmin] <= d.K)) {
            const int slot = atomicAdd(&L.nrange, 1);
            L.u.d.rent[slot] = -1 - k;
            L.u.d.rpos[slot] = cnt;
        }
    }
    __syncthreads();
    const int nrg = L.nrange;
    run_lds_scan(L.u.d.rpos, nrg, L.scan);
    const int nc = L.u.d.rpos[nrg];
    if (ts && tid == 0)
        ts[kTsDetect] = wall_clock();

    // 5. exact counts by rounds of candidates: the new run from LDS, the
    //    others from HBM
    unsigned long long tests = 0, hits = 0;
    for (int c0 = 0; c0 < nc; c0 += kSlideCand) {
        const int ncs = min(kSlideCand, nc - c0);
        for (int c = tid; c < ncs; c += NT) {
            const int cc = c0 + c;
            int lo = 0, hi = nrg - 1; // the range holding candidate cc
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (L.u.d.rpos[mid] <= cc)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            const int e = L.u.d.rent[lo];
            const int off = cc - L.u.d.rpos[lo];
            if (e < 0) {
                const int idx = L.bpre[(-1 - e) << s] + off;
                L.u.d.cx[c] = L.sx[idx];
                L.u.d.cslot[c] = slot0 + (int)L.sidx[idx];
            } else {
                const size_t gi = (size_t)L.gidx[L.erun[e]] * d.RL +
                                  (size_t)((int)(L.epos[e] & 0xffffu) + off);
                L.u.d.cx[c] = d.rx[gi];
                L.u.d.cslot[c] = d.rslot[gi];
            }
            L.u.d.ccnt[c] = 0;
        }
        __syncthreads();
        const int P = ncs * nr;
        for (int p = tid; p < P; p += NT) { // run-major, as in run_detect_body
            const int r = p / ncs, c = p - r * ncs;
            const float xp = L.u.d.cx[c];
            if (!__builtin_isfinite(xp)) // NaN and +-inf have no neighbour, not even themselves
                continue;
            const int cnt = r == newr ? run_count_direct(d, L.meta[r], L.bpre, L.sx, xp, tests)
                                      : run_count(d, L.meta[r], L.gidx[r], xp, tests);
            if (cnt)
                atomicAdd(&L.u.d.ccnt[c], cnt);
        }
        __syncthreads();
        for (int c = tid; c < ncs; c += NT) {
            const int cnt = L.u.d.ccnt[c];
            hits += (unsigned long long)min(cnt, d.K + 1);
            if (cnt <= d.K)
                run_mark_outlier(L.bitmap, L.olist, &L.nout, L.u.d.cslot[c]);
        }
        __syncthreads();
    }
    run_add_counters(tests, hits, &L.tests, &L.hits);
    run_publish(d, i, L.bitmap, L.olist, &L.nout, L.scan, nc, nrg, &L.tests, &L.hits, ts,
                err_in | L.err);

    // 6. the new run into HBM for the slides after this one (the stream
    //    orders their launches after this one's end)
    run_write_direct(d, gn, n, slot0, s, nb, c_lo, c_hi, L.bpre, L.sx, L.sidx, L.scan, nullptr,
                     nullptr, nullptr, 0);
    return true;
}

// The launch roles: each kernel holds the code of its own role only.
enum RunRole { kRoleSlide, kRoleStage, kRoleSort, kRoleDetect };

// A one-slide launch whose window does not fit the workgroup's tables: the
// run sorted into HBM, then the slide detected from there.  (Out of line it
// would give the kernel a stack in scratch memory, which measured 16 us
// more per launch.)
__device__ __forceinline__ void run_slide_hbm(const RunsDev &d, const RunJob &job, uint64_t *ts,
                                           unsigned char *dyn_lds)
{
    run_sort_body(d, job.sort_q0, 0, run_job_src(d, job, 0), ts,
                  *reinterpret_cast<RunSortLds *>(dyn_lds));
    // the run's stores reach L2 before any wave of this workgroup reads
    // them back (the detect's loads go through this CU's L1)
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    run_detect_body(d, job.det_i0, ts, *reinterpret_cast<RunDetLds *>(dyn_lds));
}

// One launch: job.nsort slides' runs sorted (rps workgroups each), then
// job.ndet slides detected (one workgroup each); or, one slide per call
// (one run per slide), one workgroup that sorts slide det_i0's run and
// detects the slide (run_slide_body, else the sort into HBM and the detect
// from there).  The detects of a launch read runs sorted by earlier
// launches only.
template <int ROLE>
__device__ __forceinline__ void runs_launch_body(const RunsDev &d, const RunJob &job,
                                                 unsigned char *dyn_lds)
{
    uint64_t *ts = job.stamps ? d.tstamp : nullptr; // stamps of split launches: HBM
    if constexpr (ROLE == kRoleSlide) {
        uint64_t t_in = 0;
        if (job.stamps && threadIdx.x == 0)
            t_in = wall_clock();
        if (run_slide_body(d, job.det_i0, run_job_src(d, job, 0), job.stamps != 0, t_in,
                           *reinterpret_cast<RunSlideLds *>(dyn_lds)))
            return;
        if (ts && threadIdx.x == 0)
            ts[kTsIngest] = t_in;
        run_slide_hbm(d, job, ts, dyn_lds);
        return;
    }
    const int nsw = job.nsort * d.rps;
    int bid = blockIdx.x;
    if (ROLE != kRoleDetect && bid < nsw) {
        const int k = bid / d.rps, part = bid - k * d.rps;
        const int64_t q = job.sort_q0 + k;
        const float *src = run_job_src(d, job, k);
        if (ts && bid == 0 && threadIdx.x == 0)
            ts[kTsIngest] = wall_clock();
        run_sort_body(d, q, part, src, bid == 0 ? ts : nullptr,
                      *reinterpret_cast<RunSortLds *>(dyn_lds));
        return;
    }
    if constexpr (ROLE != kRoleSort) {
        bid -= nsw;
        run_detect_body(d, job.det_i0 + bid, ts, *reinterpret_cast<RunDetLds *>(dyn_lds));
    }
}

// The launches by role (one body, instantiated per role: each kernel's code
// holds only what its role runs):
// k_run_slide  one slide per call (fused, one workgroup)
// k_run_stage  pipelined: 32 slides detected, the next 32 sorted
// k_run_sort   the initial window's runs, or one slide's runs (two launches
//              per slide, phase API)
// k_run_detect one slide detected (two launches per slide, phase API)
__global__ __launch_bounds__(kRunNT) void k_run_slide(RunsDev d, RunJob job)
{
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    runs_launch_body<kRoleSlide>(d, job, dyn_lds);
}
__global__ __launch_bounds__(kRunNT) void k_run_stage(RunsDev d, RunJob job)
{
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    runs_launch_body<kRoleStage>(d, job, dyn_lds);
}
__global__ __launch_bounds__(kRunNT) void k_run_sort(RunsDev d, RunJob job)
{
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    runs_launch_body<kRoleSort>(d, job, dyn_lds);
}
__global__ __launch_bounds__(kRunNT) void k_run_detect(RunsDev d, RunJob job)
{
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    runs_launch_body<kRoleDetect>(d, job, dyn_lds);
}

} // namespace rtod
