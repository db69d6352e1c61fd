// rtod_plain.h -- the plain (cell-stencil) per-slide path.
//
// One slide = four kernel launches, all state in HBM, one host round trip:
//   k_ingest_sort   new slide from pinned staging into the ring; cell keys;
//                   stable 4-bit LDS radix sort of the batch by relative key
//   k_merge_pass    (only when S exceeds one LDS tile) merges sorted runs
//   k_merge_window  incremental rebuild of the cell-sorted window:
//                   window - expired batch + new batch, by rank computation;
//                   writes the sample table of the cell-stencil index
//   k_classify      per-point cell run length -> dense fast-accept or
//                   candidate (wave-aggregated compaction)
//   k_detect        range count per candidate over its stencil rows with
//                   ballot/popcount and early exit at K+1; the last
//                   workgroup compacts the outlier bitmap into an ascending
//                   slot list written straight into pinned host memory
//
// Reference counterparts: prepare_c_non_points_queue (optixScan.cpp:967-1017),
// rebuild_gas + kGenAABB (optixScan.cpp:378-426, aabb.cu:5-45), launch +
// __intersection__cube OPT2 (optixScan.cpp:727-750, optixScan.cu:158-199),
// D2H + select (optixScan.cpp:1083-1101).
#pragma once

#include <climits>

#include "rtod_device.h"
#include "rtod_learned.h"
#include "rtod_state.h"

namespace rtod {

enum SortMode { kModeSlide = 0, kModePrime = 1 };

// largest sample table of a sorted window (entries); the stride grows with W
constexpr int kMaxSamples = 4096;

// the device's constant-rate wall clock (its rate: hipDeviceAttributeWallClockRate)
__device__ __forceinline__ uint64_t wall_clock() { return __builtin_amdgcn_s_memrealtime(); }

// One slide per call: the first workgroup of each of the slide's kernels
// stamps its start (kernels on a stream run one after another, so a
// kernel's start is the previous one's end, to the launch gap); the
// epilogue copies the stamps into the host-out buffer with the list.
__device__ __forceinline__ void stamp_start(const Dev &d, int k)
{
    if (blockIdx.x == 0 && threadIdx.x == 0)
        d.tstamp[k] = wall_clock();
}


// ---------------------------------------------------------------------------
// Key fields: non-negative per-axis cell coordinates (k + 2^B), most
// significant axis first.
// ---------------------------------------------------------------------------
template <int DIM>
__device__ __forceinline__ void key_fields(uint64_t key, uint64_t (&f)[3])
{
    constexpr int F = CellBits<DIM>::FIELD;
    constexpr uint64_t M = CellBits<DIM>::FMASK;
    f[0] = f[1] = f[2] = 0;
#pragma unroll
    for (int a = 0; a < DIM; ++a)
        f[a] = (key >> (F * (DIM - 1 - a))) & M;
}

template <int DIM>
__device__ __forceinline__ uint64_t fields_key(const uint64_t (&f)[3])
{
    constexpr int F = CellBits<DIM>::FIELD;
    uint64_t k = 0;
#pragma unroll
    for (int a = 0; a < DIM; ++a)
        k |= f[a] << (F * (DIM - 1 - a));
    return k;
}

__device__ __forceinline__ int bit_length(uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; }

// ---------------------------------------------------------------------------
// Stable LSD radix sort of the block's NT*IT 64-bit words by bits
// [sh0, sh0 + nbits), 4 bits per pass, wave-level multi-split.
// Arrangement: item i of lane l in wave w is element w*64*IT + i*64 + l, in
// arrival order.  Per item, four ballots give every lane the set of lanes
// holding its digit; its rank is the running count of that digit in the
// wave (kept in lane `digit`) plus the lanes below it in the set.  Only the
// 16 x NW (digit, wave) totals go through LDS, scanned digit-major by one
// wave, then items scatter to their ranks and are reloaded wave-striped.
// Out: s_key holds the sorted order (and the registers its striping).
// ---------------------------------------------------------------------------
template <int NT, int IT>
__device__ __forceinline__ void block_radix_sort(uint64_t (&k)[IT], int sh0, int nbits,
                                                 uint64_t *s_key, int *s_hist)
{
    constexpr int NW = NT / kWave;
    static_assert(16 * NW <= 4 * kWave, "digit table scanned by one wave");
    const int lane = lane_id(), wid = threadIdx.x / kWave;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (int sh = sh0; sh < sh0 + nbits; sh += 4) {
        int cnt = 0; // lane d < 16: items with digit d so far in this wave
        uint32_t rank[IT];
        uint32_t dig[IT];
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const uint32_t dg = (uint32_t)(k[i] >> sh) & 15u;
            dig[i] = dg;
            const uint64_t b0 = __ballot(dg & 1u), b1 = __ballot(dg & 2u);
            const uint64_t b2 = __ballot(dg & 4u), b3 = __ballot(dg & 8u);
            const uint64_t m = ((dg & 1u) ? b0 : ~b0) & ((dg & 2u) ? b1 : ~b1) &
                               ((dg & 4u) ? b2 : ~b2) & ((dg & 8u) ? b3 : ~b3);
            const int before = __shfl(cnt, (int)dg, kWave);
            rank[i] = (uint32_t)(before + __popcll(m & lt));
            const uint64_t md = ((lane & 1) ? b0 : ~b0) & ((lane & 2) ? b1 : ~b1) &
                                ((lane & 4) ? b2 : ~b2) & ((lane & 8) ? b3 : ~b3);
            cnt += __popcll(md);
        }
        if (lane < 16)
            s_hist[lane * NW + wid] = cnt;
        __syncthreads();
        if (wid == 0) { // exclusive scan of the 16*NW totals, 4 per lane
            int x[4], sum = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int e = lane * 4 + q;
                x[q] = e < 16 * NW ? s_hist[e] : 0;
                sum += x[q];
            }
            const int incl = wave_incl_scan(sum);
            int run = incl - sum;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int e = lane * 4 + q;
                if (e < 16 * NW)
                    s_hist[e] = run;
                run += x[q];
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < IT; ++i)
            s_key[s_hist[dig[i] * NW + wid] + (int)rank[i]] = k[i];
        __syncthreads();
#pragma unroll
        for (int i = 0; i < IT; ++i)
            k[i] = s_key[wid * kWave * IT + i * kWave + lane];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// k_ingest_sort: one workgroup of NT threads per LDS tile of CH = NT*IT
// points.  Slide mode: the new slide comes from st.src -- the pinned staging
// buffer or a caller's HBM buffer -- in one round of coalesced 16-byte
// loads and goes to ring slots st.u*S + j.  Prime mode: segment b of the initial window (slots b*S + j)
// comes from a device buffer.  The batch is sorted by (cell key, arrival):
// keys are re-expressed relative to the tile's per-axis minimum and packed
// into exactly the bits their ranges need, so a slide spanning a few dozen
// cells sorts in two 4-bit passes.
// Output: keys, slots, coordinates in sorted order (batch entry, or a
// scratch run when the batch spans several tiles) + the batch sample table.
// ---------------------------------------------------------------------------
template <int DIM, int NT, int IT> struct IngestLds {
    alignas(16) float coord[NT * IT * DIM];
    uint64_t key[NT * IT];
    uint16_t val[NT * IT];
    int hist[16 * (NT / kWave)];
    uint64_t red[2 * 3][NT / kWave];
};

template <int DIM, int NT, int IT>
__device__ __forceinline__ void ingest_sort_body(const Dev &d, const Step &st, int mode,
                                                 const float *src_dev, int dst_batch, int bid,
                                                 IngestLds<DIM, NT, IT> &lds)
{
    constexpr int CH = NT * IT, NW = NT / kWave;
    constexpr int IB = 13; // bits of the tile-local index (CH <= 8192)
    static_assert(CH <= (1 << IB), "tile index bits");
    float *s_coord = lds.coord;
    uint64_t *s_key = lds.key;
    uint16_t *s_val = lds.val;
    int *s_hist = lds.hist;
    auto &s_red = lds.red;

    const int tid = threadIdx.x, lane = lane_id(), wid = tid / kWave;
    const int seg = bid / d.cpb;
    const int chunk = bid % d.cpb;
    const int slot0 = (mode == kModeSlide) ? st.u * d.S : seg * d.S;
    const float *src = (mode == kModeSlide) ? st.src : src_dev + (size_t)seg * d.S * DIM;
    const int j0 = chunk * CH;
    const int n = min(CH, d.S - j0);

    // stage 1: the tile's n*DIM floats into LDS, one round of loads
    {
        const float *g = src + (size_t)j0 * DIM;
        const int nf = n * DIM;
        if ((((uintptr_t)g) & 15) == 0) {
            const int n4 = nf >> 2;
            copy_to_lds<(CH * DIM + 4 * NT - 1) / (4 * NT)>(reinterpret_cast<float4 *>(s_coord),
                                                            reinterpret_cast<const float4 *>(g),
                                                            n4);
            for (int q = (n4 << 2) + tid; q < nf; q += NT)
                s_coord[q] = g[q];
        } else {
            copy_to_lds<(CH * DIM + NT - 1) / NT>(s_coord, g, nf);
        }
    }
    __syncthreads();

    // wave-striped items: ring update, keys, per-axis min/max of key fields
    uint64_t k[IT];
    uint64_t fmin[3] = {~0ull, ~0ull, ~0ull}, fmax[3] = {0, 0, 0};
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int l = wid * kWave * IT + i * kWave + lane;
        k[i] = ~0ull;
        if (l < n) {
            float p[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int a = 0; a < DIM; ++a) {
                p[a] = s_coord[l * DIM + a];
                d.ring[(size_t)a * d.W + slot0 + j0 + l] = p[a];
            }
            k[i] = cell_key<DIM>(p, d.inv_c);
            uint64_t f[3];
            key_fields<DIM>(k[i], f);
#pragma unroll
            for (int a = 0; a < DIM; ++a) {
                fmin[a] = f[a] < fmin[a] ? f[a] : fmin[a];
                fmax[a] = f[a] > fmax[a] ? f[a] : fmax[a];
            }
        }
    }
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
        for (int dd = 1; dd < kWave; dd <<= 1) {
            const uint64_t y0 = __shfl_xor(fmin[a], dd, kWave);
            const uint64_t y1 = __shfl_xor(fmax[a], dd, kWave);
            fmin[a] = y0 < fmin[a] ? y0 : fmin[a];
            fmax[a] = y1 > fmax[a] ? y1 : fmax[a];
        }
        if (lane == 0) {
            s_red[2 * a][wid] = fmin[a];
            s_red[2 * a + 1][wid] = fmax[a];
        }
    }
    __syncthreads();
    int w[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
        uint64_t mn = ~0ull, mx = 0;
#pragma unroll
        for (int q = 0; q < NW; ++q) {
            mn = s_red[2 * a][q] < mn ? s_red[2 * a][q] : mn;
            mx = s_red[2 * a + 1][q] > mx ? s_red[2 * a + 1][q] : mx;
        }
        fmin[a] = mn;
        w[a] = bit_length(mx - mn);
    }
    const int nbits = w[0] + w[1] + w[2];
    // Sort words (relative key bits << IB) | tile index.  Keys up to 51 bits
    // sort in one stage.  Wider keys (cells spread beyond 2^51, possible
    // only in 2-D/3-D with clamped extremes) take two stable LSD stages: the
    // low 51 bits first, then the high bits, recomputed from the
    // coordinates, with the stage-1 order as the index.
    constexpr int KB = 64 - IB;
    constexpr uint64_t IMASK = (1ull << IB) - 1ull;
    const bool packed = nbits <= KB;
    auto rel_of = [&](const float *pc) -> uint64_t {
        const float p3[3] = {pc[0], DIM > 1 ? pc[1] : 0.f, DIM > 2 ? pc[2] : 0.f};
        uint64_t f[3];
        key_fields<DIM>(cell_key<DIM>(p3, d.inv_c), f);
        uint64_t rel = 0;
#pragma unroll
        for (int a = 0; a < DIM; ++a)
            rel = (rel << w[a]) | (f[a] - fmin[a]);
        return rel;
    };
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int l = wid * kWave * IT + i * kWave + lane;
        uint64_t rel = ~0ull; // padding: all ones, sorts last
        if (l < n) {
            uint64_t f[3];
            key_fields<DIM>(k[i], f);
            rel = 0;
#pragma unroll
            for (int a = 0; a < DIM; ++a)
                rel = (rel << w[a]) | (f[a] - fmin[a]);
        }
        k[i] = ((packed ? rel : (rel & ((1ull << KB) - 1ull))) << IB) | (uint64_t)l;
        s_key[l] = k[i]; // the arrival order stands when no bit varies
    }
    __syncthreads();
    block_radix_sort<NT, IT>(k, IB, packed ? nbits : KB, s_key, s_hist);
    if (!packed) {
        // stage-1 order p -> tile index, then sort by the high key bits
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const int e = tid + i * NT;
            s_val[e] = (uint16_t)(s_key[e] & IMASK);
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const int p = wid * kWave * IT + i * kWave + lane;
            const int l = s_val[p];
            const uint64_t hi = (l < n) ? (rel_of(&s_coord[l * DIM]) >> KB) : ~0ull;
            k[i] = (hi << IB) | (uint64_t)p;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < IT; ++i)
            s_key[wid * kWave * IT + i * kWave + lane] = k[i];
        __syncthreads();
        block_radix_sort<NT, IT>(k, IB, nbits - KB, s_key, s_hist);
    }

    // output in sorted order, striped (coalesced)
    const size_t stride_b = (size_t)d.nent * d.S;
    const int e = (mode == kModeSlide) ? st.enew : seg;
    uint64_t *ok;
    int32_t *os;
    float *ox;
    size_t xs;
    if (dst_batch) {
        ok = d.bkey + (size_t)e * d.S;
        os = d.bslot + (size_t)e * d.S;
        ox = d.bxyz + (size_t)e * d.S;
        xs = stride_b;
    } else {
        const size_t toff = (size_t)seg * d.S + j0 +
                            ((mode == kModeSlide && d.tile_par) ? (size_t)(st.t & 1) * d.S : 0);
        ok = d.tkey[0] + toff;
        os = d.tslot[0] + toff;
        ox = d.txyz[0] + toff;
        xs = (size_t)d.tmpn;
    }
    const int bmask = (1 << d.bs_log2) - 1;
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int pos = tid + i * NT;
        if (pos >= n)
            break;
        const uint64_t word = s_key[pos];
        int l;
        uint64_t key;
        if (packed) {
            l = (int)(word & IMASK);
            const uint64_t rel = word >> IB;
            uint64_t f[3] = {0, 0, 0};
            int sh = 0;
#pragma unroll
            for (int a = DIM - 1; a >= 0; --a) {
                f[a] = fmin[a] + ((rel >> sh) & ((1ull << w[a]) - 1ull));
                sh += w[a];
            }
            key = fields_key<DIM>(f);
        } else {
            l = s_val[word & IMASK];
            const float p3[3] = {s_coord[l * DIM], DIM > 1 ? s_coord[l * DIM + 1] : 0.f,
                                 DIM > 2 ? s_coord[l * DIM + 2] : 0.f};
            key = cell_key<DIM>(p3, d.inv_c);
        }
        ok[pos] = key;
        os[pos] = slot0 + j0 + l;
#pragma unroll
        for (int a = 0; a < DIM; ++a)
            ox[(size_t)a * xs + pos] = s_coord[l * DIM + a];
        if (dst_batch && (pos & bmask) == 0) // a batch is one tile: j0 == 0
            d.bsamp[(size_t)e * d.nbs + (pos >> d.bs_log2)] = key;
    }
}

template <int DIM, int NT, int IT>
__global__ __launch_bounds__(NT) void k_ingest_sort(Dev d, Step st, int mode, const float *src_dev,
                                                    int dst_batch)
{
    __shared__ IngestLds<DIM, NT, IT> lds;
    if (mode == kModeSlide)
        stamp_start(d, kTsIngest);
    ingest_sort_body<DIM, NT, IT>(d, st, mode, src_dev, dst_batch, blockIdx.x, lds);
}

// ---------------------------------------------------------------------------
// k_merge_pass: within each segment of seg_len elements, merge sorted runs of
// length L pairwise into runs of 2L (stable: the earlier run wins ties).
// Carries slots, coordinates (SoA planes, strides sxs/dxs) and optionally
// batch positions.  dst_batch: segment s goes to batch-ring entry
// (slide: st.enew, prime: s) and fills its sample table.
// ---------------------------------------------------------------------------
template <int DIM>
__device__ __forceinline__ void merge_pass_body(const Dev &d, const Step &st, const uint64_t *sk,
                                                const int32_t *ss, const float *sx, size_t sxs,
                                                const int32_t *sp, uint64_t *dk, int32_t *ds,
                                                float *dx, size_t dxs, int32_t *dp, int n_total,
                                                int seg_len, int L, int dst_batch, int mode,
                                                int bid)
{
    const int i = bid * 256 + threadIdx.x;
    if (i >= n_total)
        return;
    const int seg = i / seg_len;
    const int base = seg * seg_len;
    const int len = min(seg_len, n_total - base);
    const int li = i - base;
    const int r = li / L;
    const int pr = r ^ 1;
    const uint64_t key = sk[i];
    int out;
    if (pr * L >= len) {
        out = li;
    } else {
        const int p0 = base + pr * L, p1 = base + min(pr * L + L, len);
        const int rank = (r & 1) ? upper_bound_u64(sk, p0, p1, key) - p0
                                 : lower_bound_u64(sk, p0, p1, key) - p0;
        out = min(r, pr) * L + (li - r * L) + rank;
    }
    size_t o;
    if (dst_batch) {
        const int e = (mode == kModeSlide) ? st.enew : seg;
        o = (size_t)e * d.S + out;
        if ((out & ((1 << d.bs_log2) - 1)) == 0)
            d.bsamp[(size_t)e * d.nbs + (out >> d.bs_log2)] = key;
    } else {
        o = (size_t)base + out;
    }
    dk[o] = key;
    ds[o] = ss[i];
#pragma unroll
    for (int a = 0; a < DIM; ++a)
        dx[(size_t)a * dxs + o] = sx[(size_t)a * sxs + i];
    if (dp) // batch position: from the source, or the index inside its batch
        dp[o] = sp ? sp[i] : (i % d.S);
}

template <int DIM>
__global__ __launch_bounds__(256) void k_merge_pass(Dev d, Step st, const uint64_t *sk,
                                                    const int32_t *ss, const float *sx, size_t sxs,
                                                    const int32_t *sp, uint64_t *dk, int32_t *ds,
                                                    float *dx, size_t dxs, int32_t *dp,
                                                    int n_total, int seg_len, int L, int dst_batch,
                                                    int mode)
{
    merge_pass_body<DIM>(d, st, sk, ss, sx, sxs, sp, dk, ds, dx, dxs, dp, n_total, seg_len, L,
                         dst_batch, mode, blockIdx.x);
}

// ---------------------------------------------------------------------------
// k_merge_runs: one launch turning each batch segment's tile-sorted runs
// (length L = d.run_len, 1024 or 2048, the last one shorter; at most kRunMax
// runs) into the sorted batch: ring entry and sample table.  The segment's
// keys (S <= kRunMax * L <= kRunMergeCap) are staged in LDS.  Element i of run r lands at
//     (i - r*L) + sum_{r' < r} #(run r' <= key) + sum_{r' > r} #(run r' < key)
// which is stable: an earlier run holds earlier arrivals.  A thread places
// kRunMergeIT elements (strided by the workgroup size; one, since the
// search is bound by a CU's LDS bandwidth and more workgroups spread it);
// an element's counts
// in all the other runs are found together, bit by bit from L down (the
// count is the largest c whose element c-1 passes), so each of the
// log2(L)+1 steps issues up to kRunMax independent LDS reads per element.
// Slots and coordinates are loaded before the staging barrier.
// ---------------------------------------------------------------------------
constexpr int kRunMergeNT = 512, kRunMergeIT = 1, kRunLen = 1024, kRunMax = 8;
constexpr int kRunMergeCap = kRunMax * 2 * kRunLen; // 2048-point runs
constexpr int kRunMergeT = kRunMergeNT * kRunMergeIT; // elements per workgroup

// workgroups of k_merge_runs per batch segment of S points
__host__ __device__ constexpr int run_merge_wgs(int S) { return (S + kRunMergeT - 1) / kRunMergeT; }

template <int DIM, int L>
__device__ __forceinline__ void merge_runs_body_t(const Dev &d, const Step &st, int mode, int bid,
                                                  uint64_t *s_k)
{
    constexpr int IT = kRunMergeIT;
    const int S = d.S;
    const int wps = run_merge_wgs(S);
    const int seg = bid / wps, w = bid % wps;
    const size_t sxs = (size_t)d.tmpn;
    const size_t toff = (size_t)seg * S +
                        ((mode == kModeSlide && d.tile_par) ? (size_t)(st.t & 1) * S : 0);
    const uint64_t *sk = d.tkey[0] + toff;
    int32_t sl[IT];
    float cx[IT][3];
#pragma unroll
    for (int k = 0; k < IT; ++k) {
        const int i = w * kRunMergeT + k * kRunMergeNT + threadIdx.x;
        sl[k] = 0;
        cx[k][0] = cx[k][1] = cx[k][2] = 0.f;
        if (i < S) {
            sl[k] = d.tslot[0][toff + i];
#pragma unroll
            for (int a = 0; a < DIM; ++a)
                cx[k][a] = d.txyz[0][(size_t)a * sxs + toff + i];
        }
    }
    copy_to_lds<kRunMax * kRunLen / kRunMergeNT>(s_k, sk, S);
    __syncthreads();
    const int nruns = (S + L - 1) / L;
    const int e = (mode == kModeSlide) ? st.enew : seg;
    const size_t bxs = (size_t)d.nent * S;
    // the elements' searches advance together: IT * kRunMax independent
    // LDS reads per step
    uint64_t key[IT];
    int r[IT], c[IT][kRunMax], len[IT][kRunMax];
#pragma unroll
    for (int k = 0; k < IT; ++k) {
        const int i = w * kRunMergeT + k * kRunMergeNT + threadIdx.x;
        const bool on = i < S;
        key[k] = on ? s_k[i] : 0ull;
        r[k] = on ? i / L : 0;
#pragma unroll
        for (int q = 0; q < kRunMax; ++q) {
            c[k][q] = 0;
            len[k][q] = (on && q < nruns && q != r[k]) ? min(L, S - q * L) : 0;
        }
    }
#pragma unroll
    for (int b = L; b >= 1; b >>= 1) {
        uint64_t x[IT][kRunMax];
#pragma unroll
        for (int k = 0; k < IT; ++k)
#pragma unroll
            for (int q = 0; q < kRunMax; ++q)
                x[k][q] = (c[k][q] + b <= len[k][q]) ? s_k[q * L + c[k][q] + b - 1] : ~0ull;
#pragma unroll
        for (int k = 0; k < IT; ++k)
#pragma unroll
            for (int q = 0; q < kRunMax; ++q) {
                const bool pass = (c[k][q] + b <= len[k][q]) &&
                                  (q < r[k] ? x[k][q] <= key[k] : x[k][q] < key[k]);
                c[k][q] += pass ? b : 0;
            }
    }
#pragma unroll
    for (int k = 0; k < IT; ++k) {
        const int i = w * kRunMergeT + k * kRunMergeNT + threadIdx.x;
        if (i >= S)
            break;
        int o = i - r[k] * L;
#pragma unroll
        for (int q = 0; q < kRunMax; ++q)
            o += c[k][q];
        const size_t ob = (size_t)e * S + o;
        d.bkey[ob] = key[k];
        d.bslot[ob] = sl[k];
#pragma unroll
        for (int a = 0; a < DIM; ++a)
            d.bxyz[(size_t)a * bxs + ob] = cx[k][a];
        if ((o & ((1 << d.bs_log2) - 1)) == 0)
            d.bsamp[(size_t)e * d.nbs + (o >> d.bs_log2)] = key[k];
    }
}

// the context's run length (uniform): both searches fully unrolled
template <int DIM>
__device__ __forceinline__ void merge_runs_body(const Dev &d, const Step &st, int mode, int bid,
                                                uint64_t *s_k)
{
    if (d.run_len == 2 * kRunLen)
        merge_runs_body_t<DIM, 2 * kRunLen>(d, st, mode, bid, s_k);
    else
        merge_runs_body_t<DIM, kRunLen>(d, st, mode, bid, s_k);
}

template <int DIM>
__global__ __launch_bounds__(kRunMergeNT) void k_merge_runs(Dev d, Step st, int mode)
{
    __shared__ uint64_t s_k[kRunMergeCap];
    merge_runs_body<DIM>(d, st, mode, blockIdx.x, s_k);
}

// sample table of a freshly built sorted window
__global__ __launch_bounds__(256) void k_window_samples(Dev d, int par)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < d.nsamp)
        d.samples[par][s] = d.akey[par][(size_t)s << d.stride_log2];
    if (s < d.nsamp4)
        d.samples4[par][s] = d.akey[par][(size_t)s << 2];
}

// Wave-cooperative bound over a sorted table in LDS (n <= 4096 entries): a
// 64-ary search, every lane probing one position per round.  Returns the
// number of entries < key (lower) or <= key (upper), in every lane.
__device__ __forceinline__ int wave_count_lds(const uint64_t *tab, int n, uint64_t key, bool upper)
{
    const int lane = lane_id();
    int lo = 0, len = n; // answer in [lo, lo + len]
    while (len > 0) {
        const int step = (len + kWave - 1) / kWave; // probes lane*step
        const int pos = lo + lane * step;
        bool below = false;
        if (lane * step < len) {
            const uint64_t x = tab[pos];
            below = upper ? (x <= key) : (x < key);
        }
        const int c = __popcll(__ballot(below)); // probes that are below
        if (c == 0)
            return lo;
        // the answer lies after probe c-1 and at or before probe c
        lo = lo + (c - 1) * step + 1;
        len = min(step - 1, n - lo);
        if (step == 1)
            return lo;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// k_merge_window: A' = (A - E) + B, all sorted by (key, arrival).
//   A  current sorted window, E the expiring batch (a subset of A), B the new
//   batch, both batches sorted by key with arrival order on ties.
//   survivor a at index i:  pos = i - #E before a + #B with key < key_a
//   b at index j of B:      pos = j + #A with key <= k_b - #E with key <= k_b
// Workgroup c merges the A chunk [i0, i1) (chunk-boundary searches on the
// batch sample tables, the B keys the chunk touches staged in LDS) and
// places the new points whose keys fall in its key range
// [key(A[i0]), key(A[i1])) (the first chunk from the start of B, the last
// to its end): all of A before the chunk has keys <= key(A[i0]) and all of A
// after it keys >= key(A[i1]), so #A <= k_b is i0 plus a search of the
// chunk's keys in LDS, and #E <= k_b is #E before the chunk plus the chunk's
// expired points among those; no search leaves LDS.  #E before A[i0] is
// A[i0]'s own batch position when A[i0] expires.  Output positions that are
// multiples of the sample stride fill A''s sample table.
// ---------------------------------------------------------------------------
constexpr int kMergeT = 512; // window elements per merge workgroup (half the chunk
                              // of 1024 measured 8% faster on the 1-D headline)
constexpr int kMergeRB = 4;   // new points per thread loaded ahead

template <int NT> struct MergeWindowLds {
    static constexpr int IT = kMergeT / NT, T = kMergeT, CAP = 2048, MAXBS = 512;
    uint64_t k[T];
    uint64_t bs[MAXBS], es[MAXBS];
    alignas(16) uint64_t br[CAP];
    uint64_t exm[T / kWave]; // expired bits of the chunk, 64 positions a word
    int rc[IT][NT / kWave];
    int res[8];
    uint64_t kn;
    int pos0, slot0;
};

// NT threads, kMergeT / NT chunk elements each; workgroup 0 resets the
// slide's classify counters
template <int DIM, int NT>
__device__ __forceinline__ void merge_window_body(const Dev &d, const Step &st, int bid,
                                                  MergeWindowLds<NT> &lds)
{
    constexpr int IT = kMergeT / NT, T = kMergeT, CAP = 2048;
    static_assert(IT * NT == T, "merge tile");
    uint64_t *s_k = lds.k;
    uint64_t *s_bs = lds.bs, *s_es = lds.es;
    uint64_t *s_br = lds.br;
    uint64_t *s_exm = lds.exm;
    auto &s_rc = lds.rc;
    int *s_res = lds.res;
    uint64_t &s_kn = lds.kn;
    int &s_pos0 = lds.pos0, &s_slot0 = lds.slot0;

    const int tid = threadIdx.x, lane = lane_id(), wid = tid / kWave;
    const int W = d.W, S = d.S;
    const int nmb = (W + T - 1) / T; // merge workgroups
    const int par = st.par;
    const uint64_t *Ak = d.akey[par];
    const int32_t *As = d.aslot[par];
    const int32_t *Ap = d.apos[par];
    const float *Ax = d.axyz[par];
    uint64_t *Nk = d.akey[par ^ 1];
    int32_t *Ns = d.aslot[par ^ 1];
    int32_t *Np = d.apos[par ^ 1];
    float *Nx = d.axyz[par ^ 1];
    uint64_t *Nsamp = d.samples[par ^ 1];
    uint64_t *Nsamp4 = d.samples4[par ^ 1];
    const uint64_t *Ek = d.bkey + (size_t)st.eold * S;
    const uint64_t *Bk = d.bkey + (size_t)st.enew * S;
    const int32_t *Bs = d.bslot + (size_t)st.enew * S;
    const float *Bx = d.bxyz + (size_t)st.enew * S;
    const size_t bxs = (size_t)d.nent * S;
    const uint64_t *Esamp = d.bsamp + (size_t)st.eold * d.nbs;
    const uint64_t *Bsamp = d.bsamp + (size_t)st.enew * d.nbs;
    const int exp_lo = st.u * S;
    const int i0 = bid * T;
    const int i1 = min(W, i0 + T);
    const int n = i1 - i0;
    const bool first = bid == 0;
    const bool last = i1 == W;
    const int sl2 = d.stride_log2, smask = (1 << sl2) - 1;
    const int bsl = d.bs_log2, nbs = d.nbs;
    (void)nmb;

    if (bid == 0 && tid == 0) {
        SlideCnt *cn = cnt_of(d, st.t);
        cn->n_cand = 0;
        cn->n_units = 0;
        cn->n_units_i = 0;
        cn->n_done = 0;
        cn->pair_tests = 0;
        cn->hits = 0;
        cn->misses = 0;
        cn->fmin[0] = cn->fmin[1] = cn->fmin[2] = ~0ull;
    }
    if (d.ready && bid == 0 && tid < kReadyFlags)
        *ready_of(d, st.t, tid) = 0;
    // this slide's outlier bitmap, last read by the epilogue of the slide
    // kSlideSets earlier: cleared here, T / 32 words per workgroup
    {
        static_assert(T / 32 <= NT, "one word per thread");
        const int wi = bid * (T / 32) + tid;
        if (tid < T / 32 && wi < d.bm_words)
            bitmap_of(d, st.t)[wi] = 0u;
    }

    // round 1: the chunk (keys, slots, batch positions, coordinates; striped,
    // item i of thread t is chunk element i*NT + t) and the sample tables of
    // B and E
    int32_t sl[IT], ps[IT];
    float cx[IT][3];
    uint64_t kk[IT];
    bool ex[IT];
    int below[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int l = i * NT + tid;
        ex[i] = false;
        sl[i] = 0;
        ps[i] = 0;
        kk[i] = 0;
        cx[i][0] = cx[i][1] = cx[i][2] = 0.f;
        if (l < n) {
            kk[i] = Ak[i0 + l];
            sl[i] = As[i0 + l];
            ps[i] = Ap[i0 + l];
#pragma unroll
            for (int a = 0; a < DIM; ++a)
                cx[i][a] = Ax[(size_t)a * W + i0 + l];
            ex[i] = sl[i] >= exp_lo && sl[i] < exp_lo + S;
        }
    }
    copy_to_lds<2>(s_bs, Bsamp, nbs);
    copy_to_lds<2>(s_es, Esamp, nbs);
    if (tid == 0) {
        s_kn = last ? 0 : Ak[i1];
        s_pos0 = Ap[i0];
        s_slot0 = As[i0];
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int l = i * NT + tid;
        if (l < n)
            s_k[l] = kk[i];
        const uint64_t m = __ballot(ex[i]);
        below[i] = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) {
            s_rc[i][wid] = __popcll(m);
            s_exm[((i * NT) >> 6) + wid] = m; // positions i*NT + 64*wid + lane
        }
    }
    __syncthreads();

    // four chunk-boundary searches on the batch sample tables, one wave each:
    //   0: #E <= key(A[i0])  1: lb(B, key(A[i0]))  2: lb(B, key(A[i1-1]))
    //   3: lb(B, key(A[i1]))
    const uint64_t k0 = s_k[0], kl = s_k[n - 1], kn = s_kn;
    const int bstride = 1 << bsl;
    static_assert(NT / kWave >= 4, "a wave per search");
    if (wid < 4) {
        const int q = wid;
        const bool onE = q == 0;
        const bool upper = q == 0;
        const uint64_t key = (q == 2) ? kl : (q == 3) ? kn : k0;
        const int c = wave_count_lds(onE ? s_es : s_bs, nbs, key, upper);
        // finish inside one stride of the batch: lanes probe it in one round
        int res = 0;
        if (c > 0) {
            const int L = ((c - 1) << bsl) + 1, H = min(c << bsl, S);
            int cnt = 0;
            for (int r0 = 0; r0 < bstride - 1; r0 += kWave) {
                const int idx = L + r0 + lane;
                bool hit = false;
                if (r0 + lane < bstride - 1 && idx < H) {
                    const uint64_t x = onE ? Ek[idx] : Bk[idx];
                    hit = upper ? (x <= key) : (x < key);
                }
                cnt += __popcll(__ballot(hit));
            }
            res = L + cnt;
        }
        if (lane == 0)
            s_res[q] = res;
    }
    __syncthreads();
    const bool ex0 = s_slot0 >= exp_lo && s_slot0 < exp_lo + S;
    const int e_before = ex0 ? s_pos0 : s_res[0];
    const int b_lo = s_res[1], b_hi = s_res[2];
    // the new points this chunk places: B[jb0, jb1), keys in the chunk's
    // key range; the first kMergeRB per thread (keys, slots, coordinates)
    // requested with the staging of the survivors' B range [b_lo, b_hi) (a
    // slide's new points can crowd into a few chunks' key ranges: a random
    // walk's slide spans a narrow band of values)
    const int jb0 = first ? 0 : b_lo, jb1 = last ? S : s_res[3];
    const int nj = jb1 - jb0;
    uint64_t bk_[kMergeRB];
    int32_t bsl_[kMergeRB];
    float bx[kMergeRB][3];
#pragma unroll
    for (int r = 0; r < kMergeRB; ++r) {
        const int jj = tid + r * NT;
        bk_[r] = 0;
        bsl_[r] = 0;
        bx[r][0] = bx[r][1] = bx[r][2] = 0.f;
        if (jj < nj) {
            bk_[r] = Bk[jb0 + jj];
            bsl_[r] = Bs[jb0 + jj];
#pragma unroll
            for (int a = 0; a < DIM; ++a)
                bx[r][a] = Bx[(size_t)a * bxs + jb0 + jj];
        }
    }
    const int nb = b_hi - b_lo;
    const bool bfit = nb <= CAP;
    if (bfit)
        copy_to_lds<8>(s_br, Bk + b_lo, nb);
    __syncthreads();

    // B rank of each of the thread's items (#B < key), the IT searches of the
    // staged range advancing together bit by bit (independent LDS reads)
    int rbl[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i)
        rbl[i] = 0;
    if (bfit) {
        for (int b = nb > 0 ? 1 << (31 - __builtin_clz(nb)) : 0; b >= 1; b >>= 1) {
            uint64_t x[IT];
#pragma unroll
            for (int i = 0; i < IT; ++i)
                x[i] = (rbl[i] + b <= nb) ? s_br[rbl[i] + b - 1] : ~0ull;
#pragma unroll
            for (int i = 0; i < IT; ++i)
                rbl[i] += (rbl[i] + b <= nb && x[i] < kk[i]) ? b : 0;
        }
    }

    // survivors of A; #E before element l = e_before + expired elements of
    // the chunk in earlier rows, earlier waves of this row, earlier lanes
    int row_base = 0;
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        int in_row = 0, row_tot = 0;
#pragma unroll
        for (int q = 0; q < NT / kWave; ++q) {
            in_row += (q < wid) ? s_rc[i][q] : 0;
            row_tot += s_rc[i][q];
        }
        const int eb = e_before + row_base + in_row + below[i];
        row_base += row_tot;
        const int l = i * NT + tid;
        if (l >= n || ex[i])
            continue;
        const uint64_t key = kk[i];
        const int rb = bfit ? b_lo + rbl[i] : lower_bound_u64(Bk, b_lo, b_hi, key);
        const int np = i0 + l - eb + rb;
        if ((unsigned)np >= (unsigned)W) { // impossible unless state is corrupt
            atomicOr(&d.ctl->err, kErrMergeSurvivor);
            continue;
        }
        Nk[np] = key;
        Ns[np] = sl[i];
        Np[np] = ps[i];
#pragma unroll
        for (int a = 0; a < DIM; ++a)
            Nx[(size_t)a * W + np] = cx[i][a];
        if ((np & smask) == 0)
            Nsamp[np >> sl2] = key;
        if (Nsamp4 && (np & 3) == 0)
            Nsamp4[np >> 2] = key;
    }

    // the new points B[jb0, jb1): #A <= k_b = i0 + the chunk's keys <= k_b
    // (LDS), #E <= k_b = #E before the chunk + its expired points among those
    // a thread's keys ascend over its rounds and repeat in long runs (a 1-D
    // slide covers few cells): the ranks of its last key are kept
    bool have = false;
    uint64_t last_key = 0;
    int ca = 0, ce = 0;
    auto place_new = [&](int j, uint64_t key, int32_t sb, const float (&px)[3]) {
        if (!have || key != last_key) {
            ca = upper_bound_u64(s_k, 0, n, key);
            ce = 0;
            for (int w = 0; w < (ca >> 6); ++w)
                ce += __popcll(s_exm[w]);
            if (ca & 63)
                ce += __popcll(s_exm[ca >> 6] & ((1ull << (ca & 63)) - 1ull));
            have = true;
            last_key = key;
        }
        const int np = j + i0 + ca - (e_before + ce);
        if ((unsigned)np >= (unsigned)W) {
            atomicOr(&d.ctl->err, kErrMergeNew);
            return;
        }
        Nk[np] = key;
        Ns[np] = sb;
        Np[np] = j;
#pragma unroll
        for (int a = 0; a < DIM; ++a)
            Nx[(size_t)a * W + np] = px[a];
        if ((np & smask) == 0)
            Nsamp[np >> sl2] = key;
        if (Nsamp4 && (np & 3) == 0)
            Nsamp4[np >> 2] = key;
    };
#pragma unroll
    for (int r = 0; r < kMergeRB; ++r) {
        const int j = jb0 + tid + r * NT;
        if (j >= jb1)
            break;
        place_new(j, bk_[r], bsl_[r], bx[r]);
    }
    // more than kMergeRB * NT new points in this chunk's range: each round's
    // loads are issued before the previous round's stores
    int j = jb0 + tid + kMergeRB * NT;
    uint64_t key = 0;
    int32_t sb = 0;
    float px[3] = {0.f, 0.f, 0.f};
    auto load_new = [&](int jj, uint64_t &k, int32_t &b, float (&x)[3]) {
        k = Bk[jj];
        b = Bs[jj];
#pragma unroll
        for (int a = 0; a < DIM; ++a)
            x[a] = Bx[(size_t)a * bxs + jj];
    };
    if (j < jb1)
        load_new(j, key, sb, px);
    while (j < jb1) {
        const int jn = j + NT;
        uint64_t kn2 = 0;
        int32_t sn = 0;
        float pn[3] = {0.f, 0.f, 0.f};
        if (jn < jb1)
            load_new(jn, kn2, sn, pn);
        place_new(j, key, sb, px);
        j = jn;
        key = kn2;
        sb = sn;
#pragma unroll
        for (int a = 0; a < 3; ++a)
            px[a] = pn[a];
    }
}

template <int DIM>
__global__ __launch_bounds__(kMergeT) void k_merge_window(Dev d, Step st)
{
    __shared__ MergeWindowLds<kMergeT> lds;
    stamp_start(d, kTsMerge);
    merge_window_body<DIM, kMergeT>(d, st, blockIdx.x, lds);
}

// ---------------------------------------------------------------------------
// k_classify: a point is fast-accepted iff its cell is unclamped and holds
// >= K+1 points (reference: cells with num > K are skipped, optixScan.cpp:
// 987-993).  The cell is a run of equal keys in the sorted window: find the
// run start s within [i-K, i] and test key[s+K] == key.  The keys around the
// tile sit in LDS when K is small.  The other points -- all points of the
// sparse cells -- are the candidates.
//   kClsPoints  appends every candidate to the per-point list (LBVH index).
//   kClsUnits   appends cell units: a candidate starts a unit where its cell
//               run starts or at a multiple of d.unit_max; the unit runs to
//               the end of the cell or of that aligned block, so a unit is at
//               most d.unit_max candidates sharing one cell.  One atomic
//               per counter per workgroup.
//   kClsLearned kClsUnits plus, per unit, the predictor's features (own-cell
//               count, the two neighbouring cells' counts along the last
//               axis, the unit's first point's offsets from the cell centre)
//               and its prediction on the matrix cores: probable-outlier
//               units are filed at the front of the list (n_units),
//               probable-inlier units at the back (index W-1-k,
//               n_units_i); features go to the slide's replay slot under
//               the unit's list index.
// ---------------------------------------------------------------------------
constexpr int kClassifyNT = 256;
constexpr int kClassifyHalo = 512;
enum ClassifyMode { kClsPoints = 0, kClsUnits = 1, kClsLearned = 2 };
// Candidates per cell unit (a power of two, at most 16: the unit reductions
// span lanes 0..15).  Smaller units put more wavefronts on a slide's
// candidates; a unit's stencil rows are shared by fewer candidates.
// Measured per slide: 1-D best at 4 (walk W=100k: 17.5 us at 16, 16.3 at 8,
// 16.0 at 4), 3-D at 8 (gmm W=10k: 29.2, 26.3, 27.5; W=100k: 128.0, 126.2,
// 145.8).
// A batch's streams, whose detect is throughput-bound (many units per
// wavefront), keep units of 16: smaller ones only add per-unit work there.
constexpr int kUnitCap = 16; // the unit reductions span lanes 0..15
constexpr int unit_max_single(int dim) { return dim == 1 ? 4 : 8; }

template <bool MLP, int NT = kClassifyNT> struct ClassifyLds {
    static constexpr int NW = NT / kWave;
    uint64_t t[NT + 2 * kClassifyHalo];
    int wc[NW][3]; // per wave: candidates, units (front list), units (back list)
    int base[3];
    alignas(16) __bf16 h[MLP ? NW : 1][MLP ? kWave * kMlpH : 1];
    float x[MLP ? NW : 1][MLP ? kWave : 1][kFeatStride];
    float z[MLP ? NW : 1][kWave];
    float w[MLP ? kMlpNP : 1]; // the predictor's weights
};

// SIGNAL: the workgroup counts itself done in the slide's n_done once its
// unit list entries are visible at agent scope, and the last one raises the
// slide's ready flags (a detect in the same launch waits for them)
template <int DIM, int MODE, int NT = kClassifyNT, bool SIGNAL = false>
__device__ __forceinline__ void classify_body(const Dev &d, const Step &st, int bid,
                                              ClassifyLds<MODE == kClsLearned, NT> &lds)
{
    constexpr int HALO = kClassifyHalo;
    constexpr bool LEARN = MODE == kClsLearned;
    uint64_t *s_t = lds.t;
    const int i0 = bid * NT;
    if (bid == 0 && threadIdx.x == 0)
        d.ctl->t_det = st.t;
    const int i = i0 + threadIdx.x;
    const uint64_t *Nk = d.akey[st.par ^ 1];
    const int W = d.W, K = d.K;
    // keys read: [i-K, i] for the run start, [i, i+63] for a unit's end;
    // learned features reach 2K+2 either side
    const int kUnitMax = d.unit_max;
    const int H = LEARN ? max(2 * K + 2, kUnitCap) : max(K, kUnitCap);
    const bool tiled = H <= HALO;
    const int t0 = max(0, i0 - H), t1 = min(W, i0 + NT + H);
    // learned: this point's coordinates are requested first, and wave 0
    // stages the predictor's published weights in LDS while the other waves
    // stage the key tile, so the two chains of round trips overlap
    float xi[3] = {0.f, 0.f, 0.f};
    if constexpr (LEARN) {
        if (i < W) {
            const float *Nx = d.axyz[st.par ^ 1];
#pragma unroll
            for (int ax = 0; ax < DIM; ++ax)
                xi[ax] = Nx[(size_t)ax * W + i];
        }
        const int wv = threadIdx.x / kWave;
        if (wv == 0) { // all of a lane's loads in one round trip
            constexpr int PER = (kMlpNP + kWave - 1) / kWave;
            const int lane = threadIdx.x;
            // the published set, read without an acquire (which would
            // invalidate this XCD's L2 under the other waves' tile loads):
            // its index and words are read at agent scope instead.  The
            // weights only order the exact counts, so a set read while the
            // trainer publishes the next one cannot change an outcome.
            const int cur = __hip_atomic_load(&d.lr.ls->pub, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
            const float *P = d.lr.pub + (size_t)(cur & 1) * kMlpNP;
            float v[PER];
#pragma unroll
            for (int u = 0; u < PER; ++u)
                v[u] = lane + u * kWave < kMlpNP
                           ? __hip_atomic_load(&P[lane + u * kWave], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT)
                           : 0.f;
#pragma unroll
            for (int u = 0; u < PER; ++u)
                if (lane + u * kWave < kMlpNP)
                    lds.w[lane + u * kWave] = v[u];
        } else if (tiled) {
            constexpr int NTT = NT - kWave;
            for (int q = threadIdx.x - kWave; q < t1 - t0; q += 8 * NTT) {
                uint64_t v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    v[u] = q + u * NTT < t1 - t0 ? Nk[t0 + q + u * NTT] : 0ull;
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (q + u * NTT < t1 - t0)
                        s_t[q + u * NTT] = v[u];
            }
        }
    } else if (tiled) {
        copy_to_lds<(NT + 2 * HALO) / NT>(s_t, Nk + t0, t1 - t0);
    }
    __syncthreads();
    auto kat = [&](int j) -> uint64_t { return tiled ? s_t[j - t0] : Nk[j]; };
    // first index in [max(0, p-cap), p] holding `key` (key[p] == key)
    auto run_lo = [&](int p, uint64_t key, int cap) {
        int lo = max(0, p - cap), hi = p;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (kat(mid) == key)
                hi = mid;
            else
                lo = mid + 1;
        }
        return lo;
    };
    // last index in [p, min(W-1, p+cap)] holding `key` (key[p] == key)
    auto run_hi = [&](int p, uint64_t key, int cap) {
        int lo = p, hi = min(W - 1, p + cap);
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (kat(mid) == key)
                lo = mid;
            else
                hi = mid - 1;
        }
        return lo;
    };
    // ablation "no grid filter" (d.no_filter): no point is fast-accepted; with
    // a ray limit only the points in ring slots below it are candidates, each
    // a unit of its own
    bool cand = false;
    uint64_t key = 0;
    if (i < W) {
        key = kat(i);
        bool dense = false;
        if (!d.no_filter && K < W && key_guaranteed<DIM>(key)) {
            const int lo = run_lo(i, key, K);
            dense = (lo + K < W) && (kat(lo + K) == key);
        }
        cand = !dense;
        if (d.ray_lim > 0)
            cand = cand && d.aslot[st.par ^ 1][i] < d.ray_lim;
    }
    const bool singles = d.ray_lim > 0;
    const int lane = lane_id();
    const uint64_t lt = (1ull << lane) - 1ull;
    if constexpr (MODE == kClsPoints) {
        const uint64_t mask = __ballot(cand);
        const int cnt = __popcll(mask);
        // the wave's 64 positions start at a multiple of 64 (NT is one)
        const int wbase = i0 + (threadIdx.x & ~(kWave - 1));
        if (d.cmask && lane == 0 && wbase < W)
            d.cmask[wbase >> 6] = mask;
        // per workgroup, once every wave is done reading the key tile (which
        // then holds the waves' figures): one atomic for the candidate list
        // base, and the window's per-axis minimum key field -- the origin of
        // the LBVH's Morton codes, so that codes of data straddling a
        // power-of-two cell coordinate stay narrow (fewer sort passes)
        uint64_t f[3], fm[3];
        key_fields<DIM>(key, f);
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
            uint64_t m = i < W ? f[a] : ~0ull;
            for (int o = 1; o < kWave; o <<= 1) {
                const uint64_t v = __shfl_xor(m, o, kWave);
                m = v < m ? v : m;
            }
            fm[a] = m;
        }
        constexpr int NWV = NT / kWave;
        static_assert(4 * NWV + 1 <= NT, "the waves' figures fit the key tile");
        const int wid = threadIdx.x / kWave;
        __syncthreads();
        if (lane == 0) {
            s_t[4 * wid] = (uint64_t)cnt;
#pragma unroll
            for (int a = 0; a < DIM; ++a)
                s_t[4 * wid + 1 + a] = fm[a];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t tot = 0;
            for (int w = 0; w < NWV; ++w)
                tot += s_t[4 * w];
            s_t[4 * NWV] = tot ? (uint64_t)atomicAdd(&cnt_of(d, st.t)->n_cand, (int)tot) : 0;
        } else if (threadIdx.x <= DIM && (bid == 0 || threadIdx.x > 1)) {
            // axis 0 is the key's most significant field, so the window's
            // first point holds its minimum: workgroup 0 alone records it;
            // the other axes' minima are recorded where they lower the value
            // read (minima only decrease), not by every workgroup's atomic
            // on the one address
            const int a = threadIdx.x - 1;
            uint64_t m = ~0ull;
            for (int w = 0; w < NWV; ++w) {
                const uint64_t v = s_t[4 * w + 1 + a];
                m = v < m ? v : m;
            }
            unsigned long long *fa = &cnt_of(d, st.t)->fmin[a];
            if (m != ~0ull &&
                m < __hip_atomic_load(fa, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                atomicMin(fa, (unsigned long long)m);
        }
        __syncthreads();
        if (cand) {
            int base = (int)s_t[4 * NWV];
            for (int w = 0; w < wid; ++w)
                base += (int)s_t[4 * w];
            d.cand[base + __popcll(mask & lt)] = i;
        }
        return;
    } else {
        // per wave: candidate count, unit starts (and, learned, their
        // buckets); one atomic per counter per workgroup
        const int wid = threadIdx.x / kWave;
        const uint64_t cmask = __ballot(cand);
        const bool ustart =
            cand && (singles || i == 0 || (i & (kUnitMax - 1)) == 0 || kat(i - 1) != key);
        int ulen = 0;
        if (ustart) {
            const int end = min(W, (i | (kUnitMax - 1)) + 1); // end of the aligned block
            ulen = singles ? 1 : run_hi(i, key, end - 1 - i) - i + 1;
        }
        uint64_t mo = __ballot(ustart), mi = 0; // front / back list members
        float f[kMlpIn] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if constexpr (LEARN) {
            if (mo) { // the wave holds unit starts: features and prediction
                constexpr uint64_t FM = CellBits<DIM>::FMASK;
                const float capf = 1.f / (float)(K + 1);
                if (ustart) {
                    const int a = run_lo(i, key, K), b = run_hi(i, key, K);
                    f[0] = (float)min(b - a + 1, K + 1) * capf;
                    const uint64_t last = key & FM;
                    if (last > 0 && a > 0 && kat(a - 1) == key - 1)
                        f[1] = (float)min(a - run_lo(a - 1, key - 1, K), K + 1) * capf;
                    if (last < FM && b + 1 < W && kat(b + 1) == key + 1)
                        f[2] = (float)min(run_hi(b + 1, key + 1, K) - b, K + 1) * capf;
#pragma unroll
                    for (int ax = 0; ax < DIM; ++ax) {
                        const double u = (double)xi[ax] * d.inv_c;
                        const float off = (float)(u - floor(u)) - 0.5f;
                        f[3 + ax] = (off >= -0.5f && off <= 0.5f) ? off : 0.f; // non-finite -> 0
                    }
                }
                // the wave's unit starts, compacted into rows 0..m-1
                const int m = __popcll(mo);
                const int row = __popcll(mo & lt);
                if (ustart) {
#pragma unroll
                    for (int j = 0; j < kFeatStride; ++j)
                        lds.x[wid][row][j] = j < kMlpIn ? f[j < kMlpIn ? j : 0] : 0.f;
                }
                wave_lds_fence();
                const MlpRegs Wr = mlp_regs(lds.w);
                mlp_infer_rows(Wr, lds.x[wid], m, lds.h[wid], lds.z[wid]);
                const float yh = ustart ? sigmoidf(lds.z[wid][row] + Wr.b3) : 0.f;
                wave_lds_fence();
                const bool pin = ustart && yh >= kInlierCut;
                mi = __ballot(pin);
                mo = __ballot(ustart && !pin);
            }
        }
        if (lane == 0) {
            lds.wc[wid][0] = __popcll(cmask);
            lds.wc[wid][1] = __popcll(mo);
            lds.wc[wid][2] = __popcll(mi);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int tot[3] = {0, 0, 0};
            for (int w = 0; w < ClassifyLds<LEARN, NT>::NW; ++w)
                for (int q = 0; q < 3; ++q)
                    tot[q] += lds.wc[w][q];
            if (tot[0])
                atomicAdd(&cnt_of(d, st.t)->n_cand, tot[0]);
            lds.base[1] = tot[1] ? atomicAdd(&cnt_of(d, st.t)->n_units, tot[1]) : 0;
            if constexpr (LEARN)
                lds.base[2] = tot[2] ? atomicAdd(&cnt_of(d, st.t)->n_units_i, tot[2]) : 0;
        }
        __syncthreads();
        int bo = lds.base[1], bi = LEARN ? lds.base[2] : 0;
        for (int w = 0; w < wid; ++w) {
            bo += lds.wc[w][1];
            bi += lds.wc[w][2];
        }
        if (ustart) {
            const bool front = (mo >> lane) & 1ull;
            const int li = front ? bo + __popcll(mo & lt) : W - 1 - (bi + __popcll(mi & lt));
            if constexpr (SIGNAL) // coherent across XCDs without an L2 write-back
                __hip_atomic_store(reinterpret_cast<unsigned long long *>(d.units + li),
                                   (unsigned long long)(unsigned)i |
                                       ((unsigned long long)(unsigned)ulen << 32),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            else
                d.units[li] = make_int2(i, ulen);
            if constexpr (LEARN) {
                float *rx = d.lr.rep_x +
                            ((size_t)(st.t % kReplaySlides) * W + li) * kFeatStride;
                *reinterpret_cast<float4 *>(rx) = make_float4(f[0], f[1], f[2], f[3]);
                *reinterpret_cast<float4 *>(rx + 4) = make_float4(f[4], f[5], 0.f, 0.f);
            }
        }
        if constexpr (SIGNAL) {
            // the list entries were stored at agent scope and the counters
            // updated by device atomics: once every wave's are acknowledged
            // they are visible to agent-scope loads on any XCD, so the
            // workgroup counts itself done without a release (whose L2
            // write-back, under the window merge's stores, costs
            // microseconds per workgroup)
            // The last to count raises the slide's ready flags: every other
            // workgroup's entries were acknowledged before its count.
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __syncthreads();
            if (threadIdx.x == 0) {
                const int ncls = (W + NT - 1) / NT;
                lds.base[0] = __hip_atomic_fetch_add(&cnt_of(d, st.t)->n_done, 1,
                                                     __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT) == ncls - 1;
            }
            __syncthreads();
            if (lds.base[0] && threadIdx.x < kReadyFlags)
                __hip_atomic_store(ready_of(d, st.t, threadIdx.x), 1, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

template <int DIM, int MODE>
__global__ __launch_bounds__(kClassifyNT) void k_classify(Dev d, Step st)
{
    __shared__ ClassifyLds<MODE == kClsLearned> lds;
    stamp_start(d, kTsClassify);
    classify_body<DIM, MODE>(d, st, blockIdx.x, lds);
}

// ---------------------------------------------------------------------------
// Stencil of a candidate: per axis the cells overlapping [x - R', x + R'],
// R' = R (1 + 1e-5) > any |dx| the FP32 predicate accepts.  Rows are
// (k0[,k1]) fixed with the last axis spanning a contiguous key range.
// ---------------------------------------------------------------------------
template <int DIM> struct Stencil {
    int lo[3], hi[3];
    int nrows, n1;
};

template <int DIM>
__device__ __forceinline__ void make_stencil(const float (&p)[3], double inv_c, double Rp,
                                             Stencil<DIM> &st)
{
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        st.lo[a] = 0;
        st.hi[a] = 0;
    }
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
        st.lo[a] = cell_coord<DIM>((double)p[a] - Rp, inv_c);
        st.hi[a] = cell_coord<DIM>((double)p[a] + Rp, inv_c);
    }
    if constexpr (DIM == 1) {
        st.nrows = 1;
        st.n1 = 1;
    } else if constexpr (DIM == 2) {
        st.nrows = (int)(st.hi[0] - st.lo[0] + 1);
        st.n1 = 1;
    } else {
        st.n1 = (int)(st.hi[1] - st.lo[1] + 1);
        st.nrows = (int)(st.hi[0] - st.lo[0] + 1) * st.n1;
    }
}

template <int DIM>
__device__ __forceinline__ void row_keys(const Stencil<DIM> &st, int r, uint64_t &klo,
                                         uint64_t &khi)
{
    if constexpr (DIM == 1) {
        klo = pack_key<1>(st.lo[0], 0, 0);
        khi = pack_key<1>(st.hi[0], 0, 0);
    } else if constexpr (DIM == 2) {
        const int64_t k0 = st.lo[0] + r;
        klo = pack_key<2>(k0, st.lo[1], 0);
        khi = pack_key<2>(k0, st.hi[1], 0);
    } else {
        const int64_t k0 = st.lo[0] + r / st.n1;
        const int64_t k1 = st.lo[1] + r % st.n1;
        klo = pack_key<3>(k0, k1, st.lo[2]);
        khi = pack_key<3>(k0, k1, st.hi[2]);
    }
}

// 8 wavefronts share one LDS copy of the sample table; registers are capped
// for 6 wavefronts per SIMD, so three workgroups (24 wavefronts) fit a CU:
// the per-unit work is a chain of dependent memory round trips, and more
// resident units hide more of it
constexpr int kDetectNT = 512;
#define RTOD_DETECT_OCCUPANCY __attribute__((amdgpu_waves_per_eu(6, 8)))

__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Finishes bounds inside one gap of a sorted table each: bound b's answer
// lies past tab[w_L[b] - 1] and at or before tab[w_L[b] + 2^sl2 - 2]; the
// 2^sl2 - 1 positions from w_L[b] are probed (w_L[b] < 0: nothing to probe,
// positions >= lim do not exist).  The 64 lanes walk the (bound, offset)
// pairs 64 at a time, U batches of loads in flight, and each batch's ballot
// is credited to the bound's owner lane: lane b returns how many of its
// positions pass (tab[i] < key for a lower bound, <= key for an upper one).
template <int U>
__device__ __forceinline__ int refine_bounds(const uint64_t *tab, const int *w_L, int lim, int nb,
                                             int sl2, uint64_t key)
{
    const int lane = lane_id();
    const int stride = 1 << sl2;
    const int niter = (nb * stride + kWave - 1) / kWave;
    int my_cnt = 0;
    for (int q0 = 0; q0 < niter; q0 += U) {
        // a position with nothing to probe loads as ~0, which passes neither
        // test (keys use at most 63 bits)
        uint64_t xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = (q0 + u) * kWave + lane;
            const int b = e >> sl2, r = e & (stride - 1);
            const int Lb = (b < nb) ? w_L[b] : -1;
            const bool ok = (q0 + u < niter) && Lb >= 0 && r < stride - 1 && Lb + r < lim;
            xv[u] = ok ? tab[Lb + r] : ~0ull;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = q0 + u;
            const int b = (q * kWave + lane) >> sl2;
            const uint64_t kb = __shfl(key, b & 63, kWave);
            const bool less = (b & 1) ? (xv[u] <= kb) : (xv[u] < kb);
            const uint64_t m = __ballot(less);
            const int b0 = (q * kWave) >> sl2;
            if (stride >= kWave) {
                if (lane == b0)
                    my_cnt += __popcll(m);
            } else {
                const int g = lane - b0;
                if (g >= 0 && g < (kWave >> sl2))
                    my_cnt += __popcll(m & (((1ull << stride) - 1ull) << (g * stride)));
            }
        }
    }
    return my_cnt;
}

// Row bounds of a candidate's stencil: lane b < 2*nrows computes bound b
// (row b>>1, upper if b&1).  The sample table in LDS (stride S1 = 2^sl2)
// narrows each bound to one gap of S1 - 1 keys.  With S1 <= 4, or few
// enough bounds that all their gaps load in one round trip, those keys are
// probed directly; else the every-4th-key table narrows the gap to 3 keys
// first (S1/4 - 1 probes, then 3): two short coalesced rounds instead of
// S1 - 1 probes per bound, which is most of a 3-D unit's memory traffic.
template <int DIM>
__device__ __forceinline__ int row_bounds(const Dev &d, const uint64_t *Nk, const uint64_t *Nk4,
                                          const uint64_t *s_samp, const Stencil<DIM> &st,
                                          int *w_L)
{
    const int lane = lane_id();
    const int W = d.W, sl2 = d.stride_log2;
    const int nb = 2 * st.nrows;
    uint64_t key = 0;
    int lo = 0;
    if (lane < nb) {
        uint64_t klo, khi;
        row_keys<DIM>(st, lane >> 1, klo, khi);
        const bool upper = lane & 1;
        key = upper ? khi : klo;
        int hi = d.nsamp;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            const bool right = upper ? (s_samp[mid] <= key) : (s_samp[mid] < key);
            if (right)
                lo = mid + 1;
            else
                hi = mid;
        }
    }
    // the answer is 0 when no sample passes; else it lies in
    // [(lo-1)*S1 + 1, lo*S1], element (lo-1)*S1 passing.  Probed directly
    // when every bound's gap fits the loads refine_bounds keeps in flight
    // (8 batches of 64): one round trip, over the same few cache lines the
    // two rounds through the every-4th-key table would touch.
    if (!Nk4 || (nb << sl2) <= 8 * kWave) {
        if (lane < nb)
            w_L[lane] = lo > 0 ? ((lo - 1) << sl2) + 1 : -1;
        wave_sync_lds();
        const int c = refine_bounds<8>(Nk, w_L, W, nb, sl2, key);
        return (lane < nb && lo > 0) ? ((lo - 1) << sl2) + 1 + c : 0;
    }
    if (lane < nb)
        w_L[lane] = lo > 0 ? ((lo - 1) << (sl2 - 2)) + 1 : -1;
    wave_sync_lds();
    const int c4 = refine_bounds<8>(Nk4, w_L, d.nsamp4, nb, sl2 - 2, key);
    // element base passes, element base + 4 does not (or is past the end)
    const int base = lo > 0 ? ((lo - 1) << sl2) + 4 * c4 : 0;
    wave_sync_lds();
    if (lane < nb)
        w_L[lane] = lo > 0 ? base + 1 : -1;
    wave_sync_lds();
    const int c1 = refine_bounds<8>(Nk, w_L, W, nb, 2, key);
    return (lane < nb && lo > 0) ? base + 1 + c1 : 0;
}

// min / max over lanes 0..15 (a unit's candidates), broadcast to the wave
__device__ __forceinline__ int unit_min_i(int x)
{
#pragma unroll
    for (int m = 1; m < 16; m <<= 1)
        x = min(x, __shfl_xor(x, m, kWave));
    return __builtin_amdgcn_readfirstlane(x);
}

__device__ __forceinline__ int unit_max_i(int x)
{
#pragma unroll
    for (int m = 1; m < 16; m <<= 1)
        x = max(x, __shfl_xor(x, m, kWave));
    return __builtin_amdgcn_readfirstlane(x);
}

__device__ __forceinline__ float lane_f(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// Counts of a cell unit's candidates (<= d.unit_max) over the window points in the cells of
// stencil `st` (a superset of each candidate's own stencil; the predicate
// decides).  Lane c < n holds candidate c's coordinates in pc and receives
// its count, exact up to K+1 at least.  Chunks of 64 stencil points stream
// through the lanes; each chunk is tested against every candidate (one
// ballot + popcount per candidate) and the scan stops once every
// candidate's count exceeds K.  Wave-cooperative; tests += evaluations.
template <int DIM>
__device__ __forceinline__ int unit_count(const Dev &d, const uint64_t *Nk, const uint64_t *Nk4,
                                          const float *Nx, const uint64_t *s_samp, int *w_a,
                                          int *w_b,
                                          const float (&pc)[3], int n, Stencil<DIM> st,
                                          long long &tests)
{
    const int lane = lane_id();
    const int W = d.W, K = d.K;
    if (st.nrows > 32) { // cannot happen: the stencils stay within 32 rows
        if (lane == 0)
            atomicOr(&d.ctl->err, kErrStencilRows);
        st.nrows = 32;
    }
    const int bnd = row_bounds<DIM>(d, Nk, Nk4, s_samp, st, w_a);
    wave_sync_lds();
    // lane r < nrows: row r spans [bound 2r, bound 2r+1)
    const int r = lane;
    const int lo_b = __shfl(bnd, (2 * r) & 63, kWave);
    const int hi_b = __shfl(bnd, (2 * r + 1) & 63, kWave);
    const int len = (r < st.nrows) ? (hi_b - lo_b) : 0;
    const int incl = wave_incl_scan(len);
    const int total = __builtin_amdgcn_readfirstlane(__shfl(incl, 63, kWave));
    wave_sync_lds();
    if (r < st.nrows) {
        w_a[r] = lo_b;
        w_b[r] = incl - len;
    }
    wave_sync_lds();

    // PF chunks at a time: their row lookups and coordinate loads are
    // issued together, so a group costs one memory round trip (4 in 3-D:
    // three coordinate planes per chunk, within the detect's register cap;
    // 8 there measured 2% slower on the large window; 4 in 1-D, where the
    // units of 4 candidates scan short stencils)
    constexpr int PF = DIM == 2 ? 8 : 4;
    int cnt = 0;
    for (int base0 = 0; base0 < total; base0 += PF * kWave) {
        // a lane past the stencil's end holds NaN, which no predicate passes
        // (as a NaN coordinate in the window passes none)
        float qx[PF], qy[PF], qz[PF];
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            const int idx = base0 + k * kWave + lane;
            qx[k] = __builtin_nanf("");
            qy[k] = qz[k] = 0.f;
            if (idx < total) {
                int j;
                if constexpr (DIM == 1) {
                    j = w_a[0] + idx;
                } else {
                    int lo = 0, hi = st.nrows - 1;
                    while (lo < hi) { // last row with off <= idx
                        const int mid = (lo + hi + 1) >> 1;
                        if (w_b[mid] <= idx)
                            lo = mid;
                        else
                            hi = mid - 1;
                    }
                    j = w_a[lo] + (idx - w_b[lo]);
                }
                qx[k] = Nx[j];
                if constexpr (DIM > 1)
                    qy[k] = Nx[(size_t)W + j];
                if constexpr (DIM > 2)
                    qz[k] = Nx[(size_t)2 * W + j];
            }
        }
        // candidate by candidate over the group's chunks: its coordinates are
        // broadcast once per group, and it skips the groups after its own
        // count exceeds K (counts stay exact up to K+1)
        const int nk = min(PF, (total - base0 + kWave - 1) / kWave);
        int pairs = 0; // predicate evaluations of this group (uniform)
        for (int c = 0; c < n; ++c) {
            const int cc = __builtin_amdgcn_readlane(cnt, c);
            if (cc > K)
                continue;
            const float p3[3] = {lane_f(pc[0], c), DIM > 1 ? lane_f(pc[1], c) : 0.f,
                                 DIM > 2 ? lane_f(pc[2], c) : 0.f};
            // the group's chunks are counted whole: branch-free, the
            // predicate's lane mask feeds the popcount directly
            int h = 0;
#pragma unroll
            for (int k = 0; k < PF; ++k) {
                if (k < nk) { // uniform
                    const bool hit = dist2<DIM>(p3, qx[k], qy[k], qz[k]) <= d.r2;
                    h += __builtin_popcountll(__builtin_amdgcn_ballot_w64(hit));
                }
            }
            pairs += min(nk * kWave, total - base0);
            cnt += (lane == c) ? h : 0;
        }
        tests += pairs;
        if (__ballot(lane < n && cnt <= K) == 0)
            break;
    }
    return cnt;
}

// Stencil covering every candidate of a unit: per axis the union of
// [floor((x - R') / c), floor((x + R') / c)] over lanes < n.
template <int DIM>
__device__ __forceinline__ Stencil<DIM> unit_stencil(const Dev &d, const float (&pc)[3], int n)
{
    const int lane = lane_id();
    Stencil<DIM> st;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        st.lo[a] = 0;
        st.hi[a] = 0;
    }
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
        int lo = INT_MAX, hi = INT_MIN;
        if (lane < n) {
            lo = cell_coord<DIM>((double)pc[a] - d.Rp, d.inv_c);
            hi = cell_coord<DIM>((double)pc[a] + d.Rp, d.inv_c);
        }
        st.lo[a] = unit_min_i(lo);
        st.hi[a] = unit_max_i(hi);
    }
    if constexpr (DIM == 1) {
        st.nrows = 1;
        st.n1 = 1;
    } else if constexpr (DIM == 2) {
        st.nrows = st.hi[0] - st.lo[0] + 1;
        st.n1 = 1;
    } else {
        st.n1 = st.hi[1] - st.lo[1] + 1;
        st.nrows = (st.hi[0] - st.lo[0] + 1) * st.n1;
    }
    return st;
}

// The unit's own cell and its neighbours (own +- 1 per axis) within `full`.
template <int DIM>
__device__ __forceinline__ Stencil<DIM> near_stencil(const Dev &d, const Stencil<DIM> &full,
                                                     const float (&pc)[3])
{
    Stencil<DIM> nr = full;
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
        const int own = __builtin_amdgcn_readfirstlane(cell_coord<DIM>((double)pc[a], d.inv_c));
        nr.lo[a] = max(full.lo[a], own - 1);
        nr.hi[a] = min(full.hi[a], own + 1);
    }
    if constexpr (DIM == 2) {
        nr.nrows = nr.hi[0] - nr.lo[0] + 1;
    } else if constexpr (DIM == 3) {
        nr.n1 = nr.hi[1] - nr.lo[1] + 1;
        nr.nrows = (nr.hi[0] - nr.lo[0] + 1) * nr.n1;
    }
    return nr;
}

// ---------------------------------------------------------------------------
// Slide epilogue, in F parts of NT threads (one workgroup each) once the
// slide's detect phase is complete: the last detect workgroup (F = 1), or
// F workgroups of the next slide's window-merge launch.  The ascending slot
// list is written from the slide's outlier bitmap straight into pinned host
// memory; the last part to finish publishes the counters, resets them for
// the next slide and stamps the slide.  The bitmap is cut into chunks of
// kChunkWords words (thread w holds word c * kChunkWords + w of chunk c),
// chunk c taken by part
// c mod F: a part ranks its chunk's bits among the slide's (a block sum of
// the popcounts of the words before it, a block scan within it), stages the
// chunk's slots in LDS and writes them out by consecutive threads
// (coalesced writes into host memory).  The bitmap is left set: the window
// merge of the slide two later, its next user, clears it.
// ---------------------------------------------------------------------------
constexpr int kCompactRegs = 16;  // chunks whose words a thread keeps in registers
constexpr int kChunkWords = 256;  // bitmap words per chunk
constexpr int kFinishCap = 32 * kChunkWords; // slots of one chunk, staged in LDS

struct FinishLds {
    int list[kFinishCap];
    int red[4][1024 / kWave];
    int scan[1024 / kWave];
    int last;
};

// what the last part publishes besides the list: n_o / n_units, the slide's
// probable-outlier and total unit counts; use_pred, whether its detect
// followed the prediction (learned schedule)
struct SlideSummary {
    int n_o, n_units;
    bool learn, use_pred;
    bool stamps; // one slide per call: copy the kernels' start stamps into
                 // the host-out buffer
};


// block sums of a and b and block OR of m (wave-uniform m); every thread
// gets all three
template <int NT>
__device__ __forceinline__ void block_sum2(int a, int b, uint32_t m, int (*s_red)[1024 / kWave],
                                           int &sa, int &sb, uint32_t &om)
{
    const int lane = lane_id(), wid = threadIdx.x / kWave;
#pragma unroll
    for (int o = kWave / 2; o >= 1; o >>= 1) {
        a += __shfl_xor(a, o, kWave);
        b += __shfl_xor(b, o, kWave);
    }
    if (lane == 0) {
        s_red[0][wid] = a;
        s_red[1][wid] = b;
        s_red[2][wid] = (int)m;
    }
    __syncthreads();
    sa = 0;
    sb = 0;
    om = 0;
#pragma unroll
    for (int w = 0; w < NT / kWave; ++w) {
        sa += s_red[0][w];
        sb += s_red[1][w];
        om |= (uint32_t)s_red[2][w];
    }
    __syncthreads();
}

// Exclusive block scan of four words per thread, each packing four 8-bit
// counters; the caller keeps every column sum below 256, so no lane
// carries into the next.  tot: the block totals.
template <int NT>
__device__ __forceinline__ void block_excl_scan4(uint32_t (&v)[4], int (*s_red)[1024 / kWave],
                                                 uint32_t (&tot)[4])
{
    const int lane = lane_id(), wid = threadIdx.x / kWave;
    uint32_t inc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
        inc[q] = v[q];
#pragma unroll
    for (int dd = 1; dd < kWave; dd <<= 1) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t y = __shfl_up(inc[q], dd, kWave);
            inc[q] += lane >= dd ? y : 0u;
        }
    }
    if (lane == kWave - 1) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
            s_red[q][wid] = (int)inc[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        uint32_t pre = 0, t = 0;
#pragma unroll
        for (int w = 0; w < NT / kWave; ++w) {
            const uint32_t u = (uint32_t)s_red[q][w];
            pre += (w < wid) ? u : 0u;
            t += u;
        }
        tot[q] = t;
        v[q] = pre + inc[q] - v[q];
    }
    __syncthreads();
}

// a slide with at most this many outliers is listed by one part in one
// pass (block_excl_scan4 ranks the bits of all chunks at once)
constexpr int kSparseList = 255;

// chunk c (this thread's word: wd) to list positions [P, P + its bits);
// returns its bit count
template <int NT>
__device__ __forceinline__ int emit_chunk(HostOut *ho, FinishLds &fl, int c, uint32_t wd, int P)
{
    int ctot;
    int p = block_excl_scan<NT>(__popc(wd), fl.scan, ctot);
    const int word = c * kChunkWords + (int)threadIdx.x;
    while (wd) {
        fl.list[p++] = (word << 5) + __builtin_ctz(wd);
        wd &= wd - 1;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < ctot; i += NT)
        ho->list[P + i] = fl.list[i];
    __syncthreads();
    return ctot;
}

// The counters publish_slide reports, requested by one thread when the
// epilogue starts (with the bitmap loads) rather than when it ends: every
// one of them is final by then (detect and classify are complete), and
// error flags raised later (by the next slide's merge running beside the
// epilogue) stay set for the next slide.
struct PubIn {
    unsigned long long pair_tests, hits;
    int32_t err, n_cand, misses, train_steps;
    float train_loss;
};

__device__ __forceinline__ PubIn load_pub_inputs(const Dev &d, const Step &st, bool learn)
{
    PubIn in;
    SlideCnt *cn = cnt_of(d, st.t);
    in.pair_tests = __hip_atomic_load(&cn->pair_tests, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    in.hits = __hip_atomic_load(&cn->hits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    in.err = __hip_atomic_load(&d.ctl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    in.n_cand = cn->n_cand;
    in.misses = in.train_steps = 0;
    in.train_loss = 0.f;
    if (learn) {
        in.misses = __hip_atomic_load(&cn->misses, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        in.train_steps = d.lr.ls->steps_total;
        in.train_loss = d.lr.ls->loss_last;
    }
    return in;
}

// The slide's counters and learned bookkeeping, then the stamp, by one
// thread once the whole list is written (its own wave's list stores, or
// every part's, released before their tickets); total: the slide's outlier
// count.  The stamp is written after one system-scope release (cumulative
// over the list writes ordered before it): a host that sees the stamp
// (polling, acquire) sees the list and the counters.  The error flags
// reported are cleared by an atomic AND of their complement, so a flag
// raised after they were read is reported with the next slide, never lost.
__device__ __forceinline__ void publish_slide(const Dev &d, const Step &st,
                                              const SlideSummary &sm, const PubIn &in,
                                              HostOut *ho, int total)
{
    if (sm.learn) {
        LearnCtl *lc = d.lr.lc;
        const int slot = (int)(st.t % kReplaySlides);
        d.lr.rep_n[2 * slot] = sm.n_o;
        d.lr.rep_n[2 * slot + 1] = sm.n_units - sm.n_o;
        ho->misses = in.misses;
        ho->used_pred = sm.use_pred ? 1 : 0;
        ho->train_steps = in.train_steps;
        ho->train_loss = in.train_loss;
        // fallback rule: schedule slide t + 2 by prediction only if this
        // slide's miss rate (over units) is within the threshold; a slide
        // without units keeps the previous decision
        int32_t *dn = &lc->dec[(st.t + 2) % kDecSlots];
        *dn = sm.n_units > 0
                  ? (((float)in.misses <= d.lr.miss_thr * (float)sm.n_units) ? 1 : 0)
                  : lc->dec[(st.t + 1) % kDecSlots];
    } else if (d.lr.on) {
        // a learned context's slide on the plain kernels: no prediction, no
        // replay entry (its slot keeps an earlier slide's exact counts); the
        // miss-rate decision carries over to slide t + 2
        LearnCtl *lc = d.lr.lc;
        lc->dec[(st.t + 2) % kDecSlots] = lc->dec[(st.t + 1) % kDecSlots];
        ho->misses = 0;
        ho->used_pred = 0;
        ho->train_steps = d.lr.ls->steps_total;
        ho->train_loss = d.lr.ls->loss_last;
    }
    ho->learn_kernels = sm.learn ? 1 : 0;
    if (sm.stamps) {
#pragma unroll
        for (int k = 0; k < kTsPublished; ++k)
            ho->ts[k] = d.tstamp[k];
        ho->ts[kTsPublished] = wall_clock();
    }
    ho->n_units = sm.n_units;
    ho->n_out = total;
    ho->n_cand = in.n_cand;
    ho->pair_tests = in.pair_tests;
    ho->hits = in.hits;
    ho->err = in.err;
    if (in.err)
        atomicAnd(&d.ctl->err, ~in.err); // the flags reported here
    d.ctl->ticket = 0;
    d.ctl->fin_ticket = 0;
    d.ctl->t = st.t + 1;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
    __hip_atomic_store(&ho->slide, (int64_t)(st.t + 1), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_SYSTEM);
}

template <int NT>
__device__ __forceinline__ void finish_slide(const Dev &d, const Step &st, int f, int F,
                                             const SlideSummary &sm, FinishLds &fl)
{
    static_assert(NT >= kChunkWords, "a word of the chunk per thread");
    const int tid = threadIdx.x;
    const int nwords = (d.W + 31) >> 5;
    const int nch = (nwords + kChunkWords - 1) / kChunkWords;
    // thread tid holds word tid of each chunk (none past the chunk)
    const int nmine = tid < kChunkWords ? nwords : 0;
    // the bitmap's words are read with agent-scope loads: in the last detect
    // workgroup they were set by other workgroups' atomics in the same
    // kernel, possibly on other XCDs, whose L2 lines a plain load could miss
    uint32_t *bmw = bitmap_of(d, st.t);
    auto bm_load = [&](int w) {
        return __hip_atomic_load(&bmw[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    HostOut *ho = host_out(d, st.t);
    PubIn in{};
    if (tid == 0)
        in = load_pub_inputs(d, st, sm.learn);
    int P = 0, total = 0;
    const bool single = F == 1; // this part lists the whole slide
    if (nch <= kCompactRegs) {
        uint32_t wv[kCompactRegs];
#pragma unroll
        for (int k = 0; k < kCompactRegs; ++k)
            wv[k] = (k < nch && k * kChunkWords + tid < nmine) ? bm_load(k * kChunkWords + tid) : 0u;
        int before = 0, all = 0;
        uint32_t nz = 0; // bit k: chunk k holds outlier bits (this wave's words)
#pragma unroll
        for (int k = 0; k < kCompactRegs; ++k) {
            const int b = __popc(wv[k]);
            all += b;
            before += (k < f) ? b : 0;
            nz |= __ballot(wv[k] != 0u) ? (1u << k) : 0u;
        }
        block_sum2<NT>(before, all, nz, fl.red, P, total, nz);
        if (total <= kSparseList) {
            // part 0 alone, one pass: bit b of word w of chunk k goes to
            // (bits of chunks < k) + (bits of chunk k in threads < w) +
            // popc(word below b); the per-chunk prefixes over threads come
            // from one scan of the packed counts (<= 255 each)
            if (f != 0)
                return;
            uint32_t pk[4] = {0u, 0u, 0u, 0u};
#pragma unroll
            for (int k = 0; k < kCompactRegs; ++k)
                pk[k >> 2] |= (uint32_t)__popc(wv[k]) << (8 * (k & 3));
            uint32_t tot[4];
            block_excl_scan4<NT>(pk, fl.red, tot);
            int cpre = 0;
#pragma unroll
            for (int k = 0; k < kCompactRegs; ++k) {
                int p = cpre + (int)((pk[k >> 2] >> (8 * (k & 3))) & 255u);
                uint32_t bits = wv[k];
                const int word = k * kChunkWords + tid;
                while (bits) {
                    fl.list[p++] = (word << 5) + __builtin_ctz(bits);
                    bits &= bits - 1;
                }
                cpre += (int)((tot[k >> 2] >> (8 * (k & 3))) & 255u);
            }
            // one wave writes the list and publishes: one release (its
            // stores, then the stamp) and no further barrier
            __syncthreads();
            if (tid < kWave) {
                for (int i = tid; i < total; i += kWave)
                    ho->list[i] = fl.list[i];
                if (tid == 0)
                    publish_slide(d, st, sm, in, ho, total);
            }
            return;
        } else {
            for (int c = f; c < nch; c += F) {
                uint32_t wd = 0;
                int mid = 0; // bits of the chunks up to this part's next one
#pragma unroll
                for (int k = 0; k < kCompactRegs; ++k) {
                    wd = (k == c) ? wv[k] : wd;
                    mid += (k > c && k < c + F) ? __popc(wv[k]) : 0;
                }
                if ((nz >> c) & 1u) { // empty chunks are skipped (uniform)
                    const int n = emit_chunk<NT>(ho, fl, c, wd, P);
                    P += n;
                }
                const uint32_t between = (nz >> (c + 1)) & ((1u << (F - 1)) - 1u);
                if (F > 1 && c + F < nch && between) {
                    int m, unused;
                    uint32_t un2;
                    block_sum2<NT>(mid, 0, 0u, fl.red, m, unused, un2);
                    P += m;
                }
            }
        }
    } else {
        int before = 0, all = 0;
        for (int k = 0; k < nch; ++k) {
            const int wi = k * kChunkWords + tid;
            const int b = wi < nmine ? __popc(bm_load(wi)) : 0;
            all += b;
            before += (k < f) ? b : 0;
        }
        uint32_t unused;
        block_sum2<NT>(before, all, 0u, fl.red, P, total, unused);
        for (int c = f; c < nch; c += F) {
            const int wi = c * kChunkWords + tid;
            const int n = emit_chunk<NT>(ho, fl, c, wi < nmine ? bm_load(wi) : 0u, P);
            P += n;
            if (F > 1 && c + F < nch) {
                int mid = 0;
                for (int k = c + 1; k < c + F; ++k) {
                    const int wk = k * kChunkWords + tid;
                    mid += wk < nmine ? __popc(bm_load(wk)) : 0;
                }
                int m, un1;
                uint32_t un2;
                block_sum2<NT>(mid, 0, 0u, fl.red, m, un1, un2);
                P += m;
            }
        }
    }
    // every wave's list writes issued and acknowledged; a part of several
    // then releases them at system scope (a release only: no cache
    // invalidation under the window merge running beside it) before taking
    // its ticket.  (Acknowledged writes to host memory are not yet visible
    // to the host without the release.)
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (tid == 0) {
        bool last = true;
        if (!single) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
            last = atomicAdd(&d.ctl->fin_ticket, 1u) == (unsigned)F - 1;
        }
        fl.last = last;
    }
    __syncthreads();
    if (!fl.last || tid != 0)
        return;
    publish_slide(d, st, sm, in, ho, total);
}

// Part f of F of slide st's epilogue, run after its detect launch
// (pipelined slides: workgroups of the next slide's window-merge launch),
// from the counters the slide's classify and detect left in memory.
template <int NT>
__device__ __forceinline__ void finish_deferred(const Dev &d, const Step &st, int f, int F,
                                                FinishLds &fl, bool stamps = false)
{
    SlideSummary sm;
    sm.learn = st.learn != 0;
    const SlideCnt *cn = cnt_of(d, st.t);
    sm.n_o = cn->n_units;
    sm.n_units = sm.learn ? sm.n_o + cn->n_units_i : sm.n_o;
    sm.use_pred = sm.learn && d.lr.lc->dec[st.t % kDecSlots] != 0;
    sm.stamps = stamps;
    finish_slide<NT>(d, st, f, F, sm, fl);
}

// ---------------------------------------------------------------------------
// k_detect: one wavefront per cell unit (grid-stride over the unit list):
// the unit's candidates are counted together over their union stencil
// (unit_count).  Only ceil(n_units / 8) workgroups take part (the rest leave
// at once).  fin: the last of them to finish runs the slide epilogue
// (slide_epilogue: the outlier bitmap compacted into an ascending slot list
// written straight into pinned host memory); otherwise the next launch on
// the stream runs it (finish_deferred).
// LEARN: the list is the probable-outlier units [0, nO) followed by the
// probable-inlier units (indices W-1, W-2, ...): the long full counts are
// handed out first.  When the schedule is on (lc->dec of the slide), a probable-
// inlier unit (2-D/3-D) is first counted over its own cell and the
// neighbouring cells only; the full stencil is counted only if some
// candidate is still at <= K.  (In 1-D those cells are most of the stencil.)  Every count is exact; each unit's smallest count goes to the
// replay, and mispredicted units are counted for the fallback rule.
// ---------------------------------------------------------------------------
struct DetectLds {
    static constexpr int NW = kDetectNT / kWave;
    union { // the sample table, then (a batch's last workgroup) the epilogue's
        uint64_t samp[kMaxSamples];
        FinishLds fin;
    };
    int rows[NW][2][64];
    int last;
    int miss;
    int ready;
    unsigned long long tests;
    unsigned long long hits;
};

// Spins of a detect workgroup waiting for the classify of its own launch
// (at least 0.1 us each): past this the classify is taken to have failed.
// The environment variable RTOD_CLASSIFY_WAIT_SPINS, read when a context is
// created, replaces it (0 fails every such slide: a test of the error path).
constexpr int kClassifyWaitSpins = 1 << 20;

// bid / nblk: this workgroup and the number of workgroups of the detect
// phase of this stream.  WAIT: the slide's classify runs in the same launch
// in workgroups of lower index; thread 0 waits (bounded) for one of the
// slide's ready flags, and the counters and unit list entries the classify
// wrote are then read at agent scope (an acquire would invalidate this XCD's
// L2).
// The first `pre` workgroups stage the sample table before they wait (it
// comes from the window, not from the classify), so that its round trip
// overlaps the classify.
template <int DIM, bool LEARN, bool WAIT = false>
__device__ __forceinline__ void detect_body(const Dev &d, const Step &st, int bid, int nblk,
                                            bool fin, DetectLds &lds, int pre = 0)
{
    constexpr int NW = kDetectNT / kWave;
    uint64_t *s_samp = lds.samp;
    auto &s_rows = lds.rows;
    int &s_last = lds.last;
    unsigned long long &s_tests = lds.tests;
    unsigned long long &s_hits = lds.hits;
    int &s_miss = lds.miss;

    const int tid = threadIdx.x, lane = lane_id(), wid = tid / kWave;
    const int par = st.par ^ 1;
    const uint64_t *Nk = d.akey[par];
    const uint64_t *Nk4 = d.nsamp4 ? d.samples4[par] : nullptr;
    const int32_t *Ns = d.aslot[par];
    const float *Nx = d.axyz[par];
    const int W = d.W, K = d.K;
    SlideCnt *cn = cnt_of(d, st.t);
    const bool staged = WAIT && bid < pre;
    if (staged)
        copy_to_lds<kMaxSamples / kDetectNT>(s_samp, d.samples[par], d.nsamp);
    if constexpr (WAIT) {
        if (tid == 0) {
            const int32_t *flag = ready_of(d, st.t, bid % kReadyFlags);
            int ok = 0;
            for (int it = 0; it < d.wait_spins; ++it) {
                if (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                    ok = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(4);
            }
            if (!ok)
                atomicOr(&d.ctl->err, kErrClassifyWait);
            lds.ready = ok;
        }
        __syncthreads();
        if (!lds.ready)
            return;
    }
    auto ld_cnt = [&](const int32_t *p) {
        return WAIT ? __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : *p;
    };
    auto ld_unit = [&](int li) {
        if constexpr (WAIT) {
            const unsigned long long v =
                __hip_atomic_load(reinterpret_cast<const unsigned long long *>(d.units + li),
                                  __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return make_int2((int)(unsigned)v, (int)(unsigned)(v >> 32));
        } else {
            return d.units[li];
        }
    };
    const int n_o = ld_cnt(&cn->n_units);
    uint32_t *bm = bitmap_of(d, st.t);
    const int n_units = LEARN ? n_o + ld_cnt(&cn->n_units_i) : n_o;
    const bool use_pred = LEARN && d.lr.lc->dec[st.t % kDecSlots] != 0;
    const int busy = min(nblk, max(1, (n_units + NW - 1) / NW));
    if (bid >= busy)
        return;

    if (tid == 0) {
        s_tests = 0;
        s_hits = 0;
        s_miss = 0;
    }
    // a unit's list entry, then its candidates' coordinates and slots: the
    // first unit's are requested around the sample-table staging, each
    // next unit's entry while the current one is counted
    auto list_index = [&](int uu) { return (!LEARN || uu < n_o) ? uu : W - 1 - (uu - n_o); };
    const int ustep = busy * NW;
    int u = bid * NW + wid;
    int2 un = make_int2(0, 0);
    if (u < n_units)
        un = ld_unit(list_index(u));
    float pc[3] = {0.f, 0.f, 0.f};
    int32_t cslot = 0;
    auto load_candidates = [&](int2 e) {
        const int start = __builtin_amdgcn_readfirstlane(e.x);
        const int n = __builtin_amdgcn_readfirstlane(e.y);
        if (lane < n) {
#pragma unroll
            for (int a = 0; a < DIM; ++a)
                pc[a] = Nx[(size_t)a * W + start + lane];
            cslot = Ns[start + lane];
        }
    };
    if (!staged)
        copy_to_lds<kMaxSamples / kDetectNT>(s_samp, d.samples[par], d.nsamp);
    if (u < n_units)
        load_candidates(un);
    __syncthreads();

    long long tests = 0;
    int misses = 0;
    int hits = 0; // this lane's candidates' capped counts
    for (; u < n_units; u += ustep) {
        const int li = list_index(u);
        const int n = __builtin_amdgcn_readfirstlane(un.y);
        int2 un_next = make_int2(0, 0);
        if (u + ustep < n_units)
            un_next = ld_unit(list_index(u + ustep));
        const Stencil<DIM> full = unit_stencil<DIM>(d, pc, n);
        int cnt;
        if (LEARN && DIM > 1 && use_pred && u >= n_o) {
            cnt = unit_count<DIM>(d, Nk, Nk4, Nx, s_samp, s_rows[wid][0], s_rows[wid][1], pc, n,
                                  near_stencil<DIM>(d, full, pc), tests);
            if (__ballot(lane < n && cnt <= K) != 0) {
                wave_sync_lds();
                cnt = unit_count<DIM>(d, Nk, Nk4, Nx, s_samp, s_rows[wid][0], s_rows[wid][1], pc,
                                      n, full, tests);
            }
        } else {
            cnt = unit_count<DIM>(d, Nk, Nk4, Nx, s_samp, s_rows[wid][0], s_rows[wid][1], pc, n,
                                  full, tests);
        }
        const int capped = lane < n ? min(cnt, K + 1) : K + 1;
        hits += lane < n ? capped : 0;
        if (lane < n && cnt <= K)
            atomicOr(&bm[cslot >> 5], 1u << (cslot & 31));
        const int umin = unit_min_i(capped);
        if (lane == 0) {
            d.cand_cnt[u] = umin;
            if constexpr (LEARN) {
                d.lr.rep_c[(size_t)(st.t % kReplaySlides) * W + li] = umin;
                misses += ((u >= n_o) != (umin > K)) ? 1 : 0;
            }
        }
        wave_sync_lds();
        un = un_next;
        if (u + ustep < n_units)
            load_candidates(un);
    }
    if (lane == 0 && tests)
        atomicAdd(&s_tests, (unsigned long long)tests);
    if (LEARN && lane == 0 && misses)
        atomicAdd(&s_miss, misses);
    for (int dd = 1; dd < kWave; dd <<= 1)
        hits += __shfl_xor(hits, dd, kWave);
    if (lane == 0 && hits)
        atomicAdd(&s_hits, (unsigned long long)hits);
    __syncthreads();
    if (tid == 0) {
        if (s_tests)
            atomicAdd(&cn->pair_tests, s_tests);
        if (s_hits)
            atomicAdd(&cn->hits, s_hits);
        if (LEARN && s_miss)
            atomicAdd(&cn->misses, s_miss);
        if (fin) {
            __threadfence();
            s_last = (atomicAdd(&d.ctl->ticket, 1u) == (unsigned)busy - 1);
        }
    }
    if (!fin)
        return;
    __syncthreads();
    if (!s_last)
        return;
    __threadfence();
    finish_slide<kDetectNT>(d, st, 0, 1, SlideSummary{n_o, n_units, LEARN, use_pred, false},
                            lds.fin);
}

template <int DIM, bool LEARN>
__global__ __launch_bounds__(kDetectNT) RTOD_DETECT_OCCUPANCY void k_detect(Dev d, Step st,
                                                                           int fin)
{
    __shared__ DetectLds lds;
    stamp_start(d, kTsDetect);
    detect_body<DIM, LEARN>(d, st, blockIdx.x, gridDim.x, fin != 0, lds);
}

// ---------------------------------------------------------------------------
// Three launches per slide, for pipelined slides that cannot take the stage
// launches (the window merge of each slide on its own launch): slide t's
// window merge carries slide t-1's epilogue and slide t+1's tile sorts, slide
// t's detect slide t+1's run merge; it is also the prologue of the stage
// launches.
// The parts touch disjoint memory: the ingest writes batch-ring entry
// enew(t+1) and the tile scratch of slide t+1; slide t's merge reads entries
// enew(t) and eold(t) (distinct from enew(t+1) with m + 3 entries), writes
// the other window buffer and resets set t % 3 of the counters and
// bitmaps; slide t-1's epilogue reads set (t-1) % 3 and writes host-out
// buffer t-1.
// ---------------------------------------------------------------------------
// fin_prev > 0: workgroups [0, fin_prev) run slide st_prev's epilogue, in
// fin_prev parts; then nmb merge workgroups, then the tile sorts of st_next
// (as many as the grid has left)
template <int DIM, int IT>
__global__ __launch_bounds__(256) void k_merge_window_ingest(Dev d, Step st, Step st_next,
                                                            Step st_prev, int nmb, int fin_prev)
{
    __shared__ union {
        MergeWindowLds<256> mw;
        IngestLds<DIM, 256, IT> ing;
        FinishLds fin;
    } lds;
    int bid = blockIdx.x;
    if (bid < fin_prev) {
        finish_deferred<256>(d, st_prev, bid, fin_prev, lds.fin);
        return;
    }
    bid -= fin_prev;
    if (bid < nmb)
        merge_window_body<DIM, 256>(d, st, bid, lds.mw);
    else
        ingest_sort_body<DIM, 256, IT>(d, st_next, kModeSlide, nullptr, d.cpb == 1 ? 1 : 0,
                                       bid - nmb, lds.ing);
}

// slide st's epilogue in a launch of its own, one part per workgroup (a
// pipelined run stopped before the slide that would carry it)
__global__ __launch_bounds__(256) void k_finish(Dev d, Step st, int stamps)
{
    __shared__ FinishLds lds;
    if (stamps)
        stamp_start(d, kTsFinish);
    finish_deferred<256>(d, st, blockIdx.x, gridDim.x, lds, stamps != 0);
}

static_assert(kRunMergeNT == kDetectNT, "the run merge rides in detect workgroups");

// LDS of k_detect_runs, set at launch: the detect's, or the next slide's
// S keys when larger (dynamic, so a 5000-point slide takes 40 KB, not the
// 64 KB of the largest slide, and three workgroups still fit a CU)
__host__ __device__ constexpr size_t detect_runs_lds(int S)
{
    return sizeof(DetectLds) > (size_t)S * sizeof(uint64_t) ? sizeof(DetectLds)
                                                             : (size_t)S * sizeof(uint64_t);
}

template <int DIM, bool LEARN>
__global__ __launch_bounds__(kDetectNT) RTOD_DETECT_OCCUPANCY void k_detect_runs(Dev d, Step st,
                                                                                Step st_next,
                                                                                int ndet)
{
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    if ((int)blockIdx.x < ndet)
        detect_body<DIM, LEARN>(d, st, blockIdx.x, ndet, false,
                                *reinterpret_cast<DetectLds *>(dyn_lds));
    else
        merge_runs_body<DIM>(d, st_next, kModeSlide, blockIdx.x - ndet,
                             reinterpret_cast<uint64_t *>(dyn_lds));
}

// ---------------------------------------------------------------------------
// Pipelined slides on one stream (cell-stencil index, the ingest riding):
// the detect stage of slide t carries the epilogue of slide t - 1, the
// window merge of slide t + 1, the run merge of slide t + 2 and the tile
// sorts of slide t + 3; the classify of slide t is a launch of its own
// before it (learned schedule: k_classify_stage) or its first workgroups
// (plain schedule: k_slide_stage).  The parts touch disjoint memory
// (counters, bitmaps and ready flags come in kSlideSets = 3 sets, by slide
// modulo 3):
//   classify(t): reads window t; writes the candidate and unit lists,
//     counters cnt[t % 3] and (one launch) ready flags t % 3;
//   epilogue(t - 1): reads bitmap (t - 1) % 3 and cnt[(t - 1) % 3]; writes
//     host-out buffer t - 1;
//   detect(t): reads window t, the lists and cnt[t % 3]; sets bitmap t % 3;
//   merge(t + 1): reads window t and batch-ring entries enew(t + 1),
//     eold(t + 1); writes window t + 1 (the other buffers); resets
//     cnt[(t + 1) % 3] and ready flags (t + 1) % 3 and clears bitmap
//     (t + 1) % 3 (last read by epilogue(t - 2), a detect stage earlier);
//   run merge(t + 2): reads tile scratch (t + 2) & 1; writes ring entry
//     enew(t + 2);
//   tile sorts(t + 3): write tile scratch (t + 3) & 1 (one tile: ring entry
//     enew(t + 3)).
// With m + 3 ring entries, enew(t + 1), eold(t + 1), enew(t + 2) and
// enew(t + 3) are t - 2, t + 1, t - 1 and t modulo m + 3: all distinct.
// ---------------------------------------------------------------------------
template <int DIM, int MODE>
__global__ __launch_bounds__(kClassifyNT) void k_classify_stage(Dev d, Step st)
{
    __shared__ ClassifyLds<MODE == kClsLearned> lds;
    classify_body<DIM, MODE>(d, st, blockIdx.x, lds);
}

// LDS of k_detect_stage, set at launch: the largest of its parts'
template <int DIM, int TSIT>
__host__ __device__ constexpr size_t detect_stage_lds(bool fin, bool merge, bool runs, bool ts,
                                                      int S)
{
    size_t b = sizeof(DetectLds);
    if (fin && sizeof(FinishLds) > b)
        b = sizeof(FinishLds);
    if (merge && sizeof(MergeWindowLds<kDetectNT>) > b)
        b = sizeof(MergeWindowLds<kDetectNT>);
    if (runs && (size_t)S * sizeof(uint64_t) > b)
        b = (size_t)S * sizeof(uint64_t);
    if (ts && sizeof(IngestLds<DIM, kDetectNT, TSIT>) > b)
        b = sizeof(IngestLds<DIM, kDetectNT, TSIT>);
    return b;
}

// workgroups: nfin epilogue parts (slide st_prev), nmb window-merge (slide
// st_merge), nrm run-merge (slide st_runs), nts tile-sort (slide st_ts,
// tiles of kDetectNT * TSIT points), then the ndet of the detect (slide
// st), whose busy ones come first
template <int DIM, bool LEARN, int TSIT>
__global__ __launch_bounds__(kDetectNT) RTOD_DETECT_OCCUPANCY void k_detect_stage(
    Dev d, Step st, Step st_prev, Step st_merge, Step st_runs, Step st_ts, int nfin, int nmb,
    int nrm, int nts, int ndet)
{
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    int bid = blockIdx.x;
    if (bid < nfin) {
        finish_deferred<kDetectNT>(d, st_prev, bid, nfin, *reinterpret_cast<FinishLds *>(dyn_lds));
        return;
    }
    bid -= nfin;
    if (bid < nmb) {
        merge_window_body<DIM, kDetectNT>(d, st_merge, bid,
                                          *reinterpret_cast<MergeWindowLds<kDetectNT> *>(dyn_lds));
        return;
    }
    bid -= nmb;
    if (bid < nrm) {
        merge_runs_body<DIM>(d, st_runs, kModeSlide, bid, reinterpret_cast<uint64_t *>(dyn_lds));
        return;
    }
    bid -= nrm;
    if (bid < nts) {
        ingest_sort_body<DIM, kDetectNT, TSIT>(
            d, st_ts, kModeSlide, nullptr, d.cpb == 1 ? 1 : 0, bid,
            *reinterpret_cast<IngestLds<DIM, kDetectNT, TSIT> *>(dyn_lds));
        return;
    }
    bid -= nts;
    detect_body<DIM, LEARN>(d, st, bid, ndet, false, *reinterpret_cast<DetectLds *>(dyn_lds));
}

// ---------------------------------------------------------------------------
// One launch per slide (plain schedule, cell-stencil index, the ingest
// riding): k_detect_stage's parts behind ncls workgroups of slide st's
// classify, which its detect waits for inside the launch.  The classify
// workgroups have the lowest indices: workgroups are dispatched in index
// order, and a classify workgroup waits for nothing, so every classify
// workgroup gets a slot while detect workgroups wait.  The wait is bounded
// (kClassifyWaitSpins): past it the detect raises kErrClassifyWait and
// leaves, so the slide fails instead of the launch hanging.  The classify
// writes slide st's counters and unit list only; no other part of the
// launch reads them.
// ---------------------------------------------------------------------------
template <int DIM, int TSIT>
__host__ __device__ constexpr size_t slide_stage_lds(bool fin, bool merge, bool runs, bool ts,
                                                     int S)
{
    const size_t b = detect_stage_lds<DIM, TSIT>(fin, merge, runs, ts, S);
    return sizeof(ClassifyLds<false, kDetectNT>) > b ? sizeof(ClassifyLds<false, kDetectNT>) : b;
}

template <int DIM, int TSIT>
__global__ __launch_bounds__(kDetectNT) RTOD_DETECT_OCCUPANCY void k_slide_stage(
    Dev d, Step st, Step st_prev, Step st_merge, Step st_runs, Step st_ts, int ncls, int nfin,
    int nmb, int nrm, int nts, int ndet, int pre)
{
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    int bid = blockIdx.x;
    if (bid < ncls) {
        classify_body<DIM, kClsUnits, kDetectNT, true>(
            d, st, bid, *reinterpret_cast<ClassifyLds<false, kDetectNT> *>(dyn_lds));
        return;
    }
    bid -= ncls;
    if (bid < nfin) {
        finish_deferred<kDetectNT>(d, st_prev, bid, nfin, *reinterpret_cast<FinishLds *>(dyn_lds));
        return;
    }
    bid -= nfin;
    if (bid < nmb) {
        merge_window_body<DIM, kDetectNT>(d, st_merge, bid,
                                          *reinterpret_cast<MergeWindowLds<kDetectNT> *>(dyn_lds));
        return;
    }
    bid -= nmb;
    if (bid < nrm) {
        merge_runs_body<DIM>(d, st_runs, kModeSlide, bid, reinterpret_cast<uint64_t *>(dyn_lds));
        return;
    }
    bid -= nrm;
    if (bid < nts) {
        ingest_sort_body<DIM, kDetectNT, TSIT>(
            d, st_ts, kModeSlide, nullptr, d.cpb == 1 ? 1 : 0, bid,
            *reinterpret_cast<IngestLds<DIM, kDetectNT, TSIT> *>(dyn_lds));
        return;
    }
    bid -= nts;
    detect_body<DIM, false, true>(d, st, bid, ndet, false, *reinterpret_cast<DetectLds *>(dyn_lds),
                                  pre);
}


} // namespace rtod
