// rtod_plain.h -- the plain (cell-stencil) per-slide path.
//
// One slide = five device phases, all state in HBM, one host round trip:
//   k_ingest_sort   new slide from pinned staging into the ring; cell keys;
//                   LDS-staged stable radix sort of the batch by key
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

#include "rtod_device.h"
#include "rtod_state.h"

namespace rtod {

enum SortMode { kModeSlide = 0, kModePrime = 1 };

// batch-ring entry of the batch written by the current slide (generation m+t)
__device__ __forceinline__ int entry_new(const Dev &d, uint64_t t)
{
    return (int)((uint64_t(d.m) + t) % uint64_t(d.m + 1));
}
// batch-ring entry of the batch expiring at this slide (generation t)
__device__ __forceinline__ int entry_old(const Dev &d, uint64_t t)
{
    return (int)(t % uint64_t(d.m + 1));
}

// ---------------------------------------------------------------------------
// k_ingest_sort: one workgroup per LDS tile of CH = NT*IT points.
// Slide mode: points come from the pinned staging buffer of parity t, go to
// ring slots (t mod m)*S + j.  Prime mode: segment b of the initial window
// (slots b*S + j) comes from a device buffer.
// Stable LSD sort over the key bits that vary inside the tile, one split
// pass per bit (block scan of the zero flags); ties keep arrival order.
// ---------------------------------------------------------------------------
template <int DIM, int NT, int IT>
__global__ __launch_bounds__(NT) void k_ingest_sort(Dev d, int mode, const float *src_dev,
                                                    int dst_batch)
{
    constexpr int CH = NT * IT;
    __shared__ uint64_t s_key[CH];
    __shared__ int32_t s_val[CH];
    __shared__ int s_scan[NT / kWave];
    __shared__ uint64_t s_red[2][NT / kWave];

    const int tid = threadIdx.x;
    const int seg = blockIdx.x / d.cpb;
    const int chunk = blockIdx.x % d.cpb;
    const uint64_t t = (mode == kModeSlide) ? d.ctl->t : 0;
    const int slot0 = (mode == kModeSlide) ? (int)(t % uint64_t(d.m)) * d.S : seg * d.S;
    const float *src = (mode == kModeSlide) ? d.h_in + (size_t)(t & 1) * d.S * DIM
                                            : src_dev + (size_t)seg * d.S * DIM;
    const int j0 = chunk * CH;
    const int n = min(CH, d.S - j0);

    // stage 1: coalesced loads, ring update, keys into LDS
    uint64_t lor = 0, land = ~0ull;
    for (int l = tid; l < CH; l += NT) {
        uint64_t key = ~0ull;
        if (l < n) {
            const int j = j0 + l;
            float p[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int a = 0; a < DIM; ++a)
                p[a] = src[(size_t)j * DIM + a];
#pragma unroll
            for (int a = 0; a < DIM; ++a)
                d.ring[(size_t)a * d.W + slot0 + j] = p[a];
            key = cell_key<DIM>(p, d.inv_c);
            lor |= key;
            land &= key;
        }
        s_key[l] = key;
    }
    lor = wave_or(lor);
    land = wave_and(land);
    if (lane_id() == 0) {
        s_red[0][tid / kWave] = lor;
        s_red[1][tid / kWave] = land;
    }
    __syncthreads();
    uint64_t vary_or = 0, vary_and = ~0ull;
#pragma unroll
    for (int w = 0; w < NT / kWave; ++w) {
        vary_or |= s_red[0][w];
        vary_and &= s_red[1][w];
    }
    uint64_t varying = (n > 0) ? (vary_or ^ vary_and) : 0;

    // stage 2: blocked arrangement in registers
    uint64_t k[IT];
    int32_t v[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        k[i] = s_key[tid * IT + i];
        v[i] = tid * IT + i;
    }
    __syncthreads();

    while (varying) {
        const int bit = __builtin_ctzll(varying);
        varying &= varying - 1;
        int zeros = 0;
#pragma unroll
        for (int i = 0; i < IT; ++i)
            zeros += ((k[i] >> bit) & 1ull) ? 0 : 1;
        int total_z;
        int zp = block_excl_scan<NT>(zeros, s_scan, total_z);
        int op = total_z + tid * IT - zp;
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            int pos = ((k[i] >> bit) & 1ull) ? op++ : zp++;
            s_key[pos] = k[i];
            s_val[pos] = v[i];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            k[i] = s_key[tid * IT + i];
            v[i] = s_val[tid * IT + i];
        }
        __syncthreads();
    }

    // output (padding keys sorted last and dropped)
    size_t base;
    if (dst_batch) {
        int e = (mode == kModeSlide) ? entry_new(d, t) : seg;
        base = (size_t)e * d.S + j0;
    } else {
        base = (size_t)seg * d.S + j0;
    }
    uint64_t *ok = dst_batch ? d.bkey : d.tkey[0];
    int32_t *os = dst_batch ? d.bslot : d.tslot[0];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        int pos = tid * IT + i;
        if (pos < n) {
            ok[base + pos] = k[i];
            os[base + pos] = slot0 + j0 + v[i];
        }
    }
}

// ---------------------------------------------------------------------------
// k_merge_pass: within each segment of seg_len elements, merge sorted runs of
// length L pairwise into runs of 2L (stable: the earlier run wins ties).
// dst_batch: write segment s of the result into its batch-ring entry.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_merge_pass(Dev d, const uint64_t *sk, const int32_t *ss,
                                                    uint64_t *dk, int32_t *ds, int n_total,
                                                    int seg_len, int L, int dst_batch, int mode)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
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
        int e = (mode == kModeSlide) ? entry_new(d, d.ctl->t) : seg;
        o = (size_t)e * d.S + out;
    } else {
        o = (size_t)base + out;
    }
    dk[o] = key;
    ds[o] = ss[i];
}

// gather coordinates of a sorted (key, slot) array from the ring
template <int DIM>
__global__ __launch_bounds__(256) void k_gather(Dev d, int par)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.W)
        return;
    const int s = d.aslot[par][i];
#pragma unroll
    for (int a = 0; a < DIM; ++a)
        d.axyz[par][(size_t)a * d.W + i] = d.ring[(size_t)a * d.W + s];
    if ((i & ((1 << d.stride_log2) - 1)) == 0)
        d.samples[i >> d.stride_log2] = d.akey[par][i];
}

// ---------------------------------------------------------------------------
// k_merge_window: A' = (A - E) + B, all sorted by (key, arrival).
//   A  current sorted window, E the expiring batch (a subset of A, sorted by
//   key then slot), B the new batch sorted by key.
//   For a survivor a at index i: pos = i - #E before a + #B with key < key_a.
//   For b at index j in B:        pos = j + #A with key <= k_b - #E with key <= k_b.
// Workgroup c owns the A chunk [i0, i1) and the b with
// key(A[i0]) <= k_b < key(A[i1]); six leader threads narrow every search
// range once, the chunk's keys sit in LDS.  Output position multiples of the
// sample stride also fill the sample table of the cell-stencil index.
// ---------------------------------------------------------------------------
template <int DIM>
__global__ __launch_bounds__(256) void k_merge_window(Dev d)
{
    constexpr int NT = 256, IT = 4, T = NT * IT;
    __shared__ uint64_t s_k[T];
    __shared__ int s_scan[NT / kWave];
    __shared__ int s_par[8];

    const int tid = threadIdx.x;
    const uint64_t t = d.ctl->t;
    const int W = d.W, S = d.S;
    const int par = (int)(t & 1);
    const uint64_t *Ak = d.akey[par];
    const int32_t *As = d.aslot[par];
    const float *Ax = d.axyz[par];
    uint64_t *Nk = d.akey[par ^ 1];
    int32_t *Ns = d.aslot[par ^ 1];
    float *Nx = d.axyz[par ^ 1];
    const size_t eo = (size_t)entry_old(d, t) * S, en = (size_t)entry_new(d, t) * S;
    const uint64_t *Ek = d.bkey + eo;
    const int32_t *Es = d.bslot + eo;
    const uint64_t *Bk = d.bkey + en;
    const int32_t *Bs = d.bslot + en;
    const int exp_lo = (int)(t % uint64_t(d.m)) * S;
    const int i0 = blockIdx.x * T;
    const int i1 = min(W, i0 + T);
    const int n = i1 - i0;
    const bool first = blockIdx.x == 0;
    const bool last = i1 == W;
    const int smask = (1 << d.stride_log2) - 1;

    if (blockIdx.x == 0 && tid == 0)
        d.ctl->n_cand = 0;

    for (int l = tid; l < n; l += NT)
        s_k[l] = Ak[i0 + l];
    if (tid < 6) {
        const uint64_t k0 = Ak[i0];
        int r = 0;
        if (tid == 0) {
            const int s0 = As[i0];
            const bool ex = s0 >= exp_lo && s0 < exp_lo + S;
            r = ex ? lower_bound_pair(Ek, Es, 0, S, k0, s0) : upper_bound_u64(Ek, 0, S, k0);
        } else if (tid == 1) {
            r = lower_bound_u64(Bk, 0, S, k0);
        } else if (tid == 2) {
            r = lower_bound_u64(Bk, 0, S, Ak[i1 - 1]);
        } else if (tid == 3) {
            r = last ? S : lower_bound_u64(Bk, 0, S, Ak[i1]);
        } else if (tid == 4) {
            r = first ? 0 : lower_bound_u64(Ek, 0, S, k0);
        } else {
            r = last ? S : lower_bound_u64(Ek, 0, S, Ak[i1]);
        }
        s_par[tid] = r;
    }
    __syncthreads();
    const int e_before = s_par[0];
    const int b_lo = s_par[1], b_hi = s_par[2];
    const int own_lo = first ? 0 : b_lo, own_hi = s_par[3];
    const int e_lo = s_par[4], e_hi = s_par[5];

    // survivors of A (blocked items so the expired prefix is a block scan)
    int32_t sl[IT];
    bool ex[IT];
    int nex = 0;
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int l = tid * IT + i;
        ex[i] = false;
        sl[i] = 0;
        if (l < n) {
            sl[i] = As[i0 + l];
            ex[i] = sl[i] >= exp_lo && sl[i] < exp_lo + S;
            nex += ex[i] ? 1 : 0;
        }
    }
    int tot;
    int eb = e_before + block_excl_scan<NT>(nex, s_scan, tot);
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int l = tid * IT + i;
        if (l >= n)
            continue;
        if (ex[i]) {
            ++eb;
            continue;
        }
        const uint64_t key = s_k[l];
        const int rb = lower_bound_u64(Bk, b_lo, b_hi, key);
        const int np = i0 + l - eb + rb;
        if ((unsigned)np >= (unsigned)W) { // impossible unless state is corrupt
            atomicOr(&d.ctl->err, 1);
            continue;
        }
        Nk[np] = key;
        Ns[np] = sl[i];
#pragma unroll
        for (int a = 0; a < DIM; ++a)
            Nx[(size_t)a * W + np] = Ax[(size_t)a * W + i0 + l];
        if ((np & smask) == 0)
            d.samples[np >> d.stride_log2] = key;
    }

    // new points owned by this chunk
    for (int j = own_lo + tid; j < own_hi; j += NT) {
        const uint64_t key = Bk[j];
        const int32_t s = Bs[j];
        int lo = 0, hi = n;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (s_k[mid] <= key)
                lo = mid + 1;
            else
                hi = mid;
        }
        const int ubA = i0 + lo;
        const int ubE = upper_bound_u64(Ek, e_lo, e_hi, key);
        const int np = j + ubA - ubE;
        if ((unsigned)np >= (unsigned)W) {
            atomicOr(&d.ctl->err, 2);
            continue;
        }
        Nk[np] = key;
        Ns[np] = s;
#pragma unroll
        for (int a = 0; a < DIM; ++a)
            Nx[(size_t)a * W + np] = d.ring[(size_t)a * W + s];
        if ((np & smask) == 0)
            d.samples[np >> d.stride_log2] = key;
    }
}

// ---------------------------------------------------------------------------
// k_classify: a point is fast-accepted iff its cell is unclamped and holds
// >= K+1 points (reference: cells with num > K are skipped, optixScan.cpp:
// 987-993).  The cell is a run of equal keys in the sorted window: find the
// run start s within [i-K, i] and test key[s+K] == key.  Other points are
// appended to the candidate list, one atomic per wavefront.
// ---------------------------------------------------------------------------
template <int DIM>
__global__ __launch_bounds__(256) void k_classify(Dev d)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t t = d.ctl->t;
    const uint64_t *Nk = d.akey[(t & 1) ^ 1];
    const int W = d.W, K = d.K;
    bool cand = false;
    if (i < W) {
        const uint64_t key = Nk[i];
        bool dense = false;
        if (K < W && key_guaranteed<DIM>(key)) {
            int lo = max(0, i - K), hi = i;
            while (lo < hi) {
                int mid = (lo + hi) >> 1;
                if (Nk[mid] == key)
                    hi = mid;
                else
                    lo = mid + 1;
            }
            dense = (lo + K < W) && (Nk[lo + K] == key);
        }
        cand = !dense;
    }
    const uint64_t mask = __ballot(cand);
    const int cnt = __popcll(mask);
    if (cnt) {
        int base = 0;
        if (lane_id() == 0)
            base = atomicAdd(&d.ctl->n_cand, cnt);
        base = __shfl(base, 0, kWave);
        if (cand) {
            const uint64_t below = mask & ((1ull << lane_id()) - 1ull);
            d.cand[base + __popcll(below)] = i;
        }
    }
}

// ---------------------------------------------------------------------------
// Stencil of a candidate: per axis the cells overlapping [x - R', x + R'],
// R' = R (1 + 1e-5) > any |dx| the FP32 predicate accepts.  Rows are
// (k0[,k1]) fixed with the last axis spanning a contiguous key range.
// ---------------------------------------------------------------------------
template <int DIM> struct Stencil {
    int64_t lo[3], hi[3];
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

constexpr int kDetectNT = 256;
constexpr int kMaxSamples = 4096;

// Exact neighbour count of the candidate at sorted position p, capped: stops
// as soon as the count exceeds K.  Wave-cooperative (all 64 lanes call it
// with the same p).  Returns min(count, first multiple-of-64 chunk where it
// exceeded K); *tests = predicate evaluations.
template <int DIM>
__device__ __forceinline__ int range_count(const Dev &d, const uint64_t *Nk, const float *Nx,
                                           const uint64_t *s_samp, int *w_start, int *w_off,
                                           int p, long long &tests)
{
    const int lane = lane_id();
    const int W = d.W, K = d.K;
    float c[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < DIM; ++a)
        c[a] = Nx[(size_t)a * W + p];
    Stencil<DIM> st;
    make_stencil<DIM>(c, d.inv_c, d.Rp, st);
    if (st.nrows > 32) { // cannot happen: <= 5 cells per axis
        if (lane == 0)
            atomicOr(&d.ctl->err, 4);
        st.nrows = 32;
    }

    // row bounds: lanes [0,nrows) lower bounds, lanes [32,32+nrows) upper
    const int r = lane & 31;
    int bnd = 0;
    if (r < st.nrows) {
        uint64_t klo, khi;
        row_keys<DIM>(st, r, klo, khi);
        bnd = (lane < 32) ? bound2<false>(Nk, W, s_samp, d.nsamp, d.stride_log2, klo)
                          : bound2<true>(Nk, W, s_samp, d.nsamp, d.stride_log2, khi);
    }
    const int ub = __shfl(bnd, r + 32, kWave);
    int len = (lane < 32 && r < st.nrows) ? (ub - bnd) : 0;
    const int incl = wave_incl_scan(len);
    const int total = __shfl(incl, 63, kWave);
    if (total <= K) // even counting every stencil point it stays <= K
        return total;
    if (lane < 32) {
        w_start[lane] = bnd;
        w_off[lane] = incl - len;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    int cnt = 0;
    for (int base = 0; base < total; base += kWave) {
        const int idx = base + lane;
        bool hit = false;
        if (idx < total) {
            int j;
            if constexpr (DIM == 1) {
                j = w_start[0] + idx;
            } else {
                int lo = 0, hi = st.nrows - 1;
                while (lo < hi) { // last row with off <= idx
                    int mid = (lo + hi + 1) >> 1;
                    if (w_off[mid] <= idx)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                j = w_start[lo] + (idx - w_off[lo]);
            }
            const float qx = Nx[j];
            const float qy = DIM > 1 ? Nx[(size_t)W + j] : 0.f;
            const float qz = DIM > 2 ? Nx[(size_t)2 * W + j] : 0.f;
            hit = dist2<DIM>(c, qx, qy, qz) <= d.r2;
        }
        cnt += __popcll(__ballot(hit));
        tests += min(kWave, total - base);
        if (cnt > K)
            break;
    }
    return cnt;
}

// ---------------------------------------------------------------------------
// k_detect: one wavefront per candidate (grid-stride over the candidate
// list), then the last workgroup to finish compacts the bitmap.
// ---------------------------------------------------------------------------
template <int DIM>
__global__ __launch_bounds__(kDetectNT) void k_detect(Dev d)
{
    constexpr int NW = kDetectNT / kWave;
    __shared__ uint64_t s_samp[kMaxSamples];
    __shared__ int s_rows[NW][2][32];
    __shared__ int s_scan[NW];
    __shared__ int s_last;
    __shared__ unsigned long long s_tests;

    const int tid = threadIdx.x, lane = lane_id(), wid = tid / kWave;
    const uint64_t t = d.ctl->t;
    const int par = (int)(t & 1) ^ 1;
    const uint64_t *Nk = d.akey[par];
    const int32_t *Ns = d.aslot[par];
    const float *Nx = d.axyz[par];
    const int n_cand = d.ctl->n_cand;
    const int K = d.K;

    if (tid == 0)
        s_tests = 0;
    const bool busy = blockIdx.x * NW < n_cand;
    if (busy)
        for (int i = tid; i < d.nsamp; i += kDetectNT)
            s_samp[i] = d.samples[i];
    __syncthreads();

    long long tests = 0;
    if (busy) {
        for (int ci = blockIdx.x * NW + wid; ci < n_cand; ci += gridDim.x * NW) {
            const int p = __builtin_amdgcn_readfirstlane(d.cand[ci]);
            const int cnt = range_count<DIM>(d, Nk, Nx, s_samp, s_rows[wid][0], s_rows[wid][1], p,
                                             tests);
            if (lane == 0) {
                d.cand_cnt[ci] = min(cnt, K + 1);
                if (cnt <= K) {
                    const int slot = Ns[p];
                    atomicOr(&d.bitmap[slot >> 5], 1u << (slot & 31));
                }
            }
        }
    }
    if (lane == 0 && tests)
        atomicAdd(&s_tests, (unsigned long long)tests);
    __syncthreads();
    if (tid == 0) {
        if (s_tests)
            atomicAdd(&d.ctl->pair_tests, s_tests);
        __threadfence();
        s_last = (atomicAdd(&d.ctl->ticket, 1u) == gridDim.x - 1);
    }
    __syncthreads();
    if (!s_last)
        return;
    __threadfence();

    // last workgroup: ascending slot list from the bitmap, bitmap cleared
    const int nwords = (d.W + 31) >> 5;
    const int per = (nwords + kDetectNT - 1) / kDetectNT;
    const int w0 = min(nwords, tid * per), w1 = min(nwords, w0 + per);
    int mine = 0;
    for (int w = w0; w < w1; ++w)
        mine += __popc(__hip_atomic_load(&d.bitmap[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    int total;
    int pos = block_excl_scan<kDetectNT>(mine, s_scan, total);
    HostOut *ho = d.h_out;
    for (int w = w0; w < w1; ++w) {
        uint32_t bits = atomicExch(&d.bitmap[w], 0u);
        while (bits) {
            const int b = __builtin_ctz(bits);
            bits &= bits - 1;
            ho->list[pos++] = (w << 5) + b;
        }
    }
    __syncthreads();
    if (tid == 0) {
        const unsigned long long pt =
            __hip_atomic_load(&d.ctl->pair_tests, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ho->n_out = total;
        ho->n_cand = n_cand;
        ho->pair_tests = pt;
        ho->err = __hip_atomic_load(&d.ctl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ho->slide = (int64_t)(t + 1);
        d.ctl->err = 0;
        d.ctl->pair_tests = 0;
        d.ctl->ticket = 0;
        d.ctl->t = t + 1;
        __threadfence_system();
    }
}

} // namespace rtod
