// rtod_lbvh.h -- LBVH index over the candidates (RTOD_INDEX_LBVH).
//
// The reference rebuilds an OptiX GAS over one box [c-R, c+R] per candidate
// every slide (kGenAABB + optixAccelBuild, aabb.cu:5-45, optixScan.cpp:
// 378-426) and casts one zero-length ray from every window point; each box a
// ray starts in runs the exact test and atomically bumps the candidate's
// count, skipping candidates already above K (optixScan.cu:158-199).  This
// index does the same on MI355X:
//   k_lbvh_codes   Morton code of each candidate from its cell fields
//   k_lbvh_sort    stable LDS tile sort of (code, candidate) pairs
//   k_lbvh_merge   pairwise merge of sorted tiles (when U > one tile)
//   k_lbvh_build   Karras radix-tree hierarchy (one thread per internal node)
//   k_lbvh_boxes   leaf boxes rounded outward, unions bottom-up (second
//                  arrival at a node continues)
//   k_lbvh_query   groups of 64 sorted window points walk the tree together
//                  (one wavefront, 64 nodes per step); per leaf batch:
//                  early-out if the count already exceeds K, exact FP32
//                  test, atomicAdd -- counts stay exact up to K+1
//   k_lbvh_select  count <= K -> outlier bit (the slide epilogue, k_finish,
//                  lists them)
// The number of candidates U lives on the device; kernels are sized for W
// and exit past U.
// Ablations (the reference's OPTIMIZATION 0/1 and UPDATE_GAS_TYPE 0): a tree
// over the whole window (leaf items = ring slots, U = W), queried from the
// candidates or from every slot by k_lbvh_count, rebuilt each slide or built
// once at prime and refit (k_lbvh_unflag + k_lbvh_boxes).
#pragma once

#include "rtod_plain.h"

namespace rtod {

// device arrays of the index (allocated for RTOD_INDEX_LBVH contexts)
struct Lbvh {
    uint64_t *code[2]; // sorted Morton codes (ping-pong for merges)
    int32_t *item[2];  // candidate index of each sorted code
    int32_t *left, *right; // internal node children: >= 0 internal, < 0 leaf ~i
    int32_t *parent;   // [2U]: internal nodes 0..U-2, leaves at U-1+i
    float *box;        // [2U][6] lo xyz, hi xyz (internal then leaves)
    int32_t *flag;     // [U] bottom-up arrival counters
    int32_t *count;    // [U] neighbour counts per sorted leaf
    int ntiles_log2;   // merge passes for U up to W
    const int32_t *n_cand; // leaves U: the slide's candidate count (set per
                           // launch), or n_whole
    int whole;         // the tree holds the whole window: leaf items are ring
                       // slots (ablations FILTER_ONLY / NO_FILTER)
    int32_t *n_whole;  // [1] = W, the leaf count of a whole-window tree
};

// Morton interleave of up to three 21-bit (3-D) or two 31-bit (2-D) fields
__device__ __forceinline__ uint64_t spread3(uint64_t x)
{
    x &= 0x1fffffull;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

__device__ __forceinline__ uint64_t spread2(uint64_t x)
{
    x &= 0x7fffffffull;
    x = (x | x << 16) & 0x0000ffff0000ffffull;
    x = (x | x << 8) & 0x00ff00ff00ff00ffull;
    x = (x | x << 4) & 0x0f0f0f0f0f0f0f0full;
    x = (x | x << 2) & 0x3333333333333333ull;
    x = (x | x << 1) & 0x5555555555555555ull;
    return x;
}

// fields relative to org (a lower bound of the window's fields; 0: raw)
template <int DIM>
__device__ __forceinline__ uint64_t morton_of_key(uint64_t key, const uint64_t (&org)[3])
{
    uint64_t f[3];
    key_fields<DIM>(key, f);
#pragma unroll
    for (int a = 0; a < DIM; ++a)
        f[a] = f[a] >= org[a] ? f[a] - org[a] : 0;
    if constexpr (DIM == 1)
        return f[0];
    else if constexpr (DIM == 2)
        return (spread2(f[0]) << 1) | spread2(f[1]);
    else
        return (spread3(f[0]) << 2) | (spread3(f[1]) << 1) | spread3(f[2]);
}

template <int DIM>
__global__ __launch_bounds__(256) void k_lbvh_codes(Dev d, Step st, Lbvh L)
{
    stamp_start(d, kTsIndex);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int U = *L.n_cand;
    if (i >= U)
        return;
    // the slide's code origin (classify); unset (~0, or 0 at prime): raw fields
    const SlideCnt *cn = cnt_of(d, st.t);
    uint64_t org[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
        org[a] = cn->fmin[a] == ~0ull ? 0 : cn->fmin[a];
    if (L.whole) { // leaf i: ring slot i, the point's cell from its coordinates
        float p[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int a = 0; a < DIM; ++a)
            p[a] = d.ring[(size_t)a * d.W + i];
        L.code[0][i] = morton_of_key<DIM>(cell_key<DIM>(p, d.inv_c), org);
    } else {
        const uint64_t *Nk = d.akey[st.par ^ 1];
        L.code[0][i] = morton_of_key<DIM>(Nk[d.cand[i]], org);
    }
    L.item[0][i] = i;
}

// Stable sort of one tile of CH (code, item) pairs: relative codes packed
// with the tile index, two LSD stages when the codes span more than 51 bits.
template <int NT, int IT>
__global__ __launch_bounds__(NT) void k_lbvh_sort(Dev d, Lbvh L)
{
    constexpr int CH = NT * IT, NW = NT / kWave, IB = 13, KB = 64 - IB;
    constexpr uint64_t IMASK = (1ull << IB) - 1ull;
    static_assert(CH <= (1 << IB), "tile index bits");
    __shared__ uint64_t s_key[CH];
    __shared__ uint64_t s_code[CH];
    __shared__ uint16_t s_val[CH];
    __shared__ int s_hist[16 * NW];
    __shared__ uint64_t s_red[2][NW];
    const int U = *L.n_cand;
    const int j0 = blockIdx.x * CH;
    if (j0 >= U)
        return;
    const int n = min(CH, U - j0);
    const int tid = threadIdx.x, lane = lane_id(), wid = tid / kWave;
    uint64_t k[IT];
    uint64_t mn = ~0ull, mx = 0;
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int l = wid * kWave * IT + i * kWave + lane;
        k[i] = ~0ull;
        if (l < n) {
            k[i] = L.code[0][j0 + l];
            mn = k[i] < mn ? k[i] : mn;
            mx = k[i] > mx ? k[i] : mx;
        }
        s_code[l] = k[i];
    }
    for (int dd = 1; dd < kWave; dd <<= 1) {
        const uint64_t a = __shfl_xor(mn, dd, kWave), b = __shfl_xor(mx, dd, kWave);
        mn = a < mn ? a : mn;
        mx = b > mx ? b : mx;
    }
    if (lane == 0) {
        s_red[0][wid] = mn;
        s_red[1][wid] = mx;
    }
    __syncthreads();
    mn = ~0ull;
    mx = 0;
#pragma unroll
    for (int q = 0; q < NW; ++q) {
        mn = s_red[0][q] < mn ? s_red[0][q] : mn;
        mx = s_red[1][q] > mx ? s_red[1][q] : mx;
    }
    const int nbits = bit_length(mx - mn);
    const bool packed = nbits <= KB;
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int l = wid * kWave * IT + i * kWave + lane;
        const uint64_t rel = (l < n) ? k[i] - mn : ~0ull;
        k[i] = ((packed ? rel : (rel & ((1ull << KB) - 1ull))) << IB) | (uint64_t)l;
        s_key[l] = k[i];
    }
    __syncthreads();
    block_radix_sort<NT, IT>(k, IB, packed ? nbits : KB, s_key, s_hist);
    if (!packed) {
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
            const uint64_t hi = (l < n) ? ((s_code[l] - mn) >> KB) : ~0ull;
            k[i] = (hi << IB) | (uint64_t)p;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < IT; ++i)
            s_key[wid * kWave * IT + i * kWave + lane] = k[i];
        __syncthreads();
        block_radix_sort<NT, IT>(k, IB, nbits - KB, s_key, s_hist);
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int pos = tid + i * NT;
        if (pos >= n)
            break;
        const uint64_t w = s_key[pos];
        const int l = packed ? (int)(w & IMASK) : (int)s_val[w & IMASK];
        L.code[1][j0 + pos] = s_code[l];
        L.item[1][j0 + pos] = L.item[0][j0 + l];
    }
}

// merge sorted runs of length Lr of (code, item), stable; src -> dst
__global__ __launch_bounds__(256) void k_lbvh_merge(Dev d, Lbvh L, int Lr, int src)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int U = *L.n_cand;
    if (i >= U)
        return; // Lr >= U still copies (pr * Lr >= U): buffers stay in step
    const uint64_t *sk = L.code[src];
    const int r = i / Lr, pr = r ^ 1;
    const uint64_t key = sk[i];
    int out;
    if (pr * Lr >= U) {
        out = i;
    } else {
        const int p0 = pr * Lr, p1 = min(pr * Lr + Lr, U);
        const int rank = (r & 1) ? upper_bound_u64(sk, p0, p1, key) - p0
                                 : lower_bound_u64(sk, p0, p1, key) - p0;
        out = min(r, pr) * Lr + (i - r * Lr) + rank;
    }
    L.code[src ^ 1][out] = key;
    L.item[src ^ 1][out] = L.item[src][i];
}

// copy the sorted codes back to buffer 1 when the merges ended in buffer 0
__global__ __launch_bounds__(256) void k_lbvh_settle(Dev d, Lbvh L, int src)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int U = *L.n_cand;
    if (src == 1 || i >= U)
        return;
    L.code[1][i] = L.code[0][i];
    L.item[1][i] = L.item[0][i];
}

// common prefix length of the augmented keys (code, index); -1 out of range
__device__ __forceinline__ int lbvh_delta(const uint64_t *code, int U, int i, int j)
{
    if (j < 0 || j >= U)
        return -1;
    const uint64_t a = code[i], b = code[j];
    if (a != b)
        return __builtin_clzll(a ^ b);
    return 64 + __builtin_clz((uint32_t)(i ^ j));
}

// Karras (2012) hierarchy: internal node i covers a key range [min(i,j),
// max(i,j)] found by exponential + binary search; its split is the last
// position sharing more than the range's common prefix.
__global__ __launch_bounds__(256) void k_lbvh_build(Dev d, Lbvh L)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int U = *L.n_cand;
    if (i == 0 && U > 0)
        L.parent[0] = -1;
    if (i >= U - 1)
        return;
    const uint64_t *c = L.code[1];
    const int dl = lbvh_delta(c, U, i, i + 1) - lbvh_delta(c, U, i, i - 1);
    const int dir = dl >= 0 ? 1 : -1;
    const int dmin = lbvh_delta(c, U, i, i - dir);
    int lmax = 2;
    while (lbvh_delta(c, U, i, i + lmax * dir) > dmin)
        lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (lbvh_delta(c, U, i, i + (l + t) * dir) > dmin)
            l += t;
    const int j = i + l * dir;
    const int dnode = lbvh_delta(c, U, i, j);
    int s = 0;
    for (int div = 2;; div *= 2) {
        const int t = (l + div - 1) / div;
        if (lbvh_delta(c, U, i, i + (s + t) * dir) > dnode)
            s += t;
        if (t == 1)
            break;
    }
    const int gamma = i + s * dir + min(dir, 0);
    const int lo = min(i, j), hi = max(i, j);
    const int left = (lo == gamma) ? ~gamma : gamma;
    const int right = (hi == gamma + 1) ? ~(gamma + 1) : gamma + 1;
    L.left[i] = left;
    L.right[i] = right;
    L.parent[left >= 0 ? left : (U - 1 + ~left)] = i;
    L.parent[right >= 0 ? right : (U - 1 + ~right)] = i;
    L.flag[i] = 0;
}

// Leaf boxes [x - R', x + R'] rounded outward, then unions up the tree: the
// second child to reach a node computes its box and continues.  Boxes are
// written with agent-scope atomic stores and read with agent-scope atomic
// loads, which the XCDs' L2 caches do not hold back; a child counts its
// arrival at the parent only once its box stores are acknowledged, so the
// second arrival reads both children's boxes.  No device-scope fence per
// level (each one writes back the L2: microseconds, times the tree depth).
template <int DIM>
__global__ __launch_bounds__(256) void k_lbvh_boxes(Dev d, Step st, Lbvh L)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int U = *L.n_cand;
    if (i >= U)
        return;
    auto put = [](float *b, const float (&lo)[3], const float (&hi)[3]) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            __hip_atomic_store(&b[a], lo[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&b[3 + a], hi[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    };
    auto get = [](const float *b, int a) {
        return __hip_atomic_load(&b[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    // the leaf's point: a candidate in the sorted window, or (whole-window
    // tree) the point in ring slot item
    const float *Nx = L.whole ? d.ring : d.axyz[st.par ^ 1];
    const int p = L.whole ? L.item[1][i] : d.cand[L.item[1][i]];
    const int leaf = U - 1 + i;
    float lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = hi[a] = 0.f;
        if (a < DIM) {
            const double x = (double)Nx[(size_t)a * d.W + p];
            lo[a] = __double2float_rd(x - d.Rp);
            hi[a] = __double2float_ru(x + d.Rp);
        }
    }
    L.count[i] = 0;
    put(L.box + (size_t)leaf * 6, lo, hi);
    int node = L.parent[leaf];
    while (node >= 0) {
        if (atomicAdd(&L.flag[node], 1) == 0)
            return; // the sibling subtree finishes this node
        const int a0 = L.left[node], b0 = L.right[node];
        const float *ba = L.box + (size_t)(a0 >= 0 ? a0 : U - 1 + ~a0) * 6;
        const float *bb = L.box + (size_t)(b0 >= 0 ? b0 : U - 1 + ~b0) * 6;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(get(ba, a), get(bb, a));
            hi[a] = fmaxf(get(ba, 3 + a), get(bb, 3 + a));
        }
        put(L.box + (size_t)node * 6, lo, hi);
        node = L.parent[node];
    }
}

// Node ids of the walk: internal nodes 0..U-2, leaf j at U-1+j (the layout
// of L.parent and L.box); a child entry c of L.left/L.right is c >= 0 for an
// internal node, ~j for leaf j.
__device__ __forceinline__ int lbvh_node(int c, int U) { return c >= 0 ? c : U - 1 + ~c; }

// ---------------------------------------------------------------------------
// Group walk: one wavefront walks the tree for a group of up to 64 query
// points (consecutive in the sorted window, so the group is compact) against
// the group's bounding box, 64 nodes at a time: each lane tests one node's
// box for overlap, and the overlapping internal nodes push both children on
// the wave's node stack in LDS, the overlapping leaves go to a leaf list;
// every 64 leaves are handed to batch(n) (leaf indices in w.leaf[0..n)),
// which tests them against every query of the group -- 64 lanes wide -- and
// returns false to end the walk.  One thread per query walking alone gives a
// few hundred wavefronts at W = 10k, each a chain of dependent loads; this
// gives every level of the walk one round trip for 64 nodes.
// Stack bound: a step pops the top min(64, top) entries and pushes at most
// 128, so a push leaves at most 64 entries behind once the next step pops,
// and the pushes still present have strictly increasing least depth; with
// D levels the stack never holds more than 64 D + 128 entries.  A Karras
// tree over 63-bit codes with index tie-breaks has at most 98 levels.
// ---------------------------------------------------------------------------
constexpr int kLbvhNT = 128;                   // two wavefronts per workgroup
constexpr int kLbvhStack = 64 * 98 + 128;      // nodes per wavefront
struct LbvhWaveLds {
    int32_t stack[kLbvhStack];
    int32_t leaf[2 * kWave];
};

template <int DIM, typename F>
__device__ __forceinline__ void lbvh_group_walk(const Lbvh &L, int U, const float (&glo)[3],
                                                const float (&ghi)[3], LbvhWaveLds &w,
                                                int32_t *err, F batch)
{
    const int lane = lane_id();
    const uint64_t lt = (1ull << lane) - 1ull;
    int top = 1, nleaf = 0; // wave-uniform
    if (lane == 0)
        w.stack[0] = 0; // the root (leaf 0 itself when U == 1)
    wave_sync_lds();
    while (top > 0) {
        const int n = min(top, kWave);
        top -= n;
        int node = 0, l = 0, r = 0;
        bool hit = false;
        if (lane < n) {
            node = w.stack[top + lane];
            const float *b = L.box + (size_t)node * 6;
            float bl[3], bh[3];
#pragma unroll
            for (int a = 0; a < DIM; ++a) {
                bl[a] = b[a];
                bh[a] = b[3 + a];
            }
            if (node < U - 1) { // requested with the box: one round trip
                l = L.left[node];
                r = L.right[node];
            }
            hit = true;
#pragma unroll
            for (int a = 0; a < DIM; ++a)
                hit = hit && bl[a] <= ghi[a] && bh[a] >= glo[a];
        }
        const bool leaf = node >= U - 1;
        const uint64_t im = __ballot(hit && !leaf), lm = __ballot(hit && leaf);
        const int ni = __popcll(im), nl = __popcll(lm);
        if (top + 2 * ni > kLbvhStack) { // cannot happen for a Karras tree
            if (lane == 0)
                atomicOr(err, kErrLbvhWalk);
            return;
        }
        wave_sync_lds(); // the pops are read before the pushes overwrite them
        if (hit && !leaf) {
            const int k = top + 2 * __popcll(im & lt);
            w.stack[k] = lbvh_node(l, U);
            w.stack[k + 1] = lbvh_node(r, U);
        }
        if (hit && leaf)
            w.leaf[nleaf + __popcll(lm & lt)] = node - (U - 1);
        top += 2 * ni;
        nleaf += nl;
        wave_sync_lds();
        if (nleaf >= kWave) {
            if (!batch(kWave))
                return;
            const int rest = nleaf - kWave;
            const int v = lane < rest ? w.leaf[kWave + lane] : 0;
            wave_sync_lds();
            if (lane < rest)
                w.leaf[lane] = v;
            nleaf = rest;
            wave_sync_lds();
        }
    }
    if (nleaf > 0)
        (void)batch(nleaf);
}

// the group's query points: lane q holds query q's coordinates; the box
// over the active lanes (inactive: +inf / -inf, so an all-inactive group
// overlaps nothing)
template <int DIM>
__device__ __forceinline__ void group_box(const float (&x)[3], bool active, float (&glo)[3],
                                          float (&ghi)[3])
{
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float lo = (a < DIM && active) ? x[a] : __builtin_inff();
        float hi = (a < DIM && active) ? x[a] : -__builtin_inff();
        for (int m = 1; m < kWave; m <<= 1) {
            lo = fminf(lo, __shfl_xor(lo, m, kWave));
            hi = fmaxf(hi, __shfl_xor(hi, m, kWave));
        }
        glo[a] = a < DIM ? lo : 0.f;
        ghi[a] = a < DIM ? hi : 0.f;
    }
}

// Default index (boxes of the slide's candidates): every window point is a
// query; a group is G <= 64 consecutive sorted-window points (one wavefront
// each: smaller groups have smaller boxes, so fewer leaves each, and give
// more wavefronts to a small window).  Per batch of
// leaves (one per lane): the candidate's count is read (racy; counts only
// grow, so one already past K stays past it and is skipped), each of the
// group's queries is broadcast and tested against every lane's candidate
// (exact FP32 predicate; it implies the query lies in the candidate's box),
// and each lane adds its candidate's hits with one atomic -- counts stay
// exact up to K+1.
template <int DIM>
__global__ __launch_bounds__(kLbvhNT) void k_lbvh_query(Dev d, Step st, Lbvh L, int G)
{
    __shared__ LbvhWaveLds lds[kLbvhNT / kWave];
    stamp_start(d, kTsDetect);
    const int lane = lane_id(), wid = threadIdx.x / kWave;
    const int q0 = (blockIdx.x * (kLbvhNT / kWave) + wid) * G;
    const int U = *L.n_cand;
    if (U == 0 || q0 >= d.W)
        return;
    const float *Nx = d.axyz[st.par ^ 1];
    const int nq = min(G, d.W - q0);
    const bool active = lane < nq;
    float x[3] = {0.f, 0.f, 0.f};
    if (active)
#pragma unroll
        for (int a = 0; a < DIM; ++a)
            x[a] = Nx[(size_t)a * d.W + q0 + lane];
    float glo[3], ghi[3];
    group_box<DIM>(x, active, glo, ghi);
    const int K = d.K;
    long long tests = 0;
    LbvhWaveLds &w = lds[wid];
    lbvh_group_walk<DIM>(L, U, glo, ghi, w, &d.ctl->err, [&](int nb) {
        const bool mine = lane < nb;
        const int j = mine ? w.leaf[lane] : 0;
        float p[3] = {0.f, 0.f, 0.f};
        bool open = false;
        if (mine) {
            open = __hip_atomic_load(&L.count[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= K;
            const int cp = d.cand[L.item[1][j]];
#pragma unroll
            for (int a = 0; a < DIM; ++a)
                p[a] = Nx[(size_t)a * d.W + cp];
        }
        int acc = 0;
        for (int qi = 0; qi < nq; ++qi) {
            const float qx = lane_f(x[0], qi);
            const float qy = DIM > 1 ? lane_f(x[1], qi) : 0.f;
            const float qz = DIM > 2 ? lane_f(x[2], qi) : 0.f;
            acc += (open && dist2<DIM>(p, qx, qy, qz) <= d.r2) ? 1 : 0;
        }
        tests += (long long)__popcll(__ballot(open)) * nq;
        if (acc)
            atomicAdd(&L.count[j], acc);
        return true;
    });
    if (lane == 0 && tests)
        atomicAdd(&cnt_of(d, st.t)->pair_tests, (unsigned long long)tests);
}

// Whole-window tree (the reference's OPTIMIZATION 0/1: a GAS over every
// window point, rays counting hits and terminating once past K,
// optixScan.cu:80-156): each query counts the leaves whose point passes the
// FP32 test (its own leaf included); the group's walk ends once every query
// of the group has passed K, and a count <= K sets the query point's outlier
// bit.  The queries, grouped by G <= 64 consecutive sorted-window positions
// (G divides 64) so that a group is compact: cands != 0, the slide's candidates (their bits in
// d.cmask, OPTIMIZATION 1); else the points whose ring slot is below nq
// (OPTIMIZATION 0, nq = min(W, launch_ray_num or W): rays past W repeat
// slots and are expanded on the host).
template <int DIM>
__global__ __launch_bounds__(kLbvhNT) void k_lbvh_count(Dev d, Step st, Lbvh L, int cands, int nq,
                                                        int G)
{
    __shared__ LbvhWaveLds lds[kLbvhNT / kWave];
    stamp_start(d, kTsDetect);
    const int lane = lane_id(), wid = threadIdx.x / kWave;
    const int U = *L.n_whole;
    SlideCnt *cn = cnt_of(d, st.t);
    if (!cands && blockIdx.x == 0 && threadIdx.x == 0)
        cn->n_cand = nq; // every query is a candidate
    const int p0 = (blockIdx.x * (kLbvhNT / kWave) + wid) * G;
    if (p0 >= d.W)
        return;
    const float *Nx = d.axyz[st.par ^ 1];
    const int32_t *Ns = d.aslot[st.par ^ 1];
    const int qp = p0 + lane;
    // G divides 64: the group's positions lie in one mask word
    const uint64_t cm = cands ? d.cmask[p0 >> 6] >> (p0 & 63) : ~0ull;
    int slot = 0;
    bool active = false;
    if (lane < G && qp < d.W && ((cm >> lane) & 1ull)) {
        slot = Ns[qp];
        active = cands || slot < nq;
    }
    float x[3] = {0.f, 0.f, 0.f};
    if (active)
#pragma unroll
        for (int a = 0; a < DIM; ++a)
            x[a] = Nx[(size_t)a * d.W + qp];
    float glo[3], ghi[3];
    group_box<DIM>(x, active, glo, ghi);
    const uint64_t amask = __ballot(active);
    const int K = d.K;
    int cnt = 0;
    long long tests = 0;
    LbvhWaveLds &w = lds[wid];
    if (amask)
        lbvh_group_walk<DIM>(L, U, glo, ghi, w, &d.ctl->err, [&](int nb) {
            const bool mine = lane < nb;
            float p[3] = {0.f, 0.f, 0.f};
            if (mine) {
                const int r = L.item[1][w.leaf[lane]];
#pragma unroll
                for (int a = 0; a < DIM; ++a)
                    p[a] = d.ring[(size_t)a * d.W + r];
            }
            uint64_t open = amask & __ballot(cnt <= K);
            tests += (long long)__popcll(open) * nb;
            while (open) {
                const int qi = __builtin_ctzll(open);
                open &= open - 1ull;
                const float qx = lane_f(x[0], qi);
                const float qy = DIM > 1 ? lane_f(x[1], qi) : 0.f;
                const float qz = DIM > 2 ? lane_f(x[2], qi) : 0.f;
                const int c = __popcll(__ballot(mine && dist2<DIM>(p, qx, qy, qz) <= d.r2));
                cnt += lane == qi ? c : 0;
            }
            return (amask & __ballot(cnt <= K)) != 0;
        });
    int hits = active ? min(cnt, K + 1) : 0;
    if (active) {
        if (cnt <= K)
            atomicOr(&bitmap_of(d, st.t)[slot >> 5], 1u << (slot & 31));
    }
    for (int dd = 1; dd < kWave; dd <<= 1)
        hits += __shfl_xor(hits, dd, kWave);
    if (lane == 0 && tests)
        atomicAdd(&cn->pair_tests, (unsigned long long)tests);
    if (lane == 0 && hits)
        atomicAdd(&cn->hits, (unsigned long long)hits);
}

// refit of a whole-window tree: internal nodes' arrival counters cleared
// before k_lbvh_boxes recomputes every box over the kept topology
__global__ __launch_bounds__(256) void k_lbvh_unflag(Dev d, Lbvh L)
{
    stamp_start(d, kTsIndex);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < *L.n_whole - 1)
        L.flag[i] = 0;
}

// outlier bits from the leaf counts (the slide epilogue, k_finish, follows)
__global__ __launch_bounds__(kDetectNT) void k_lbvh_select(Dev d, Step st, Lbvh L)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int U = *L.n_cand;
    int h = 0;
    if (i < U) {
        const int c = L.count[i];
        h = min(c, d.K + 1);
        d.cand_cnt[L.item[1][i]] = min(c, d.K + 1);
        if (c <= d.K) {
            const int slot = d.aslot[st.par ^ 1][d.cand[L.item[1][i]]];
            atomicOr(&bitmap_of(d, st.t)[slot >> 5], 1u << (slot & 31));
        }
    }
    for (int dd = 1; dd < kWave; dd <<= 1)
        h += __shfl_xor(h, dd, kWave);
    if (lane_id() == 0 && h)
        atomicAdd(&cnt_of(d, st.t)->hits, (unsigned long long)h);
}

} // namespace rtod
