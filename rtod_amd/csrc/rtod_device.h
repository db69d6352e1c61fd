// rtod_device.h -- device-side building blocks shared by every RTOD kernel:
// the FP32 neighbour predicate, cell keys, wave/block scans and searches.
//
// Compiled with -ffp-contract=off: the only fused operations are the explicit
// __builtin_fmaf of the predicate, so the device and the CPU oracle round the
// same way bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <climits>
#include <stdint.h>

namespace rtod {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Predicate.  The reference tests FP64 `|dx| < R` (1-D) and `sqdist < R2`
// (3-D) in __intersection__cube (optixScan.cu:170-189); this build's contract
// is FP32 d2 <= r2 with r2 = R*R rounded once, identical on host and device.
// ---------------------------------------------------------------------------
template <int DIM>
__device__ __forceinline__ float dist2(const float (&p)[3], float qx, float qy, float qz)
{
    float dx = p[0] - qx;
    float d2 = dx * dx;
    if constexpr (DIM > 1) {
        float dy = p[1] - qy;
        d2 = __builtin_fmaf(dy, dy, d2);
    }
    if constexpr (DIM > 2) {
        float dz = p[2] - qz;
        d2 = __builtin_fmaf(dz, dz, d2);
    }
    return d2;
}

// ---------------------------------------------------------------------------
// Cells.  Side c = R (1 - 1e-5) / sqrt(d); per-axis coordinate
// floor(x * (1/c)) computed in double and clamped to [-2^B, 2^B - 1]
// (B = 30 for d <= 2, 20 for d = 3).  The reference uses side R / sqrt(d)
// with global min/max offsets (optixScan.cpp:947-965); the shrink makes any
// two points of one unclamped cell provably closer than R under the FP32
// predicate, so a cell with >= K+1 points is dense without a bbox pass.
// Keys: d=1 k0+2^30; d=2 (k0+2^30)<<31 | (k1+2^30); d=3 three 21-bit fields,
// k2 least significant, so each (k0,k1) row of cells is a contiguous key run.
// ---------------------------------------------------------------------------
template <int DIM> struct CellBits {
    static constexpr int B = (DIM == 3) ? 20 : 30;
    static constexpr int FIELD = B + 1;
    static constexpr uint64_t FMASK = (1ull << FIELD) - 1;
};

// The clamped value fits int32 (|k| <= 2^30), so the conversion is the
// single native f64 -> i32 instruction.
template <int DIM>
__device__ __forceinline__ int cell_coord(double v, double inv_c)
{
    constexpr int B = CellBits<DIM>::B;
    double f = floor(v * inv_c);
    const double lo = -(double)(1ll << B);
    const double hi = (double)((1ll << B) - 1);
    if (!(f > lo))
        f = lo;
    else if (!(f < hi))
        f = hi;
    return (int)f;
}

template <int DIM>
__device__ __forceinline__ uint64_t pack_key(int64_t k0, int64_t k1, int64_t k2)
{
    constexpr int B = CellBits<DIM>::B;
    constexpr int64_t bias = 1ll << B;
    if constexpr (DIM == 1) {
        return (uint64_t)(k0 + bias);
    } else if constexpr (DIM == 2) {
        return ((uint64_t)(k0 + bias) << 31) | (uint64_t)(k1 + bias);
    } else {
        return ((uint64_t)(k0 + bias) << 42) | ((uint64_t)(k1 + bias) << 21) |
               (uint64_t)(k2 + bias);
    }
}

template <int DIM>
__device__ __forceinline__ uint64_t cell_key(const float (&p)[3], double inv_c)
{
    const int64_t k0 = cell_coord<DIM>((double)p[0], inv_c);
    const int64_t k1 = DIM > 1 ? cell_coord<DIM>((double)p[1], inv_c) : 0;
    const int64_t k2 = DIM > 2 ? cell_coord<DIM>((double)p[2], inv_c) : 0;
    return pack_key<DIM>(k0, k1, k2);
}

// A key whose every field lies strictly inside the clamp range: the cell's
// diameter bound holds, so its count alone decides "dense".
template <int DIM>
__device__ __forceinline__ bool key_guaranteed(uint64_t key)
{
    constexpr int F = CellBits<DIM>::FIELD;
    constexpr uint64_t M = CellBits<DIM>::FMASK;
    bool ok = true;
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
        uint64_t f = (key >> (F * (DIM - 1 - a))) & M;
        ok = ok && (f > 0) && (f < M);
    }
    return ok;
}

// ---------------------------------------------------------------------------
// Scans and searches
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

// Wave-wide scans and reductions on DPP moves (register to register between
// lanes, a few cycles each) instead of ds_bpermute shuffles (an LDS round
// trip each): row_shr:1, 2, 4, 8 within each 16-lane row, then row_bcast:15
// (lane 15 of rows 0 and 2 into rows 1 and 3) and row_bcast:31 (lane 31 into
// rows 2 and 3).  A lane whose DPP source does not exist receives `id`, the
// operation's identity, so after the six steps lane i holds the inclusive
// prefix over lanes 0..i and lane 63 the whole wave's result.
enum : int {
    kDppRowShr1 = 0x111,
    kDppRowShr2 = 0x112,
    kDppRowShr4 = 0x114,
    kDppRowShr8 = 0x118,
    kDppRowBcast15 = 0x142,
    kDppRowBcast31 = 0x143
};

template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_move(int id, int v)
{
    return __builtin_amdgcn_update_dpp(id, v, CTRL, ROWS, 0xf, false);
}

template <typename OP>
__device__ __forceinline__ int wave_prefix(int x, int id, OP op)
{
    x = op(x, dpp_move<kDppRowShr1, 0xf>(id, x));
    x = op(x, dpp_move<kDppRowShr2, 0xf>(id, x));
    x = op(x, dpp_move<kDppRowShr4, 0xf>(id, x));
    x = op(x, dpp_move<kDppRowShr8, 0xf>(id, x));
    x = op(x, dpp_move<kDppRowBcast15, 0xa>(id, x));
    x = op(x, dpp_move<kDppRowBcast31, 0xc>(id, x));
    return x;
}

__device__ __forceinline__ int wave_incl_scan(int x)
{
    return wave_prefix(x, 0, [](int a, int b) { return a + b; });
}

// whole-wave reductions (every lane gets the result)
__device__ __forceinline__ int wave_min_i32(int x)
{
    return __builtin_amdgcn_readlane(
        wave_prefix(x, INT_MAX, [](int a, int b) { return a < b ? a : b; }), kWave - 1);
}
__device__ __forceinline__ int wave_max_i32(int x)
{
    return __builtin_amdgcn_readlane(
        wave_prefix(x, INT_MIN, [](int a, int b) { return a > b ? a : b; }), kWave - 1);
}
__device__ __forceinline__ int wave_or_i32(int x)
{
    return __builtin_amdgcn_readlane(wave_prefix(x, 0, [](int a, int b) { return a | b; }),
                                     kWave - 1);
}
// float min/max over the non-NaN values (fminf/fmaxf drop NaN operands)
__device__ __forceinline__ float wave_min_f32(float x)
{
    const int r = wave_prefix(__float_as_int(x), __float_as_int(__builtin_inff()), [](int a, int b) {
        return __float_as_int(fminf(__int_as_float(a), __int_as_float(b)));
    });
    return __int_as_float(__builtin_amdgcn_readlane(r, kWave - 1));
}
__device__ __forceinline__ float wave_max_f32(float x)
{
    const int r = wave_prefix(__float_as_int(x), __float_as_int(-__builtin_inff()),
                              [](int a, int b) {
                                  return __float_as_int(fmaxf(__int_as_float(a), __int_as_float(b)));
                              });
    return __int_as_float(__builtin_amdgcn_readlane(r, kWave - 1));
}

// whole-wave sum of a 64-bit counter on 32-bit DPP moves: the high words,
// bits 8..31 and bits 0..7 summed apart (64 values of 24 bits, or of 8,
// cannot carry out of 32 bits), exact while the total fits 64 bits
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long x)
{
    auto add = [](int a, int b) { return (int)((unsigned)a + (unsigned)b); };
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane(
        wave_prefix((int)(unsigned)(x >> 32), 0, add), kWave - 1);
    const unsigned mid = (unsigned)__builtin_amdgcn_readlane(
        wave_prefix((int)((unsigned)(x >> 8) & 0xffffffu), 0, add), kWave - 1);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane(
        wave_prefix((int)((unsigned)x & 0xffu), 0, add), kWave - 1);
    return ((unsigned long long)hi << 32) + ((unsigned long long)mid << 8) + lo;
}

// Exclusive block scan of one int per thread; s_tmp holds NT/64 ints.
template <int NT>
__device__ __forceinline__ int block_excl_scan(int x, int *s_tmp, int &total)
{
    const int lane = lane_id(), wid = threadIdx.x / kWave;
    int inc = wave_incl_scan(x);
    if (lane == kWave - 1)
        s_tmp[wid] = inc;
    __syncthreads();
    int pre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NT / kWave; ++w) {
        int v = s_tmp[w];
        pre += (w < wid) ? v : 0;
        tot += v;
    }
    __syncthreads();
    total = tot;
    return pre + inc - x;
}

__device__ __forceinline__ uint64_t wave_or(uint64_t x)
{
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1)
        x |= __shfl_xor(x, d, kWave);
    return x;
}

__device__ __forceinline__ uint64_t wave_and(uint64_t x)
{
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1)
        x &= __shfl_xor(x, d, kWave);
    return x;
}

// Block-wide copy of n elements global -> LDS with U loads in flight per
// thread: all U loads of a round are issued before any LDS store, so the
// copy costs ceil(n / (U * blockDim)) memory round trips instead of one per
// element per thread.
template <int U, typename T>
__device__ __forceinline__ void copy_to_lds(T *__restrict__ s, const T *__restrict__ g, int n)
{
    const int nt = blockDim.x;
    for (int base = threadIdx.x; base < n; base += U * nt) {
        T v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = base + u * nt;
            if (i < n)
                v[u] = g[i];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = base + u * nt;
            if (i < n)
                s[i] = v[u];
        }
    }
}

// first index in [lo,hi) with a[i] >= k (hi if none)
__device__ __forceinline__ int lower_bound_u64(const uint64_t *a, int lo, int hi, uint64_t k)
{
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] < k)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// first index in [lo,hi) with a[i] > k (hi if none)
__device__ __forceinline__ int upper_bound_u64(const uint64_t *a, int lo, int hi, uint64_t k)
{
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] <= k)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// first index in [lo,hi) with (ak[i], as[i]) >= (k, s) lexicographically
__device__ __forceinline__ int lower_bound_pair(const uint64_t *ak, const int32_t *as, int lo,
                                                int hi, uint64_t k, int32_t s)
{
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        uint64_t km = ak[mid];
        bool less = (km < k) || (km == k && as[mid] < s);
        if (less)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Two-level search over the cell-sorted key array a[0,n): s_samp[s] = a[s*stride]
// (ns entries, held in LDS) narrows the range to one stride, then a short
// global search finishes.  lower (STRICT=false) or upper (STRICT=true) bound.
template <bool UPPER>
__device__ __forceinline__ int bound2(const uint64_t *a, int n, const uint64_t *s_samp, int ns,
                                      int stride_log2, uint64_t k)
{
    int lo = 0, hi = ns;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        bool go_right = UPPER ? (s_samp[mid] <= k) : (s_samp[mid] < k);
        if (go_right)
            lo = mid + 1;
        else
            hi = mid;
    }
    if (lo == 0)
        return 0;
    int L = ((lo - 1) << stride_log2) + 1;
    int H = lo << stride_log2;
    if (H > n)
        H = n;
    return UPPER ? upper_bound_u64(a, L, H, k) : lower_bound_u64(a, L, H, k);
}

} // namespace rtod
