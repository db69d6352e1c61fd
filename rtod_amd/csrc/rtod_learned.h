// rtod_learned.h -- the learned work scheduler's predictor.
//
// A 6 -> 32 -> 32 -> 1 MLP (ReLU, ReLU, sigmoid) maps cheap per-candidate
// features to a predicted neighbour count; the detect kernel uses it only to
// choose the order and the path of each exact range count (probable inliers
// try their nearest cells first and usually stop early), never to decide an
// outcome.  Both hidden layers run on the matrix cores as
// v_mfma_f32_16x16x32_bf16 (BF16 operands, FP32 accumulation); the output
// layer, biases and activations are FP32.  Training is online: a single
// workgroup, launched on a stream of its own after each slide, takes Adam
// steps on minibatches drawn from the exact counts of the previous slides
// (20 kept) while the next slide's latency-bound ingest and merge leave CUs
// idle; new weights are published through a double buffer.
//
// MFMA 16x16x32 lane layout (lane l of 64): A operand holds row l%16,
// k = (l/16)*8 + i (i < 8); B operand holds column l%16, the same k; the
// 4-float result holds column l%16, rows (l/16)*4 + v.
//
// Features (all in [-0.5, 1]): the candidate's own-cell count, the counts of
// the cells just below and above it along the last axis (each capped at K+1
// and divided by K+1), and its offset from the cell centre along each axis
// (frac(x / c) - 0.5; zero for missing axes).
// Target: capped count y = 1 for an inlier (count > K), 0.5 * count / (K+1)
// for an outlier; predicted inlier iff the output >= 0.75.
// Parameters (flat FP32): W1[6][32] b1[32] W2[32][32] b2[32] W3[32] b3.
#pragma once

#include "rtod_device.h"
#include "rtod_state.h"

namespace rtod {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMlpIn = 6, kMlpH = 32;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpIn * kMlpH;
constexpr int kOffW2 = kOffB1 + kMlpH;
constexpr int kOffB2 = kOffW2 + kMlpH * kMlpH;
constexpr int kOffW3 = kOffB2 + kMlpH;
constexpr int kOffB3 = kOffW3 + kMlpH;
constexpr int kMlpNP = kOffB3 + 1;
static_assert(kMlpNP == 1313, "parameter count");

constexpr int kReplaySlides = 20;   // training data: the previous 20 slides
constexpr int kFeatStride = 8;      // floats per stored feature row
constexpr float kInlierCut = 0.75f; // predicted inlier iff output >= this
constexpr int kTrainNT = 256;       // training workgroup: 4 waves
constexpr int kTrainRows = 128;     // samples per step: 32 per wave
constexpr int kTrainSteps = 2;      // Adam steps per training launch
constexpr int kTrainStepsBatch = 8; // per launch after a batch step (a step is ms long)
constexpr float kAdamLr = 5e-3f, kAdamB1 = 0.9f, kAdamB2 = 0.999f, kAdamEps = 1e-7f;

__device__ __forceinline__ f32x4 mfma_bf16(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }

// logistic function from the hardware exp2 and reciprocal (~1 ulp each)
__device__ __forceinline__ float sigmoidf(float z) { return __frcp_rn(1.f + __expf(-z)); }

__device__ __forceinline__ float mlp_target(int cnt, int K)
{
    return cnt > K ? 1.f : 0.5f * (float)cnt / (float)(K + 1);
}

// B operands of the hidden layers, column tile t (columns 16t..16t+15).
// Layer 1: B[k = feature][n = unit] = W1[k][n], zero for k >= 6.
__device__ __forceinline__ bf16x8 b_w1(const float *P, int t)
{
    const int l = lane_id(), n = (l & 15) + 16 * t, k0 = (l >> 4) * 8;
    bf16x8 b;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int k = k0 + i;
        b[i] = to_bf16(k < kMlpIn ? P[kOffW1 + k * kMlpH + n] : 0.f);
    }
    return b;
}

// Layer 2 forward: B[k = input unit][n = output unit] = W2[k][n].
__device__ __forceinline__ bf16x8 b_w2(const float *P, int t)
{
    const int l = lane_id(), n = (l & 15) + 16 * t, k0 = (l >> 4) * 8;
    bf16x8 b;
#pragma unroll
    for (int i = 0; i < 8; ++i)
        b[i] = to_bf16(P[kOffW2 + (k0 + i) * kMlpH + n]);
    return b;
}

// Layer 2 backward (dh1 = d2 W2^T): B[k = output unit][n = input unit] = W2[n][k].
__device__ __forceinline__ bf16x8 b_w2t(const float *P, int t)
{
    const int l = lane_id(), n = (l & 15) + 16 * t, k0 = (l >> 4) * 8;
    bf16x8 b;
#pragma unroll
    for (int i = 0; i < 8; ++i)
        b[i] = to_bf16(P[kOffW2 + n * kMlpH + k0 + i]);
    return b;
}

__device__ __forceinline__ void wave_lds_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum over the 16 lanes of each lane group (l/16): every lane gets its
// group's total.
__device__ __forceinline__ float group16_sum(float x)
{
#pragma unroll
    for (int m = 1; m < 16; m <<= 1)
        x += __shfl_xor(x, m, kWave);
    return x;
}

// Sum over the four lane groups: lane l gets the total of lanes l%16 + 16g.
__device__ __forceinline__ float cross_group_sum(float x)
{
    x += __shfl_xor(x, 16, kWave);
    x += __shfl_xor(x, 32, kWave);
    return x;
}

// ---------------------------------------------------------------------------
// Inference for up to 64 rows of features held by one wavefront in LDS:
// s_x[r][0..7] (zero padded) for r < m -> s_z[r] = the pre-sigmoid output.
// Only ceil(m/16) 16-row tiles run, all tiles of a layer back to back, so a
// layer costs one LDS round trip.  s_h: 64*32 bf16, private to the wave.
// Every lane of the wave must call it (m uniform).
// ---------------------------------------------------------------------------
// A lane's slices of the weights for inference (MFMA B operands, biases,
// output weights), loaded once so the loads can be issued early.
struct MlpRegs {
    bf16x8 w1a, w1b, w2a, w2b;
    float b1a, b1b, b2a, b2b, w3a, w3b, b3;
};

__device__ __forceinline__ MlpRegs mlp_regs(const float *P)
{
    const int n0 = lane_id() & 15;
    MlpRegs r;
    r.w1a = b_w1(P, 0);
    r.w1b = b_w1(P, 1);
    r.w2a = b_w2(P, 0);
    r.w2b = b_w2(P, 1);
    r.b1a = P[kOffB1 + n0];
    r.b1b = P[kOffB1 + n0 + 16];
    r.b2a = P[kOffB2 + n0];
    r.b2b = P[kOffB2 + n0 + 16];
    r.w3a = P[kOffW3 + n0];
    r.w3b = P[kOffW3 + n0 + 16];
    r.b3 = P[kOffB3];
    return r;
}

__device__ __forceinline__ void mlp_infer_rows(const MlpRegs &Wr, const float (*s_x)[kFeatStride],
                                               int m, __bf16 *s_h, float *s_z)
{
    const int l = lane_id(), n0 = l & 15, grp = l >> 4;
    const int tiles = (m + 15) >> 4;
    const bf16x8 w1a = Wr.w1a, w1b = Wr.w1b, w2a = Wr.w2a, w2b = Wr.w2b;
    const float b1a = Wr.b1a, b1b = Wr.b1b, b2a = Wr.b2a, b2b = Wr.b2b;
    const float w3a = Wr.w3a, w3b = Wr.w3b;
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    // layer 1: A = feature rows (k < 8 in lane group 0, zero elsewhere)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        if (r < tiles) {
            bf16x8 a;
#pragma unroll
            for (int i = 0; i < 8; ++i)
                a[i] = to_bf16(grp == 0 ? s_x[r * 16 + n0][i] : 0.f);
            const f32x4 c0 = mfma_bf16(a, w1a, zero);
            const f32x4 c1 = mfma_bf16(a, w1b, zero);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = r * 16 + grp * 4 + v;
                s_h[row * kMlpH + n0] = to_bf16(fmaxf(c0[v] + b1a, 0.f));
                s_h[row * kMlpH + n0 + 16] = to_bf16(fmaxf(c1[v] + b1b, 0.f));
            }
        }
    }
    wave_lds_fence();
    // layer 2 and the output dot product
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        if (r < tiles) {
            const bf16x8 a2 =
                *reinterpret_cast<const bf16x8 *>(&s_h[(r * 16 + n0) * kMlpH + grp * 8]);
            const f32x4 d0 = mfma_bf16(a2, w2a, zero);
            const f32x4 d1 = mfma_bf16(a2, w2b, zero);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                float part = fmaxf(d0[v] + b2a, 0.f) * w3a + fmaxf(d1[v] + b2b, 0.f) * w3b;
                part = group16_sum(part);
                if (n0 == 0)
                    s_z[r * 16 + grp * 4 + v] = part;
            }
        }
    }
    wave_lds_fence();
}

// One wavefront, lane l's features f -> predicted output (all 64 rows).
// s_x: 64 x 8 floats, s_h: 64*32 bf16, s_z: 64 floats, private to the wave.
__device__ __forceinline__ float mlp_infer_wave(const float *P, const float (&f)[kMlpIn],
                                                float (*s_x)[kFeatStride], __bf16 *s_h,
                                                float *s_z)
{
    const int l = lane_id();
#pragma unroll
    for (int j = 0; j < kFeatStride; ++j)
        s_x[l][j] = j < kMlpIn ? f[j < kMlpIn ? j : 0] : 0.f;
    wave_lds_fence();
    const MlpRegs Wr = mlp_regs(P);
    mlp_infer_rows(Wr, s_x, kWave, s_h, s_z);
    const float z = s_z[l] + Wr.b3;
    wave_lds_fence();
    return sigmoidf(z);
}

// ---------------------------------------------------------------------------
// Training: per-wave LDS of one 32-sample slice, block LDS of the weights.
// ---------------------------------------------------------------------------
struct alignas(16) TrainWaveLds {
    float xs[32][kFeatStride];
    float ys[32];
    __bf16 h1[32][kMlpH];
    __bf16 d2[32][kMlpH];
    __bf16 d1[32][kMlpH];
};

struct TrainLds {
    float p[kMlpNP + 3]; // weights
    float g[kMlpNP + 3]; // gradient of the current step
    float loss;
    int pref[kReplaySlides + 1];
};
constexpr int kTrainMaxStreams = 1024; // a batch's replays sampled by one trainer

// Gradient of the mean loss over n_total samples, for the 32 samples of this
// wave (rows row0 .. row0+31; rows >= n_total contribute nothing).  w.xs and
// w.ys hold the samples (xs zero-padded).  Adds into G (LDS, atomics);
// returns the sum of squared errors of the wave's valid rows.
__device__ __forceinline__ float mlp_wave_grads(const float *P, float *G, TrainWaveLds &w,
                                                int row0, int n_total)
{
    const int l = lane_id(), n0 = l & 15, grp = l >> 4;
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    const bf16x8 w1[2] = {b_w1(P, 0), b_w1(P, 1)};
    const bf16x8 w2[2] = {b_w2(P, 0), b_w2(P, 1)};
    const bf16x8 w2t[2] = {b_w2t(P, 0), b_w2t(P, 1)};
    float b1[2], b2[2], w3[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        b1[t] = P[kOffB1 + n0 + 16 * t];
        b2[t] = P[kOffB2 + n0 + 16 * t];
        w3[t] = P[kOffW3 + n0 + 16 * t];
    }
    const float b3 = P[kOffB3];
    const float inv_n = 1.f / (float)n_total;

    // layer 1
    f32x4 a1[2][2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        bf16x8 a;
#pragma unroll
        for (int i = 0; i < 8; ++i)
            a[i] = to_bf16(grp == 0 ? w.xs[r * 16 + n0][i] : 0.f);
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            a1[r][t] = mfma_bf16(a, w1[t], zero);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                a1[r][t][v] += b1[t];
                w.h1[r * 16 + grp * 4 + v][n0 + 16 * t] = to_bf16(fmaxf(a1[r][t][v], 0.f));
            }
        }
    }
    wave_lds_fence();
    // layer 2 and output
    f32x4 a2[2][2];
    float dz[2][4];
    float sse = 0.f;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&w.h1[r * 16 + n0][grp * 8]);
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            a2[r][t] = mfma_bf16(a, w2[t], zero);
#pragma unroll
            for (int v = 0; v < 4; ++v)
                a2[r][t][v] += b2[t];
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            float part = fmaxf(a2[r][0][v], 0.f) * w3[0] + fmaxf(a2[r][1][v], 0.f) * w3[1];
            const float z = group16_sum(part) + b3;
            const int row = r * 16 + grp * 4 + v;
            const bool valid = row0 + row < n_total;
            const float yh = sigmoidf(z);
            const float e = yh - w.ys[row];
            dz[r][v] = valid ? 2.f * e * yh * (1.f - yh) * inv_n : 0.f;
            if (valid && n0 == 0)
                sse += e * e;
        }
    }
    // output layer gradient, d2 = dz * W3 * relu'(a2)
    float gw3[2] = {0.f, 0.f}, gb2[2] = {0.f, 0.f}, gb3 = 0.f;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            if (n0 == 0)
                gb3 += dz[r][v];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const float a = a2[r][t][v];
                gw3[t] += fmaxf(a, 0.f) * dz[r][v];
                const float d2v = a > 0.f ? dz[r][v] * w3[t] : 0.f;
                gb2[t] += d2v;
                w.d2[r * 16 + grp * 4 + v][n0 + 16 * t] = to_bf16(d2v);
            }
        }
    }
    gb3 = cross_group_sum(gb3);
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        gw3[t] = cross_group_sum(gw3[t]);
        gb2[t] = cross_group_sum(gb2[t]);
    }
    if (l < 16) {
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            atomicAdd(&G[kOffW3 + n0 + 16 * t], gw3[t]);
            atomicAdd(&G[kOffB2 + n0 + 16 * t], gb2[t]);
        }
    }
    if (l == 0)
        atomicAdd(&G[kOffB3], gb3);
    wave_lds_fence();
    // dh1 = d2 W2^T on the matrix cores; d1 = dh1 * relu'(a1)
    float gb1[2] = {0.f, 0.f};
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&w.d2[r * 16 + n0][grp * 8]);
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const f32x4 c = mfma_bf16(a, w2t[t], zero);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float d1v = a1[r][t][v] > 0.f ? c[v] : 0.f;
                gb1[t] += d1v;
                w.d1[r * 16 + grp * 4 + v][n0 + 16 * t] = to_bf16(d1v);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        gb1[t] = cross_group_sum(gb1[t]);
        if (l < 16)
            atomicAdd(&G[kOffB1 + n0 + 16 * t], gb1[t]);
    }
    wave_lds_fence();
    // dW2[i][o] = sum_s h1[s][i] d2[s][o]: the 32 samples are the MFMA's k
#pragma unroll
    for (int ti = 0; ti < 2; ++ti) {
        bf16x8 a;
#pragma unroll
        for (int j = 0; j < 8; ++j)
            a[j] = w.h1[grp * 8 + j][n0 + 16 * ti];
#pragma unroll
        for (int to = 0; to < 2; ++to) {
            bf16x8 b;
#pragma unroll
            for (int j = 0; j < 8; ++j)
                b[j] = w.d2[grp * 8 + j][n0 + 16 * to];
            const f32x4 c = mfma_bf16(a, b, zero);
#pragma unroll
            for (int v = 0; v < 4; ++v)
                atomicAdd(&G[kOffW2 + (16 * ti + grp * 4 + v) * kMlpH + n0 + 16 * to], c[v]);
        }
    }
    // dW1[f][i] = sum_s x[s][f] d1[s][i]
    {
        bf16x8 a;
#pragma unroll
        for (int j = 0; j < 8; ++j)
            a[j] = to_bf16(n0 < kMlpIn ? w.xs[grp * 8 + j][n0 < kMlpIn ? n0 : 0] : 0.f);
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            bf16x8 b;
#pragma unroll
            for (int j = 0; j < 8; ++j)
                b[j] = w.d1[grp * 8 + j][n0 + 16 * t];
            const f32x4 c = mfma_bf16(a, b, zero);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int fi = grp * 4 + v;
                if (fi < kMlpIn)
                    atomicAdd(&G[kOffW1 + fi * kMlpH + n0 + 16 * t], c[v]);
            }
        }
    }
    return sse;
}

__device__ __forceinline__ uint32_t mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// ---------------------------------------------------------------------------
// Training: one workgroup of kTrainNT threads, launched on a stream of its
// own after a slide, overlapping the next slide's ingest and merge on CUs
// they leave idle.  `steps` Adam steps, each on kTrainRows samples drawn by
// `sample` (which fills a wave's 32 rows of xs / ys).  The FP32 master
// weights and the moments are the trainer's own; the result is published to
// the inactive half of the double-buffered weights and then made current,
// so a classify never waits for training and never reads a half-written set.
// ---------------------------------------------------------------------------
template <typename Sample>
__device__ __forceinline__ void train_core(const Learn &L, int steps, TrainWaveLds *lds_wave,
                                           TrainLds &s, Sample &&sample)
{
    LearnShared *ls = L.ls;
    const int tid = threadIdx.x, lane = lane_id(), wid = tid / kWave;
    TrainWaveLds &w = lds_wave[wid];
    constexpr int PER = (kMlpNP + kTrainNT - 1) / kTrainNT;
    float am[PER], av[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int p = tid + q * kTrainNT;
        am[q] = av[q] = 0.f;
        if (p < kMlpNP) {
            s.p[p] = L.params[p];
            s.g[p] = 0.f;
            am[q] = L.adam_m[p];
            av[q] = L.adam_v[p];
        }
    }
    if (tid == 0)
        s.loss = 0.f;
    __syncthreads();
    int t_adam = ls->adam_t;
    for (int it = 0; it < steps; ++it) {
        if (lane < 32) {
            const uint32_t h = mix32((uint32_t)t_adam * 0x9e3779b9u + (uint32_t)(wid * 32 + lane));
            sample(h, w.xs[lane], w.ys[lane]);
        }
        wave_lds_fence();
        float sse = mlp_wave_grads(s.p, s.g, w, wid * 32, kTrainRows);
        sse = cross_group_sum(sse);
        if (lane == 0)
            atomicAdd(&s.loss, sse);
        __syncthreads();
        ++t_adam;
        const float c1 = 1.f - powf(kAdamB1, (float)t_adam);
        const float c2 = 1.f - powf(kAdamB2, (float)t_adam);
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int p = tid + q * kTrainNT;
            if (p < kMlpNP) {
                const float g = s.g[p];
                s.g[p] = 0.f;
                am[q] = kAdamB1 * am[q] + (1.f - kAdamB1) * g;
                av[q] = kAdamB2 * av[q] + (1.f - kAdamB2) * g * g;
                const float mh = am[q] / c1, vh = av[q] / c2;
                s.p[p] -= kAdamLr * mh / (sqrtf(vh) + kAdamEps);
            }
        }
        __syncthreads();
    }
    const int cur = __hip_atomic_load(&ls->pub, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    float *pub = L.pub + (size_t)((cur + 1) & 1) * kMlpNP;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int p = tid + q * kTrainNT;
        if (p < kMlpNP) {
            L.params[p] = s.p[p];
            pub[p] = s.p[p];
            L.adam_m[p] = am[q];
            L.adam_v[p] = av[q];
        }
    }
    __threadfence();
    __syncthreads();
    if (tid == 0) {
        ls->adam_t = t_adam;
        ls->steps_total += steps;
        ls->loss_last = s.loss / (float)(steps * kTrainRows);
        __threadfence();
        __hip_atomic_store(&ls->pub, cur + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// one replay row: features of unit-list index li in slot j, target from its
// capped count
__device__ __forceinline__ void load_replay_row(const Learn &L, int W, int K, int j, int li,
                                                float *xs, float &y)
{
    const size_t e = (size_t)j * W + li;
    const float4 x0 = *reinterpret_cast<const float4 *>(&L.rep_x[e * kFeatStride]);
    const float2 x1 = *reinterpret_cast<const float2 *>(&L.rep_x[e * kFeatStride + 4]);
    xs[0] = x0.x;
    xs[1] = x0.y;
    xs[2] = x0.z;
    xs[3] = x0.w;
    xs[4] = x1.x;
    xs[5] = x1.y;
    xs[6] = 0.f;
    xs[7] = 0.f;
    y = mlp_target(L.rep_c[e], K);
}

// Replay slots a training launch leaves alone: those of slides t_next ..
// t_next + kReplaySkip - 1, which may be queued or running while it trains
// (up to kHostOuts slides are in flight); a slide past them waits for the
// launch before it rewrites its slot.
constexpr int kReplaySkip = kHostOuts;
static_assert(kReplaySkip < kReplaySlides, "replay left to train on");

__device__ __forceinline__ bool replay_skipped(int j, uint64_t t_next)
{
    const int first = (int)(t_next % kReplaySlides);
    return (j - first + kReplaySlides) % kReplaySlides < kReplaySkip;
}

// One stream's replay, uniform over its stored units, the slots of the
// slides from t_next on skipped (replay_skipped).
__global__ __launch_bounds__(kTrainNT) void k_train(Dev d, uint64_t t_next, int steps)
{
    __shared__ TrainWaveLds lds_wave[kTrainNT / kWave];
    __shared__ TrainLds s;
    const Learn &L = d.lr;
    if (L.ls->frozen)
        return;
    if (threadIdx.x == 0) {
        int acc = 0;
        s.pref[0] = 0;
        for (int j = 0; j < kReplaySlides; ++j) {
            if (!replay_skipped(j, t_next))
                acc += L.rep_n[2 * j] + L.rep_n[2 * j + 1];
            s.pref[j + 1] = acc;
        }
    }
    __syncthreads();
    const int N = s.pref[kReplaySlides];
    if (N <= 0)
        return;
    const int W = d.W, K = d.K;
    train_core(L, steps, lds_wave, s, [&](uint32_t h, float *xs, float &y) {
        const int u = (int)(h % (uint32_t)N);
        int j = 0, hi = kReplaySlides - 1; // last slot with pref[j] <= u
        while (j < hi) {
            const int mid = (j + hi + 1) >> 1;
            if (s.pref[mid] <= u)
                j = mid;
            else
                hi = mid - 1;
        }
        const int off = u - s.pref[j];
        const int nO = L.rep_n[2 * j];
        load_replay_row(L, W, K, j, off < nO ? off : W - 1 - (off - nO), xs, y);
    });
}


// ---------------------------------------------------------------------------
// Test entry kernels: the same device code on caller-given inputs.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mlp_forward(const float *P, const float *X, int n,
                                                     float *out)
{
    __shared__ __attribute__((aligned(16))) __bf16 s_h[4][kWave * kMlpH];
    __shared__ float s_x[4][kWave][kFeatStride];
    __shared__ float s_z[4][kWave];
    const int wid = threadIdx.x / kWave;
    const int i = blockIdx.x * 256 + threadIdx.x;
    float f[kMlpIn];
#pragma unroll
    for (int j = 0; j < kMlpIn; ++j)
        f[j] = i < n ? X[(size_t)i * kMlpIn + j] : 0.f;
    const float y = mlp_infer_wave(P, f, s_x[wid], s_h[wid], s_z[wid]);
    if (i < n)
        out[i] = y;
}

// One workgroup; n <= kTrainRows samples.  grad[kMlpNP], loss[0] = mean SSE.
__global__ __launch_bounds__(kTrainNT) void k_mlp_grad(const float *P, const float *X,
                                                       const float *Y, int n, float *grad,
                                                       float *loss)
{
    __shared__ TrainWaveLds s_w[kTrainNT / kWave];
    __shared__ float s_p[kMlpNP + 3], s_g[kMlpNP + 3];
    __shared__ float s_loss;
    const int tid = threadIdx.x, lane = lane_id(), wid = tid / kWave;
    for (int p = tid; p < kMlpNP; p += kTrainNT) {
        s_p[p] = P[p];
        s_g[p] = 0.f;
    }
    if (tid == 0)
        s_loss = 0.f;
    TrainWaveLds &w = s_w[wid];
    if (lane < 32) {
        const int r = wid * 32 + lane;
#pragma unroll
        for (int j = 0; j < kFeatStride; ++j)
            w.xs[lane][j] = (r < n && j < kMlpIn) ? X[(size_t)r * kMlpIn + j] : 0.f;
        w.ys[lane] = r < n ? Y[r] : 0.f;
    }
    __syncthreads();
    float sse = mlp_wave_grads(s_p, s_g, w, wid * 32, n);
    sse = cross_group_sum(sse);
    if (lane == 0)
        atomicAdd(&s_loss, sse);
    __syncthreads();
    for (int p = tid; p < kMlpNP; p += kTrainNT)
        grad[p] = s_g[p];
    if (tid == 0)
        loss[0] = s_loss / (float)n;
}

} // namespace rtod
