/*
 * synth.c -- deterministic synthetic streams for the measured workloads.
 *
 * xoshiro256** (Blackman & Vigna) seeded by splitmix64, Box-Muller normals
 * in double, points cast to float32.
 *   rtod_synth_walk: per-axis random walk from `start` with N(0, sigma)
 *     increments; with probability p_spike a point is displaced (the walk
 *     itself continues undisplaced) by U(spike_lo, spike_hi) along a random
 *     direction (1-D: +-1 equiprobable).
 *   rtod_synth_gmm: ncomp Gaussians with centres U[lo,hi]^dim and per-axis
 *     sigma U(sig_lo, sig_hi), centres drifting linearly by `drift` in a
 *     random direction over the stream; a fraction noise_frac of uniform
 *     background points in [lo,hi]^dim; components chosen uniformly.
 * Output: out[n][dim] row-major.  Compiled with -ffp-contract=off.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#define RTOD_SYNTH_API __attribute__((visibility("default")))

static uint64_t splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

typedef struct {
    uint64_t s[4];
    int have_spare;
    double spare;
} rng_t;

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static void rng_seed(rng_t *r, uint64_t seed)
{
    uint64_t sm = seed;
    for (int i = 0; i < 4; i++)
        r->s[i] = splitmix64(&sm);
    r->have_spare = 0;
    r->spare = 0.0;
}

/* xoshiro256** (Blackman & Vigna) */
static uint64_t rng_next(rng_t *r)
{
    uint64_t *s = r->s;
    uint64_t result = rotl64(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}

static double rng_u01(rng_t *r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }

static double rng_uniform(rng_t *r, double lo, double hi) { return lo + (hi - lo) * rng_u01(r); }

/* Box-Muller, both outputs used */
static double rng_normal(rng_t *r)
{
    if (r->have_spare) {
        r->have_spare = 0;
        return r->spare;
    }
    double u1 = 1.0 - rng_u01(r); /* (0,1] */
    double u2 = rng_u01(r);
    double rad = sqrt(-2.0 * log(u1));
    double th = 6.283185307179586476925 * u2;
    r->spare = rad * sin(th);
    r->have_spare = 1;
    return rad * cos(th);
}

RTOD_SYNTH_API void rtod_synth_walk(uint64_t seed, int64_t n, int dim, double start, double sigma,
                                      double p_spike, double spike_lo, double spike_hi,
                                      float *out)
{
    rng_t r;
    rng_seed(&r, seed);
    double w[3] = {start, start, start};
    for (int64_t i = 0; i < n; i++) {
        if (i > 0)
            for (int a = 0; a < dim; a++)
                w[a] += sigma * rng_normal(&r);
        double off[3] = {0.0, 0.0, 0.0};
        if (rng_u01(&r) < p_spike) {
            double mag = rng_uniform(&r, spike_lo, spike_hi);
            if (dim == 1) {
                off[0] = (rng_u01(&r) < 0.5) ? -mag : mag;
            } else {
                double v[3] = {0, 0, 0}, nrm = 0.0;
                for (int a = 0; a < dim; a++) {
                    v[a] = rng_normal(&r);
                    nrm += v[a] * v[a];
                }
                nrm = sqrt(nrm);
                if (nrm == 0.0) {
                    v[0] = 1.0;
                    nrm = 1.0;
                }
                for (int a = 0; a < dim; a++)
                    off[a] = mag * v[a] / nrm;
            }
        }
        for (int a = 0; a < dim; a++)
            out[i * dim + a] = (float)(w[a] + off[a]);
    }
}

RTOD_SYNTH_API void rtod_synth_gmm(uint64_t seed, int64_t n, int dim, int ncomp, double lo, double hi,
                                     double sig_lo, double sig_hi, double noise_frac,
                                     double drift, float *out)
{
    rng_t r;
    rng_seed(&r, seed);
    if (ncomp < 1)
        ncomp = 1;
    double *cen = (double *)malloc(sizeof(double) * (size_t)ncomp * 3);
    double *sig = (double *)malloc(sizeof(double) * (size_t)ncomp * 3);
    double *dv = (double *)malloc(sizeof(double) * (size_t)ncomp * 3);
    for (int k = 0; k < ncomp; k++) {
        for (int a = 0; a < dim; a++)
            cen[k * 3 + a] = rng_uniform(&r, lo, hi);
        for (int a = 0; a < dim; a++)
            sig[k * 3 + a] = rng_uniform(&r, sig_lo, sig_hi);
        double nrm = 0.0;
        for (int a = 0; a < dim; a++) {
            dv[k * 3 + a] = rng_normal(&r);
            nrm += dv[k * 3 + a] * dv[k * 3 + a];
        }
        nrm = sqrt(nrm);
        for (int a = 0; a < dim; a++)
            dv[k * 3 + a] = nrm > 0.0 ? drift * dv[k * 3 + a] / nrm : 0.0;
    }
    for (int64_t i = 0; i < n; i++) {
        double frac = n > 1 ? (double)i / (double)(n - 1) : 0.0;
        if (rng_u01(&r) < noise_frac) {
            for (int a = 0; a < dim; a++)
                out[i * dim + a] = (float)rng_uniform(&r, lo, hi);
        } else {
            int k = (int)(rng_u01(&r) * ncomp);
            if (k >= ncomp)
                k = ncomp - 1;
            for (int a = 0; a < dim; a++) {
                double v = cen[k * 3 + a] + dv[k * 3 + a] * frac + sig[k * 3 + a] * rng_normal(&r);
                out[i * dim + a] = (float)v;
            }
        }
    }
    free(cen);
    free(sig);
    free(dv);
}

