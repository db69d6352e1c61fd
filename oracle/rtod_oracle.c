/*
 * rtod_oracle.c -- CPU restatement of RTOD's per-slide outlier detection.
 *
 * TEST INFRASTRUCTURE ONLY (checker + timed CPU baseline); see rtod_oracle.h
 * for the contract and the reference lines each function follows.  Built with
 * -ffp-contract=off so that the only fused operations are the explicit fmaf()
 * calls of the predicate.
 */
#define _GNU_SOURCE
#include "rtod_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------ */
/* predicate (contract; reference optixScan.cu:170-189 uses FP64 `<`)  */
/* ------------------------------------------------------------------ */
float rtod_o_d2(const float *p, const float *q, int dim)
{
    float dx = p[0] - q[0];
    float d2 = dx * dx;
    if (dim > 1) {
        float dy = p[1] - q[1];
        d2 = fmaf(dy, dy, d2);
    }
    if (dim > 2) {
        float dz = p[2] - q[2];
        d2 = fmaf(dz, dz, d2);
    }
    return d2;
}

static inline int pred_contract(const float *p, const float *q, int dim, float r2)
{
    return rtod_o_d2(p, q, dim) <= r2;
}

/* Reference rule: src/ods/optixScan.cu:170-189 (FP64, strict), R2 = R*R in
 * double (optixScan.cpp:713-714); inputs are the f32 values widened. */
static inline int pred_reference(const float *p, const float *q, int dim, double R, double R2)
{
    if (dim == 1)
        return fabs((double)p[0] - (double)q[0]) < R;
    double s = 0.0;
    for (int a = 0; a < dim; a++) {
        double o = (double)p[a] - (double)q[a];
        s = s + o * o;
    }
    return s < R2;
}

/* ------------------------------------------------------------------ */
/* small parallel-for on pthreads                                       */
/* ------------------------------------------------------------------ */
typedef void (*range_fn)(void *ctx, int64_t lo, int64_t hi, int tid);

typedef struct {
    range_fn fn;
    void *ctx;
    int64_t n, chunk;
    atomic_llong next;
    int tid;
} pf_job;

typedef struct {
    pf_job *job;
    int tid;
} pf_arg;

static void *pf_worker(void *p)
{
    pf_arg *a = (pf_arg *)p;
    pf_job *j = a->job;
    for (;;) {
        int64_t lo = atomic_fetch_add(&j->next, j->chunk);
        if (lo >= j->n)
            break;
        int64_t hi = lo + j->chunk < j->n ? lo + j->chunk : j->n;
        j->fn(j->ctx, lo, hi, a->tid);
    }
    return NULL;
}

static void parallel_for(int nthreads, int64_t n, int64_t chunk, range_fn fn, void *ctx)
{
    if (n <= 0)
        return;
    if (nthreads <= 1 || n <= chunk) {
        fn(ctx, 0, n, 0);
        return;
    }
    pf_job job;
    job.fn = fn;
    job.ctx = ctx;
    job.n = n;
    job.chunk = chunk;
    atomic_init(&job.next, 0);
    pthread_t th[256];
    pf_arg args[256];
    if (nthreads > 256)
        nthreads = 256;
    for (int t = 1; t < nthreads; t++) {
        args[t].job = &job;
        args[t].tid = t;
        pthread_create(&th[t], NULL, pf_worker, &args[t]);
    }
    args[0].job = &job;
    args[0].tid = 0;
    pf_worker(&args[0]);
    for (int t = 1; t < nthreads; t++)
        pthread_join(th[t], NULL);
}

/* persistent spinning pool for the per-slide range counts of the grid
 * oracle (thread creation per slide would dominate the baseline) */
typedef struct {
    pthread_t th[256];
    int nthreads;
    atomic_int gen, done, stop;
    pf_job job;
} pool_t;

typedef struct {
    pool_t *pool;
    int tid;
} pool_arg;

static void pool_run_share(pf_job *j, int tid)
{
    for (;;) {
        int64_t lo = atomic_fetch_add(&j->next, j->chunk);
        if (lo >= j->n)
            break;
        int64_t hi = lo + j->chunk < j->n ? lo + j->chunk : j->n;
        j->fn(j->ctx, lo, hi, tid);
    }
}

static void *pool_worker(void *p)
{
    pool_arg *a = (pool_arg *)p;
    pool_t *pl = a->pool;
    int tid = a->tid;
    free(a);
    int seen = 0;
    for (;;) {
        long spins = 0;
        while (atomic_load_explicit(&pl->gen, memory_order_acquire) == seen) {
            if (atomic_load(&pl->stop))
                return NULL;
            if (++spins > 200000) {
                struct timespec ts = {0, 20000};
                nanosleep(&ts, NULL);
            } else {
                __builtin_ia32_pause();
            }
        }
        seen = atomic_load(&pl->gen);
        pool_run_share(&pl->job, tid);
        atomic_fetch_add_explicit(&pl->done, 1, memory_order_release);
    }
}

static pool_t *pool_create(int nthreads)
{
    pool_t *pl = (pool_t *)calloc(1, sizeof(pool_t));
    pl->nthreads = nthreads > 256 ? 256 : nthreads;
    atomic_init(&pl->gen, 0);
    atomic_init(&pl->done, 0);
    atomic_init(&pl->stop, 0);
    for (int t = 1; t < pl->nthreads; t++) {
        pool_arg *a = (pool_arg *)malloc(sizeof(pool_arg));
        a->pool = pl;
        a->tid = t;
        pthread_create(&pl->th[t], NULL, pool_worker, a);
    }
    return pl;
}

static void pool_for(pool_t *pl, int64_t n, int64_t chunk, range_fn fn, void *ctx)
{
    if (n <= 0)
        return;
    if (!pl || pl->nthreads <= 1 || n <= chunk) {
        fn(ctx, 0, n, 0);
        return;
    }
    pl->job.fn = fn;
    pl->job.ctx = ctx;
    pl->job.n = n;
    pl->job.chunk = chunk;
    atomic_store(&pl->job.next, 0);
    atomic_store(&pl->done, 0);
    atomic_fetch_add_explicit(&pl->gen, 1, memory_order_release);
    pool_run_share(&pl->job, 0);
    while (atomic_load_explicit(&pl->done, memory_order_acquire) < pl->nthreads - 1)
        __builtin_ia32_pause();
}

static void pool_destroy(pool_t *pl)
{
    if (!pl)
        return;
    atomic_store(&pl->stop, 1);
    for (int t = 1; t < pl->nthreads; t++)
        pthread_join(pl->th[t], NULL);
    free(pl);
}

static int cmp_i32(const void *a, const void *b)
{
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/* ------------------------------------------------------------------ */
/* brute force (optixScan.cpp:817-838, set semantics)                  */
/* ------------------------------------------------------------------ */
typedef struct {
    const float *win;
    int64_t W;
    int dim, K, semantics;
    float r2;
    double R, R2;
    int32_t *counts;
} brute_ctx;

static void brute_range(void *p, int64_t lo, int64_t hi, int tid)
{
    (void)tid;
    brute_ctx *c = (brute_ctx *)p;
    const int K1 = c->K + 1;
    for (int64_t i = lo; i < hi; i++) {
        const float *pi = c->win + i * c->dim;
        /* scan outward from i: neighbours of a stream point tend to be near
         * it in time, so early exit comes sooner; the order is irrelevant to
         * the result. */
#define BRUTE_HIT(j)                                                              \
    (c->semantics == 0 ? pred_contract(pi, c->win + (j) * c->dim, c->dim, c->r2)  \
                       : pred_reference(pi, c->win + (j) * c->dim, c->dim, c->R, c->R2))
        int cnt = BRUTE_HIT(i);
        for (int64_t h = 1; cnt < K1 && (i - h >= 0 || i + h < c->W); h++) {
            if (i + h < c->W) {
                cnt += BRUTE_HIT(i + h);
                if (cnt >= K1)
                    break;
            }
            if (i - h >= 0)
                cnt += BRUTE_HIT(i - h);
        }
#undef BRUTE_HIT
        c->counts[i] = cnt;
    }
}

int64_t rtod_o_brute(const float *win, int64_t W, int dim, float R, int K, int semantics,
                     int nthreads, int32_t *out_slots, int32_t *out_counts)
{
    if (W <= 0 || dim < 1 || dim > 3 || K < 0)
        return -1;
    brute_ctx c;
    c.win = win;
    c.W = W;
    c.dim = dim;
    c.K = K;
    c.semantics = semantics;
    c.r2 = R * R;
    c.R = (double)R;
    c.R2 = (double)R * (double)R;
    int32_t *counts = out_counts ? out_counts : (int32_t *)malloc(sizeof(int32_t) * (size_t)W);
    c.counts = counts;
    parallel_for(nthreads, W, 64, brute_range, &c);
    int64_t n = 0;
    for (int64_t i = 0; i < W; i++)
        if (counts[i] <= K)
            out_slots[n++] = (int32_t)i;
    if (!out_counts)
        free(counts);
    return n;
}

/* ------------------------------------------------------------------ */
/* 1-D sorted-window oracle                                            */
/* ------------------------------------------------------------------ */
typedef struct {
    float v;
    int32_t slot;
} vs_t;

static int cmp_vs(const void *a, const void *b)
{
    const vs_t *x = (const vs_t *)a, *y = (const vs_t *)b;
    if (x->v < y->v)
        return -1;
    if (x->v > y->v)
        return 1;
    return (x->slot > y->slot) - (x->slot < y->slot);
}

int64_t rtod_o_sorted1d(const float *win, int64_t W, float R, int K, int32_t *out_slots,
                        int32_t *out_counts)
{
    if (W <= 0)
        return -1;
    vs_t *a = (vs_t *)malloc(sizeof(vs_t) * (size_t)W);
    for (int64_t i = 0; i < W; i++) {
        a[i].v = win[i];
        a[i].slot = (int32_t)i;
    }
    qsort(a, (size_t)W, sizeof(vs_t), cmp_vs);
    const float r2 = R * R;
    int64_t n = 0;
    for (int64_t i = 0; i < W; i++) {
        const float p = a[i].v;
        /* lo: first j in [0,i] with pred true (pred is a suffix of [0,i]) */
        int64_t L = 0, H = i;
        while (L < H) {
            int64_t mid = L + (H - L) / 2;
            if (pred_contract(&p, &a[mid].v, 1, r2))
                H = mid;
            else
                L = mid + 1;
        }
        int64_t lo = L;
        /* hi: last j in [i,W) with pred true (pred is a prefix of [i,W)) */
        L = i;
        H = W - 1;
        while (L < H) {
            int64_t mid = L + (H - L + 1) / 2;
            if (pred_contract(&p, &a[mid].v, 1, r2))
                L = mid;
            else
                H = mid - 1;
        }
        int64_t hi = L;
        int64_t cnt = hi - lo + 1;
        if (out_counts)
            out_counts[a[i].slot] = (int32_t)(cnt > K + 1 ? K + 1 : cnt);
        if (cnt <= K)
            out_slots[n++] = a[i].slot;
    }
    free(a);
    qsort(out_slots, (size_t)n, sizeof(int32_t), cmp_i32);
    return n;
}

/* ------------------------------------------------------------------ */
/* grid oracle                                                          */
/* ------------------------------------------------------------------ */
/* Cell side c = R (1 - 1e-5) / sqrt(d), cell coordinate floor(x * (1/c)) in
 * double, clamped to (-2^B, 2^B - 1) with B = 30 (d <= 2) or 20 (d = 3).
 * Two points of one unclamped cell differ by < c (1 + 1e-6) per axis, so
 * their FP32 d2 is below fl(R*R): a cell with >= K+1 points is provably
 * dense (reference rule `num > K`, optixScan.cpp:987-993).  Clamped cells
 * are never taken as dense.  Key packing: d=1 k0+2^30; d=2 (k0+2^30)<<31 |
 * (k1+2^30); d=3 21-bit fields, k2 least significant. */
#define CELL_SHRINK 1e-5
#define STENCIL_GROW 1e-5

static inline int cell_bits(int dim) { return dim == 3 ? 20 : 30; }

static inline int64_t cell_coord(double v, double inv_c, int B)
{
    double f = floor(v * inv_c);
    double lo = -ldexp(1.0, B), hi = ldexp(1.0, B) - 1.0;
    if (!(f > lo))
        f = lo;
    else if (!(f < hi))
        f = hi;
    return (int64_t)f;
}

static inline uint64_t pack_key(const int64_t *k, int dim, int B)
{
    uint64_t bias = 1ull << B;
    if (dim == 1)
        return (uint64_t)(k[0] + (int64_t)bias);
    if (dim == 2)
        return ((uint64_t)(k[0] + (int64_t)bias) << 31) | (uint64_t)(k[1] + (int64_t)bias);
    return ((uint64_t)(k[0] + (int64_t)bias) << 42) | ((uint64_t)(k[1] + (int64_t)bias) << 21) |
           (uint64_t)(k[2] + (int64_t)bias);
}

static inline int guaranteed_coords(const int64_t *k, int dim, int B)
{
    int64_t lo = -(1ll << B), hi = (1ll << B) - 1;
    for (int a = 0; a < dim; a++)
        if (!(k[a] > lo && k[a] < hi))
            return 0;
    return 1;
}

static void grid_geometry(int dim, float R, double *c, double *inv_c, double *Rp)
{
    *c = (double)R * (1.0 - CELL_SHRINK) / sqrt((double)dim);
    *inv_c = 1.0 / *c;
    *Rp = (double)R * (1.0 + STENCIL_GROW);
}

uint64_t rtod_o_cell_key(const float *p, int dim, float R)
{
    double c, inv_c, Rp;
    grid_geometry(dim, R, &c, &inv_c, &Rp);
    int B = cell_bits(dim);
    int64_t k[3] = {0, 0, 0};
    for (int a = 0; a < dim; a++)
        k[a] = cell_coord((double)p[a], inv_c, B);
    return pack_key(k, dim, B);
}

typedef struct {
    int32_t *buf;
    int32_t cap, head, n;
} fifo_t;

static void fifo_push(fifo_t *f, int32_t v)
{
    if (f->n == f->cap) {
        int32_t ncap = f->cap ? f->cap * 2 : 8;
        int32_t *nb = (int32_t *)malloc(sizeof(int32_t) * (size_t)ncap);
        for (int32_t i = 0; i < f->n; i++)
            nb[i] = f->buf[(f->head + i) % f->cap];
        free(f->buf);
        f->buf = nb;
        f->cap = ncap;
        f->head = 0;
    }
    f->buf[(f->head + f->n) % f->cap] = v;
    f->n++;
}

static int32_t fifo_pop(fifo_t *f)
{
    int32_t v = f->buf[f->head];
    f->head = (f->head + 1) % f->cap;
    f->n--;
    return v;
}

struct rtod_o_grid {
    int dim, K, nthreads, B;
    int64_t W, S, m, t;
    float R, r2;
    double c, inv_c, Rp;
    float *ring;
    int32_t *cell_of_slot;
    /* cells */
    int64_t ncells, cells_cap;
    uint64_t *ckey;
    int32_t *ccount;
    uint8_t *cguar;
    fifo_t *cfifo;
    int32_t *cact_pos;
    int32_t *active;
    int64_t nactive;
    /* hash: cell index + 1, 0 = empty */
    int32_t *htab;
    uint64_t hcap;
    /* per slide */
    int32_t *cand;
    int32_t *cand_cnt;
    int64_t ncand, ndense;
    int64_t hits; /* sum over candidates of min(count, K + 1), last slide */
    int64_t *tests_per_thread;
    int64_t pair_tests;
    pool_t *pool;
};

static inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static int64_t grid_lookup(const rtod_o_grid *g, uint64_t key)
{
    uint64_t h = mix64(key) & (g->hcap - 1);
    for (;;) {
        int32_t e = g->htab[h];
        if (e == 0)
            return -1;
        if (g->ckey[e - 1] == key)
            return e - 1;
        h = (h + 1) & (g->hcap - 1);
    }
}

static void grid_rehash(rtod_o_grid *g, uint64_t newcap)
{
    free(g->htab);
    g->hcap = newcap;
    g->htab = (int32_t *)calloc((size_t)newcap, sizeof(int32_t));
    for (int64_t ci = 0; ci < g->ncells; ci++) {
        uint64_t h = mix64(g->ckey[ci]) & (newcap - 1);
        while (g->htab[h] != 0)
            h = (h + 1) & (newcap - 1);
        g->htab[h] = (int32_t)(ci + 1);
    }
}

/* drop empty cells, renumber, rehash */
static void grid_compact(rtod_o_grid *g)
{
    int64_t *remap = (int64_t *)malloc(sizeof(int64_t) * (size_t)g->ncells);
    int64_t n = 0;
    for (int64_t ci = 0; ci < g->ncells; ci++) {
        if (g->ccount[ci] > 0) {
            remap[ci] = n;
            g->ckey[n] = g->ckey[ci];
            g->ccount[n] = g->ccount[ci];
            g->cguar[n] = g->cguar[ci];
            g->cfifo[n] = g->cfifo[ci];
            n++;
        } else {
            remap[ci] = -1;
            free(g->cfifo[ci].buf);
        }
    }
    for (int64_t s = 0; s < g->W; s++)
        if (g->cell_of_slot[s] >= 0)
            g->cell_of_slot[s] = (int32_t)remap[g->cell_of_slot[s]];
    g->ncells = n;
    g->nactive = 0;
    for (int64_t ci = 0; ci < n; ci++) {
        g->cact_pos[ci] = (int32_t)g->nactive;
        g->active[g->nactive++] = (int32_t)ci;
    }
    free(remap);
    uint64_t cap = g->hcap;
    while ((uint64_t)n * 4 > cap)
        cap *= 2;
    grid_rehash(g, cap);
}

static int64_t grid_insert_cell(rtod_o_grid *g, uint64_t key, int guar)
{
    if ((uint64_t)(g->ncells + 1) * 2 > g->hcap)
        grid_compact(g);
    if (g->ncells == g->cells_cap) {
        int64_t nc = g->cells_cap * 2;
        g->ckey = (uint64_t *)realloc(g->ckey, sizeof(uint64_t) * (size_t)nc);
        g->ccount = (int32_t *)realloc(g->ccount, sizeof(int32_t) * (size_t)nc);
        g->cguar = (uint8_t *)realloc(g->cguar, (size_t)nc);
        g->cfifo = (fifo_t *)realloc(g->cfifo, sizeof(fifo_t) * (size_t)nc);
        g->cact_pos = (int32_t *)realloc(g->cact_pos, sizeof(int32_t) * (size_t)nc);
        g->active = (int32_t *)realloc(g->active, sizeof(int32_t) * (size_t)nc);
        g->cells_cap = nc;
    }
    int64_t ci = g->ncells++;
    g->ckey[ci] = key;
    g->ccount[ci] = 0;
    g->cguar[ci] = (uint8_t)guar;
    memset(&g->cfifo[ci], 0, sizeof(fifo_t));
    g->cact_pos[ci] = -1;
    uint64_t h = mix64(key) & (g->hcap - 1);
    while (g->htab[h] != 0)
        h = (h + 1) & (g->hcap - 1);
    g->htab[h] = (int32_t)(ci + 1);
    return ci;
}

static void grid_add_point(rtod_o_grid *g, int32_t slot)
{
    const float *p = g->ring + (int64_t)slot * g->dim;
    int64_t k[3] = {0, 0, 0};
    for (int a = 0; a < g->dim; a++)
        k[a] = cell_coord((double)p[a], g->inv_c, g->B);
    uint64_t key = pack_key(k, g->dim, g->B);
    int64_t ci = grid_lookup(g, key);
    if (ci < 0)
        ci = grid_insert_cell(g, key, guaranteed_coords(k, g->dim, g->B));
    fifo_push(&g->cfifo[ci], slot);
    if (g->ccount[ci]++ == 0) {
        g->cact_pos[ci] = (int32_t)g->nactive;
        g->active[g->nactive++] = (int32_t)ci;
    }
    g->cell_of_slot[slot] = (int32_t)ci;
}

static void grid_remove_point(rtod_o_grid *g, int32_t slot)
{
    int32_t ci = g->cell_of_slot[slot];
    if (ci < 0)
        return;
    /* points leave a cell in arrival order: the expired slot is the FIFO
     * front (the reference decrements FixQueue::num, state.h:15-57) */
    int32_t v = fifo_pop(&g->cfifo[ci]);
    (void)v;
    if (--g->ccount[ci] == 0) {
        int32_t pos = g->cact_pos[ci];
        int32_t last = g->active[--g->nactive];
        g->active[pos] = last;
        g->cact_pos[last] = pos;
        g->cact_pos[ci] = -1;
    }
    g->cell_of_slot[slot] = -1;
}

rtod_o_grid *rtod_o_grid_create(int dim, int64_t W, int64_t S, float R, int K, int nthreads)
{
    if (dim < 1 || dim > 3 || W <= 0 || S <= 0 || S > W || W % S != 0 || K < 1 || !(R > 0.0f))
        return NULL;
    rtod_o_grid *g = (rtod_o_grid *)calloc(1, sizeof(rtod_o_grid));
    g->dim = dim;
    g->K = K;
    g->nthreads = nthreads < 1 ? 1 : nthreads;
    g->B = cell_bits(dim);
    g->W = W;
    g->S = S;
    g->m = W / S;
    g->R = R;
    g->r2 = R * R;
    grid_geometry(dim, R, &g->c, &g->inv_c, &g->Rp);
    g->ring = (float *)calloc((size_t)(W * dim), sizeof(float));
    g->cell_of_slot = (int32_t *)malloc(sizeof(int32_t) * (size_t)W);
    for (int64_t s = 0; s < W; s++)
        g->cell_of_slot[s] = -1;
    g->cells_cap = 1024;
    g->ckey = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)g->cells_cap);
    g->ccount = (int32_t *)malloc(sizeof(int32_t) * (size_t)g->cells_cap);
    g->cguar = (uint8_t *)malloc((size_t)g->cells_cap);
    g->cfifo = (fifo_t *)malloc(sizeof(fifo_t) * (size_t)g->cells_cap);
    g->cact_pos = (int32_t *)malloc(sizeof(int32_t) * (size_t)g->cells_cap);
    g->active = (int32_t *)malloc(sizeof(int32_t) * (size_t)g->cells_cap);
    g->hcap = 4096;
    g->htab = (int32_t *)calloc((size_t)g->hcap, sizeof(int32_t));
    g->cand = (int32_t *)malloc(sizeof(int32_t) * (size_t)W);
    g->cand_cnt = (int32_t *)malloc(sizeof(int32_t) * (size_t)W);
    g->tests_per_thread = (int64_t *)calloc(256, sizeof(int64_t));
    g->pool = g->nthreads > 1 ? pool_create(g->nthreads) : NULL;
    return g;
}

void rtod_o_grid_destroy(rtod_o_grid *g)
{
    if (!g)
        return;
    for (int64_t ci = 0; ci < g->ncells; ci++)
        free(g->cfifo[ci].buf);
    free(g->ring);
    free(g->cell_of_slot);
    free(g->ckey);
    free(g->ccount);
    free(g->cguar);
    free(g->cfifo);
    free(g->cact_pos);
    free(g->active);
    free(g->htab);
    free(g->cand);
    free(g->cand_cnt);
    free(g->tests_per_thread);
    pool_destroy(g->pool);
    free(g);
}

int rtod_o_grid_prime(rtod_o_grid *g, const float *pts)
{
    for (int64_t s = 0; s < g->W; s++)
        if (g->cell_of_slot[s] >= 0)
            grid_remove_point(g, (int32_t)s);
    memcpy(g->ring, pts, sizeof(float) * (size_t)(g->W * g->dim));
    for (int64_t s = 0; s < g->W; s++)
        grid_add_point(g, (int32_t)s);
    g->t = 0;
    return 0;
}

/* exact range count of one candidate with early exit at K+1
 * (optixScan.cu:158-199 counts hits per candidate box; early-outs at > K) */
static int grid_count(const rtod_o_grid *g, int32_t slot, int64_t *tests)
{
    const int dim = g->dim, K1 = g->K + 1;
    const float *p = g->ring + (int64_t)slot * dim;
    int64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int a = 0; a < dim; a++) {
        lo[a] = cell_coord((double)p[a] - g->Rp, g->inv_c, g->B);
        hi[a] = cell_coord((double)p[a] + g->Rp, g->inv_c, g->B);
    }
    int cnt = 0;
    int64_t k[3];
    for (k[0] = lo[0]; k[0] <= hi[0]; k[0]++)
        for (k[1] = lo[1]; k[1] <= hi[1]; k[1]++)
            for (k[2] = lo[2]; k[2] <= hi[2]; k[2]++) {
                int64_t ci = grid_lookup(g, pack_key(k, dim, g->B));
                if (ci < 0)
                    continue;
                const fifo_t *f = &g->cfifo[ci];
                for (int32_t i = 0; i < f->n; i++) {
                    int32_t q = f->buf[(f->head + i) % f->cap];
                    (*tests)++;
                    if (pred_contract(p, g->ring + (int64_t)q * dim, dim, g->r2))
                        if (++cnt >= K1)
                            return cnt;
                }
            }
    return cnt;
}

static void grid_count_range(void *ctx, int64_t lo, int64_t hi, int tid)
{
    rtod_o_grid *g = (rtod_o_grid *)ctx;
    int64_t tests = 0;
    for (int64_t i = lo; i < hi; i++)
        g->cand_cnt[i] = grid_count(g, g->cand[i], &tests);
    g->tests_per_thread[tid] += tests;
}

int64_t rtod_o_grid_slide(rtod_o_grid *g, const float *new_pts, int32_t *out_slots)
{
    const int64_t S = g->S, dim = g->dim;
    const int64_t base = (g->t % g->m) * S; /* optixScan.cpp:1053-1057, :1106 */
    for (int64_t j = 0; j < S; j++)
        grid_remove_point(g, (int32_t)(base + j));
    memcpy(g->ring + base * dim, new_pts, sizeof(float) * (size_t)(S * dim));
    for (int64_t j = 0; j < S; j++)
        grid_add_point(g, (int32_t)(base + j));

    /* classify cells: dense (>= K+1 points, unclamped) or candidates */
    g->ncand = 0;
    g->ndense = 0;
    for (int64_t a = 0; a < g->nactive; a++) {
        int32_t ci = g->active[a];
        if (g->ccount[ci] > g->K && g->cguar[ci]) {
            g->ndense++;
            continue;
        }
        const fifo_t *f = &g->cfifo[ci];
        for (int32_t i = 0; i < f->n; i++)
            g->cand[g->ncand++] = f->buf[(f->head + i) % f->cap];
    }
    for (int t = 0; t < g->nthreads && t < 256; t++)
        g->tests_per_thread[t] = 0;
    if (g->ncand >= 256)
        pool_for(g->pool, g->ncand, 32, grid_count_range, g);
    else
        grid_count_range(g, 0, g->ncand, 0);
    g->pair_tests = 0;
    for (int t = 0; t < g->nthreads && t < 256; t++)
        g->pair_tests += g->tests_per_thread[t];

    int64_t n = 0;
    g->hits = 0;
    for (int64_t i = 0; i < g->ncand; i++) {
        g->hits += g->cand_cnt[i]; /* min(count, K + 1): the count stops at K + 1 */
        if (g->cand_cnt[i] <= g->K)
            out_slots[n++] = g->cand[i];
    }
    qsort(out_slots, (size_t)n, sizeof(int32_t), cmp_i32);
    g->t++;
    return n;
}

void rtod_o_grid_stats(const rtod_o_grid *g, int64_t *out5)
{
    out5[0] = g->ncand;
    out5[1] = g->nactive;
    out5[2] = g->pair_tests;
    out5[3] = g->ndense;
    out5[4] = g->hits;
}

/* ------------------------------------------------------------------ */
/* stream driver                                                        */
/* ------------------------------------------------------------------ */
static double now_us(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec * 1e6 + (double)ts.tv_nsec * 1e-3;
}

int64_t rtod_o_run_stream(int method, const float *xs, int dim, int64_t W, int64_t S, int64_t T,
                          float R, int K, int nthreads, int32_t *n_out, int32_t *out, int64_t cap,
                          double *us_per_slide)
{
    if (dim < 1 || dim > 3 || W <= 0 || S <= 0 || W % S != 0 || T < 0 || K < 1 || !(R > 0.0f))
        return -1;
    if (method == 2 && dim != 1)
        return -1;
    int32_t *buf = (int32_t *)malloc(sizeof(int32_t) * (size_t)W);
    int64_t total = 0;
    double acc = 0.0;
    if (method == 0) {
        rtod_o_grid *g = rtod_o_grid_create(dim, W, S, R, K, nthreads);
        if (!g) {
            free(buf);
            return -1;
        }
        rtod_o_grid_prime(g, xs);
        for (int64_t t = 1; t <= T; t++) {
            double t0 = now_us();
            int64_t n = rtod_o_grid_slide(g, xs + (W + (t - 1) * S) * dim, buf);
            acc += now_us() - t0;
            n_out[t - 1] = (int32_t)n;
            if (total + n > cap) {
                rtod_o_grid_destroy(g);
                free(buf);
                return -1;
            }
            memcpy(out + total, buf, sizeof(int32_t) * (size_t)n);
            total += n;
        }
        rtod_o_grid_destroy(g);
    } else {
        float *win = (float *)malloc(sizeof(float) * (size_t)(W * dim));
        for (int64_t t = 1; t <= T; t++) {
            for (int64_t gi = t * S; gi < t * S + W; gi++)
                memcpy(win + (gi % W) * dim, xs + gi * dim, sizeof(float) * (size_t)dim);
            double t0 = now_us();
            int64_t n;
            if (method == 1)
                n = rtod_o_brute(win, W, dim, R, K, 0, nthreads, buf, NULL);
            else if (method == 2)
                n = rtod_o_sorted1d(win, W, R, K, buf, NULL);
            else
                n = rtod_o_brute(win, W, dim, R, K, 1, nthreads, buf, NULL);
            acc += now_us() - t0;
            n_out[t - 1] = (int32_t)n;
            if (n < 0 || total + n > cap) {
                free(win);
                free(buf);
                return -1;
            }
            memcpy(out + total, buf, sizeof(int32_t) * (size_t)n);
            total += n;
        }
        free(win);
    }
    if (us_per_slide)
        *us_per_slide = T > 0 ? acc / (double)T : 0.0;
    free(buf);
    return total;
}

/* ------------------------------------------------------------------ */
/* contract vs reference semantics                                      */
/* ------------------------------------------------------------------ */
/* For one window, the full neighbour count of every point under both rules
 * -- the contract (FP32 d2 <= r2, pred_contract) and the reference's FP64
 * strict test (pred_reference, optixScan.cu:170-189) -- and the points whose
 * outlier status (count <= K) differs.  1-D: both neighbour sets of a point
 * are index intervals of the sorted window containing the point (each
 * predicate is monotone in |x_i - x_j|), found by binary search.  2-D/3-D:
 * cells of side R (1 + 1e-4), every pair of neighbouring cells enumerated
 * (a pair passing either rule is closer than R (1 + 1e-6) in exact
 * arithmetic, so it lies in adjacent cells).  out3: [0] points whose status
 * differs, [1] points whose counts differ, [2] ordered pairs on which the
 * rules differ. */
typedef struct {
    uint64_t key;
    int32_t slot;
} ks_t;

static int cmp_ks(const void *a, const void *b)
{
    const ks_t *x = (const ks_t *)a, *y = (const ks_t *)b;
    if (x->key != y->key)
        return x->key < y->key ? -1 : 1;
    return (x->slot > y->slot) - (x->slot < y->slot);
}

typedef struct {
    const float *win;
    int dim, K;
    int64_t W;
    float r2;
    double R, R2, inv_c;
    const vs_t *a1;  /* 1-D: window sorted by value */
    const ks_t *ks;  /* 2-D/3-D: window sorted by cell key */
    int64_t acc[256][3];
} sdiff_ctx;

static int64_t ks_lower(const ks_t *ks, int64_t n, uint64_t key)
{
    int64_t L = 0, H = n;
    while (L < H) {
        int64_t mid = L + (H - L) / 2;
        if (ks[mid].key < key)
            L = mid + 1;
        else
            H = mid;
    }
    return L;
}

static uint64_t sdiff_key(const int64_t *k, int dim)
{
    /* 21 bits per axis around a bias; coordinates beyond are clamped (a
     * pair there is compared with the clamped cells' points, still exactly) */
    uint64_t key = 0;
    for (int a = 0; a < dim; a++) {
        int64_t v = k[a] + (1ll << 20);
        if (v < 0)
            v = 0;
        if (v > (1ll << 21) - 1)
            v = (1ll << 21) - 1;
        key = (key << 21) | (uint64_t)v;
    }
    return key;
}

static void sdiff_range(void *p, int64_t lo, int64_t hi, int tid)
{
    sdiff_ctx *c = (sdiff_ctx *)p;
    int64_t st = 0, cd = 0, pd = 0;
    for (int64_t i = lo; i < hi; i++) {
        int64_t n32 = 0, n64 = 0;
        if (c->dim == 1) {
            const vs_t *a = c->a1;
            const float x = a[i].v;
            int64_t b[4];
            for (int r = 0; r < 2; r++) {
                /* first j in [0,i] passing, last j in [i,W) passing */
                int64_t L = 0, H = i;
                while (L < H) {
                    int64_t mid = L + (H - L) / 2;
                    int ok = r == 0 ? pred_contract(&x, &a[mid].v, 1, c->r2)
                                    : pred_reference(&x, &a[mid].v, 1, c->R, c->R2);
                    if (ok)
                        H = mid;
                    else
                        L = mid + 1;
                }
                b[2 * r] = L;
                L = i;
                H = c->W - 1;
                while (L < H) {
                    int64_t mid = L + (H - L + 1) / 2;
                    int ok = r == 0 ? pred_contract(&x, &a[mid].v, 1, c->r2)
                                    : pred_reference(&x, &a[mid].v, 1, c->R, c->R2);
                    if (ok)
                        L = mid;
                    else
                        H = mid - 1;
                }
                b[2 * r + 1] = L;
            }
            n32 = b[1] - b[0] + 1;
            n64 = b[3] - b[2] + 1;
            pd += llabs(b[0] - b[2]) + llabs(b[1] - b[3]);
        } else {
            const float *p0 = c->win + (int64_t)c->ks[i].slot * c->dim;
            int64_t k[3] = {0, 0, 0};
            for (int a = 0; a < c->dim; a++)
                k[a] = (int64_t)floor((double)p0[a] * c->inv_c);
            int64_t nb = 1;
            for (int a = 0; a < c->dim; a++)
                nb *= 3;
            for (int64_t q = 0; q < nb; q++) {
                int64_t kk[3] = {0, 0, 0}, qq = q;
                for (int a = 0; a < c->dim; a++) {
                    kk[a] = k[a] + (qq % 3) - 1;
                    qq /= 3;
                }
                const uint64_t key = sdiff_key(kk, c->dim);
                for (int64_t j = ks_lower(c->ks, c->W, key); j < c->W && c->ks[j].key == key;
                     j++) {
                    const float *q0 = c->win + (int64_t)c->ks[j].slot * c->dim;
                    const int a32 = pred_contract(p0, q0, c->dim, c->r2);
                    const int a64 = pred_reference(p0, q0, c->dim, c->R, c->R2);
                    n32 += a32;
                    n64 += a64;
                    pd += a32 != a64;
                }
            }
        }
        cd += n32 != n64;
        st += (n32 <= c->K) != (n64 <= c->K);
    }
    c->acc[tid][0] += st;
    c->acc[tid][1] += cd;
    c->acc[tid][2] += pd;
}

int64_t rtod_o_semantics_diff(const float *win, int64_t W, int dim, float R, int K, int nthreads,
                              int64_t *out3)
{
    if (!win || W <= 0 || dim < 1 || dim > 3 || K < 1 || !(R > 0.0f))
        return -1;
    if (nthreads < 1)
        nthreads = 1;
    if (nthreads > 256)
        nthreads = 256;
    sdiff_ctx *c = (sdiff_ctx *)calloc(1, sizeof(sdiff_ctx));
    c->win = win;
    c->dim = dim;
    c->K = K;
    c->W = W;
    c->r2 = R * R;
    c->R = (double)R;
    c->R2 = (double)R * (double)R;
    c->inv_c = 1.0 / ((double)R * (1.0 + 1e-4));
    vs_t *a1 = NULL;
    ks_t *ks = NULL;
    if (dim == 1) {
        a1 = (vs_t *)malloc(sizeof(vs_t) * (size_t)W);
        for (int64_t i = 0; i < W; i++) {
            a1[i].v = win[i];
            a1[i].slot = (int32_t)i;
        }
        qsort(a1, (size_t)W, sizeof(vs_t), cmp_vs);
        c->a1 = a1;
    } else {
        ks = (ks_t *)malloc(sizeof(ks_t) * (size_t)W);
        for (int64_t i = 0; i < W; i++) {
            int64_t k[3] = {0, 0, 0};
            for (int a = 0; a < dim; a++)
                k[a] = (int64_t)floor((double)win[i * dim + a] * c->inv_c);
            ks[i].key = sdiff_key(k, dim);
            ks[i].slot = (int32_t)i;
        }
        qsort(ks, (size_t)W, sizeof(ks_t), cmp_ks);
        c->ks = ks;
    }
    parallel_for(nthreads, W, 256, sdiff_range, c);
    int64_t r[3] = {0, 0, 0};
    for (int t = 0; t < 256; t++)
        for (int q = 0; q < 3; q++)
            r[q] += c->acc[t][q];
    if (out3)
        for (int q = 0; q < 3; q++)
            out3[q] = r[q];
    free(a1);
    free(ks);
    free(c);
    return r[0];
}

int64_t rtod_o_semantics_diff_stream(const float *xs, int dim, int64_t W, int64_t S, int64_t T,
                                     int64_t every, float R, int K, int nthreads,
                                     int64_t *per_slide, int64_t *out3)
{
    if (dim < 1 || dim > 3 || W <= 0 || S <= 0 || T < 0 || every < 1)
        return -1;
    float *win = (float *)malloc(sizeof(float) * (size_t)(W * dim));
    int64_t tot[3] = {0, 0, 0};
    for (int64_t t = 1; t <= T; t++) {
        if (per_slide)
            per_slide[t - 1] = -1;
        if ((t - 1) % every != 0)
            continue;
        for (int64_t gi = t * S; gi < t * S + W; gi++)
            memcpy(win + (gi % W) * dim, xs + gi * dim, sizeof(float) * (size_t)dim);
        int64_t o[3];
        if (rtod_o_semantics_diff(win, W, dim, R, K, nthreads, o) < 0) {
            free(win);
            return -1;
        }
        if (per_slide)
            per_slide[t - 1] = o[0];
        for (int q = 0; q < 3; q++)
            tot[q] += o[q];
    }
    free(win);
    if (out3)
        for (int q = 0; q < 3; q++)
            out3[q] = tot[q];
    return tot[0];
}

/* ------------------------------------------------------------------ */
/* MLP reference (learned scheduler's predictor)                        */
/* ------------------------------------------------------------------ */
#define NIN 6
#define NH 32
#define OFF_W1 0
#define OFF_B1 (NIN * NH)
#define OFF_W2 (OFF_B1 + NH)
#define OFF_B2 (OFF_W2 + NH * NH)
#define OFF_W3 (OFF_B2 + NH)
#define OFF_B3 (OFF_W3 + NH)

float rtod_o_bf16_round(float x)
{
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) { /* inf / nan: truncate, keep nan quiet */
        if (u & 0x007fffffu)
            u |= 0x00400000u;
        u &= 0xffff0000u;
    } else {
        const uint32_t lsb = (u >> 16) & 1u;
        u += 0x7fffu + lsb;
        u &= 0xffff0000u;
    }
    float r;
    memcpy(&r, &u, 4);
    return r;
}

static inline float q(float x, int bf16) { return bf16 ? rtod_o_bf16_round(x) : x; }

/* one sample forward; keeps pre-activations for the backward pass */
static float mlp_fwd1(const float *P, const float *x, int bf16, float *a1, float *h1, float *a2,
                      float *h2)
{
    for (int j = 0; j < NH; j++) {
        float s = 0.f;
        for (int i = 0; i < NIN; i++)
            s += q(x[i], bf16) * q(P[OFF_W1 + i * NH + j], bf16);
        a1[j] = s + P[OFF_B1 + j];
        h1[j] = a1[j] > 0.f ? a1[j] : 0.f;
    }
    for (int j = 0; j < NH; j++) {
        float s = 0.f;
        for (int i = 0; i < NH; i++)
            s += q(h1[i], bf16) * q(P[OFF_W2 + i * NH + j], bf16);
        a2[j] = s + P[OFF_B2 + j];
        h2[j] = a2[j] > 0.f ? a2[j] : 0.f;
    }
    float z = 0.f; /* output layer: FP32 dot product on both paths */
    for (int i = 0; i < NH; i++)
        z += h2[i] * P[OFF_W3 + i];
    z += P[OFF_B3];
    return 1.f / (1.f + expf(-z));
}

void rtod_o_mlp_forward(const float *params, const float *X, int n, int bf16, float *pred)
{
    float a1[NH], h1[NH], a2[NH], h2[NH];
    for (int s = 0; s < n; s++)
        pred[s] = mlp_fwd1(params, X + s * NIN, bf16, a1, h1, a2, h2);
}

float rtod_o_mlp_grad(const float *params, const float *X, const float *y, int n, int bf16,
                      float *grad)
{
    const float *P = params;
    memset(grad, 0, sizeof(float) * RTOD_O_MLP_NP);
    float a1[NH], h1[NH], a2[NH], h2[NH], d2[NH], d1[NH];
    double loss = 0.0;
    for (int s = 0; s < n; s++) {
        const float *x = X + s * NIN;
        const float yh = mlp_fwd1(P, x, bf16, a1, h1, a2, h2);
        const float e = yh - y[s];
        loss += (double)e * e;
        const float dz = 2.f * e * yh * (1.f - yh) / (float)n;
        /* layer 3 */
        for (int i = 0; i < NH; i++)
            grad[OFF_W3 + i] += h2[i] * dz;
        grad[OFF_B3] += dz;
        for (int i = 0; i < NH; i++)
            d2[i] = a2[i] > 0.f ? dz * P[OFF_W3 + i] : 0.f;
        /* layer 2 */
        for (int i = 0; i < NH; i++)
            for (int j = 0; j < NH; j++)
                grad[OFF_W2 + i * NH + j] += q(h1[i], bf16) * q(d2[j], bf16);
        for (int j = 0; j < NH; j++)
            grad[OFF_B2 + j] += d2[j];
        for (int i = 0; i < NH; i++) {
            float t = 0.f;
            for (int j = 0; j < NH; j++)
                t += q(d2[j], bf16) * q(P[OFF_W2 + i * NH + j], bf16);
            d1[i] = a1[i] > 0.f ? t : 0.f;
        }
        /* layer 1 */
        for (int i = 0; i < NIN; i++)
            for (int j = 0; j < NH; j++)
                grad[OFF_W1 + i * NH + j] += q(x[i], bf16) * q(d1[j], bf16);
        for (int j = 0; j < NH; j++)
            grad[OFF_B1 + j] += d1[j];
    }
    return (float)(loss / (double)n);
}

double rtod_o_mlp_loss64(const double *P, const float *X, const float *y, int n)
{
    double loss = 0.0;
    for (int s = 0; s < n; s++) {
        const float *x = X + s * NIN;
        double h1[NH], h2[NH];
        for (int j = 0; j < NH; j++) {
            double t = P[OFF_B1 + j];
            for (int i = 0; i < NIN; i++)
                t += (double)x[i] * P[OFF_W1 + i * NH + j];
            h1[j] = t > 0.0 ? t : 0.0;
        }
        for (int j = 0; j < NH; j++) {
            double t = P[OFF_B2 + j];
            for (int i = 0; i < NH; i++)
                t += h1[i] * P[OFF_W2 + i * NH + j];
            h2[j] = t > 0.0 ? t : 0.0;
        }
        double z = P[OFF_B3];
        for (int i = 0; i < NH; i++)
            z += h2[i] * P[OFF_W3 + i];
        const double yh = 1.0 / (1.0 + exp(-z));
        loss += (yh - y[s]) * (yh - y[s]);
    }
    return loss / (double)n;
}
