/*
 * expected.c -- closed-form expected join output for the seeded workloads
 * (checker only; never linked into the product).
 *
 * The generators (pythia-core_amd/csrc/pj_datagen.h) state for every probe
 * tuple which build tuple, if any, holds its key. For a primary-key /
 * foreign-key join with projection [B$0, B$1, P$1] the output multiset is
 * therefore { (key, build payload, probe payload) : probe tuple j matches },
 * one tuple per matching probe tuple (the closed form of a PK-FK join), so
 * its count and order-independent checksum follow without building any hash
 * table. This pins the device join at sizes the restated HashJoinOp cannot
 * finish in seconds.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "pj_datagen.h"

typedef struct {
    const pj_pkfk_params* p;
    uint64_t b, e, sum, cnt;
} pkfk_job;

static void* pkfk_worker(void* arg)
{
    pkfk_job* j = (pkfk_job*)arg;
    uint64_t s = 0, c = 0;
    for (uint64_t i = j->b; i < j->e; ++i) {
        uint64_t k, pp, bk, bp;
        uint64_t idx = pj_pkfk_probe(j->p, i, &k, &pp);
        if (idx == ~0ull)
            continue;
        pj_pkfk_build(j->p, idx, &bk, &bp);
        uint64_t t[3] = {k, bp, pp};
        s += or_tuple_mix(t, 24);
        c++;
    }
    j->sum = s;
    j->cnt = c;
    return NULL;
}

int or_expected_pkfk(uint64_t seed, uint64_t domain, uint64_t n_build, uint64_t n_probe, uint64_t n_miss,
                     int threads, uint64_t* sum, uint64_t* count)
{
    pj_pkfk_params p = {seed, domain, n_build, n_probe, n_miss};
    if (threads < 1)
        threads = 1;
    if (threads > 64)
        threads = 64;
    pthread_t th[64];
    pkfk_job jobs[64];
    for (int t = 0; t < threads; ++t) {
        jobs[t].p = &p;
        jobs[t].b = n_probe * (uint64_t)t / (uint64_t)threads;
        jobs[t].e = n_probe * (uint64_t)(t + 1) / (uint64_t)threads;
        pthread_create(&th[t], NULL, pkfk_worker, &jobs[t]);
    }
    uint64_t s = 0, c = 0;
    for (int t = 0; t < threads; ++t) {
        pthread_join(th[t], NULL);
        s += jobs[t].sum;
        c += jobs[t].cnt;
    }
    *sum = s;
    *count = c;
    return 0;
}

/* Join j of the batch workload: rows (key, build payload32, probe
 * payload32), 16 bytes. */
typedef struct {
    pj_pkfk_params p;
    uint64_t b, e, sum, cnt;
} batch_job;

static void* batch_worker(void* arg)
{
    batch_job* j = (batch_job*)arg;
    uint64_t s = 0, c = 0;
    for (uint64_t i = j->b; i < j->e; ++i) {
        uint64_t k, pp, bk, bp;
        uint64_t idx = pj_pkfk_probe(&j->p, i, &k, &pp);
        if (idx == ~0ull)
            continue;
        pj_pkfk_build(&j->p, idx, &bk, &bp);
        uint64_t t[2] = {k, (bp & 0xFFFFFFFFull) | (pp << 32)};
        s += or_tuple_mix(t, 16);
        c++;
    }
    j->sum = s;
    j->cnt = c;
    return NULL;
}

int or_expected_batch_join(uint64_t seed, uint32_t jidx, uint64_t n_build, uint64_t n_probe, int threads,
                           uint64_t* sum, uint64_t* count)
{
    pj_pkfk_params p = {pj_batch_seed(seed, jidx), 1ull << 40, n_build, n_probe, 0};
    if (threads < 1)
        threads = 1;
    if (threads > 64)
        threads = 64;
    pthread_t th[64];
    batch_job jobs[64];
    for (int t = 0; t < threads; ++t) {
        jobs[t].p = p;
        jobs[t].b = n_probe * (uint64_t)t / (uint64_t)threads;
        jobs[t].e = n_probe * (uint64_t)(t + 1) / (uint64_t)threads;
        pthread_create(&th[t], NULL, batch_worker, &jobs[t]);
    }
    uint64_t s = 0, c = 0;
    for (int t = 0; t < threads; ++t) {
        pthread_join(th[t], NULL);
        s += jobs[t].sum;
        c += jobs[t].cnt;
    }
    *sum = s;
    *count = c;
    return 0;
}

/* Host generation of the same relations (for the CPU-side join runs). */
int or_gen_pkfk(uint64_t seed, uint64_t domain, uint64_t n_build, uint64_t n_probe, uint64_t n_miss,
                void* build, void* probe)
{
    pj_pkfk_params p = {seed, domain, n_build, n_probe, n_miss};
    uint64_t* b = (uint64_t*)build;
    uint64_t* q = (uint64_t*)probe;
    for (uint64_t i = 0; i < n_build; ++i)
        pj_pkfk_build(&p, i, &b[2 * i], &b[2 * i + 1]);
    for (uint64_t j = 0; j < n_probe; ++j)
        pj_pkfk_probe(&p, j, &q[2 * j], &q[2 * j + 1]);
    return 0;
}

/* Probe tuples [j0, j1) of the full probe relation of n_probe tuples. */
int or_gen_pkfk_probe_range(uint64_t seed, uint64_t domain, uint64_t n_build, uint64_t n_probe,
                            uint64_t n_miss, uint64_t j0, uint64_t j1, void* probe)
{
    pj_pkfk_params p = {seed, domain, n_build, n_probe, n_miss};
    uint64_t* q = (uint64_t*)probe;
    for (uint64_t j = j0; j < j1; ++j)
        pj_pkfk_probe(&p, j, &q[2 * (j - j0)], &q[2 * (j - j0) + 1]);
    return 0;
}

typedef struct {
    const double* cdf;
    uint64_t seed, n_build, b, e, sum, cnt;
} pz_job;

static void* pz_worker(void* arg)
{
    pz_job* j = (pz_job*)arg;
    uint64_t s = 0;
    for (uint64_t i = j->b; i < j->e; ++i) {
        uint32_t a, b, ba, bb;
        uint64_t pp, bp;
        uint64_t idx = pj_pair_zipf_probe(j->cdf, j->n_build, j->seed, i, &a, &b, &pp);
        pj_pair_build(j->seed, idx, &ba, &bb, &bp);
        unsigned char t[24];
        memcpy(t, &a, 4);
        memcpy(t + 4, &b, 4);
        memcpy(t + 8, &bp, 8);
        memcpy(t + 16, &pp, 8);
        s += or_tuple_mix(t, 24);
    }
    j->sum = s;
    j->cnt = j->e - j->b;
    return NULL;
}

/* Every Zipf probe matches exactly one build tuple (unique composite keys),
 * so the output of projection [B$0, B$1, B$2, P$2] is one (a, b, bp, pp) per
 * probe tuple. */
int or_expected_pairs_zipf(uint64_t seed, uint64_t n_build, uint64_t n_probe, const double* cdf, int threads,
                           uint64_t* sum, uint64_t* count)
{
    if (threads < 1)
        threads = 1;
    if (threads > 64)
        threads = 64;
    pthread_t th[64];
    pz_job jobs[64];
    for (int t = 0; t < threads; ++t) {
        jobs[t].cdf = cdf;
        jobs[t].seed = seed;
        jobs[t].n_build = n_build;
        jobs[t].b = n_probe * (uint64_t)t / (uint64_t)threads;
        jobs[t].e = n_probe * (uint64_t)(t + 1) / (uint64_t)threads;
        pthread_create(&th[t], NULL, pz_worker, &jobs[t]);
    }
    uint64_t s = 0, c = 0;
    for (int t = 0; t < threads; ++t) {
        pthread_join(th[t], NULL);
        s += jobs[t].sum;
        c += jobs[t].cnt;
    }
    *sum = s;
    *count = c;
    return 0;
}

/* Host generation of the same relations, (a, b, payload) packed in 16 B. */
int or_gen_pairs_zipf(uint64_t seed, uint64_t n_build, uint64_t n_probe, const double* cdf, void* build,
                      void* probe)
{
    unsigned char* B = (unsigned char*)build;
    unsigned char* P = (unsigned char*)probe;
    for (uint64_t i = 0; i < n_build; ++i) {
        uint32_t a, b;
        uint64_t pay;
        pj_pair_build(seed, i, &a, &b, &pay);
        memcpy(B + 16 * i, &a, 4);
        memcpy(B + 16 * i + 4, &b, 4);
        memcpy(B + 16 * i + 8, &pay, 8);
    }
    for (uint64_t j = 0; j < n_probe; ++j) {
        uint32_t a, b;
        uint64_t pay;
        pj_pair_zipf_probe(cdf, n_build, seed, j, &a, &b, &pay);
        memcpy(P + 16 * j, &a, 4);
        memcpy(P + 16 * j + 4, &b, 4);
        memcpy(P + 16 * j + 8, &pay, 8);
    }
    return 0;
}

/* ---- clustered keys with Zipf foreign keys ------------------------------ */

/* guide[g] = first rank r with cdf[r] > g / G, so the rank of u lies
 * between guide[g - 1] and guide[g + 2] for g = floor(u * G). */
static uint32_t* make_guide(const double* cdf, uint64_t n, uint64_t G)
{
    uint32_t* guide = (uint32_t*)malloc((G + 1) * sizeof(uint32_t));
    uint64_t r = 0;
    for (uint64_t g = 0; g <= G; ++g) {
        const double t = (double)g / (double)G;
        while (r < n - 1 && !(cdf[r] > t))
            r++;
        guide[g] = (uint32_t)r;
    }
    return guide;
}

/* The same rank as pj_zipf_rank (same uniforms, same redraws, first CDF
 * entry > u), searched inside a guide-table bracket. */
static uint64_t zipf_rank_guided(const double* cdf, uint64_t n, const uint32_t* guide, uint64_t G, uint64_t seed,
                                 uint64_t j)
{
    const double last = cdf[n - 1];
    for (uint64_t attempt = 0;; ++attempt) {
        uint64_t r = pj_rand64(seed, 5 + attempt, j) >> 2;
        double u = (double)r / 4611686018427387904.0;
        if (u >= last)
            continue;
        uint64_t g = (uint64_t)(u * (double)G);
        if (g > G)
            g = G;
        uint64_t lo = g >= 1 ? guide[g - 1] : 0, hi = g + 2 <= G ? guide[g + 2] : n - 1;
        if (hi > n - 1)
            hi = n - 1;
        while (lo < hi) {
            uint64_t mid = lo + (hi - lo) / 2;
            if (cdf[mid] > u)
                hi = mid;
            else
                lo = mid + 1;
        }
        return lo;
    }
}

typedef struct {
    const pj_cluster_table* t;
    const double* cdf;
    const uint32_t* guide;
    uint64_t G, b, e, sum, cnt;
} cl_job;

static void* cl_worker(void* arg)
{
    cl_job* jb = (cl_job*)arg;
    uint64_t s = 0;
    for (uint64_t j = jb->b; j < jb->e; ++j) {
        uint64_t key, pp, bk, bp;
        uint64_t rank = zipf_rank_guided(jb->cdf, jb->t->n_keys, jb->guide, jb->G, jb->t->seed, j);
        uint64_t idx = pj_clustered_probe_with_rank(jb->t, rank, j, &key, &pp);
        pj_clustered_build(jb->t, idx, &bk, &bp);
        uint64_t tup[3] = {key, bp, pp};
        s += or_tuple_mix(tup, 24);
    }
    jb->sum = s;
    jb->cnt = jb->e - jb->b;
    return NULL;
}

/* Expected output of projection [B$0, B$1, P$1] over the clustered workload:
 * one (key, bp, pp) per probe tuple (every probe key is a build key). */
int or_expected_clustered(const void* table, uint64_t n_probe, const double* cdf, int threads, uint64_t* sum,
                          uint64_t* count)
{
    const pj_cluster_table* t = (const pj_cluster_table*)table;
    const uint64_t G = 1ull << 24;
    uint32_t* guide = make_guide(cdf, t->n_keys, G);
    if (threads < 1)
        threads = 1;
    if (threads > 64)
        threads = 64;
    pthread_t th[64];
    cl_job jobs[64];
    for (int k = 0; k < threads; ++k) {
        jobs[k].t = t;
        jobs[k].cdf = cdf;
        jobs[k].guide = guide;
        jobs[k].G = G;
        jobs[k].b = n_probe * (uint64_t)k / (uint64_t)threads;
        jobs[k].e = n_probe * (uint64_t)(k + 1) / (uint64_t)threads;
        pthread_create(&th[k], NULL, cl_worker, &jobs[k]);
    }
    uint64_t s = 0, c = 0;
    for (int k = 0; k < threads; ++k) {
        pthread_join(th[k], NULL);
        s += jobs[k].sum;
        c += jobs[k].cnt;
    }
    free(guide);
    *sum = s;
    *count = c;
    return 0;
}

/* Probe ranks through the guide table, for checking it against plain
 * binary search. */
int or_zipf_ranks_guided(const double* cdf, uint64_t n, uint64_t seed, uint64_t m, uint64_t* ranks)
{
    const uint64_t G = 1ull << 16;
    uint32_t* guide = make_guide(cdf, n, G);
    for (uint64_t j = 0; j < m; ++j)
        ranks[j] = zipf_rank_guided(cdf, n, guide, G, seed, j);
    free(guide);
    return 0;
}
