// pj_datagen.h -- counter-based, seeded workload generators shared by host
// and device code. Tuple i of a relation is a pure function of (seed, i), so
// a relation generated on the GPU is bit-identical to one generated on the
// host, in any order and on any number of threads.
//
// The reference's generators (operators/generator_int.cpp:101-124,
// operators/generator_zipf.cpp:44-156) are time-seeded and not reproducible
// (util/reentrant_random.inl:58); these keep their semantics (uniform keys,
// Zipf over ranks with a rank -> key permutation, a double-precision CDF
// searched by binary search with 62-bit uniforms) with explicit seeds.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define PJ_HD __host__ __device__ inline
#else
#define PJ_HD static inline
#endif

// splitmix64 finaliser (Steele, Lea, Flood 2014; Vigna's constants).
PJ_HD uint64_t pj_fin64(uint64_t x)
{
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

// Independent 64-bit uniform stream `stream` of generator `seed`.
PJ_HD uint64_t pj_rand64(uint64_t seed, uint64_t stream, uint64_t i)
{
    return pj_fin64(pj_fin64(seed ^ (stream * 0xd1b54a32d192ed03ull)) + i * 0x9e3779b97f4a7c15ull);
}

PJ_HD unsigned pj_ceil_log2(uint64_t n)
{
    unsigned k = 0;
    while (k < 64 && (1ull << k) < n)
        ++k;
    return k;
}

// A keyed bijection on [0, 2^k): four rounds of odd multiply, xorshift and
// add, each a bijection mod 2^k.
PJ_HD uint64_t pj_bij(uint64_t x, unsigned k, uint64_t key)
{
    const uint64_t mask = k >= 64 ? ~0ull : ((1ull << k) - 1);
    const unsigned sh = (k + 1) / 2;
    const uint64_t mul[4] = {0x9e3779b97f4a7c15ull, 0xbf58476d1ce4e5b9ull,
                             0x94d049bb133111ebull, 0xd6e8feb86659fd93ull};
    for (int r = 0; r < 4; ++r) {
        x = (x * mul[r]) & mask;
        if (sh)
            x ^= x >> sh;
        x = (x + pj_fin64(key + (uint64_t)r)) & mask;
    }
    return x;
}

// A keyed permutation of [0, n) by cycle walking over pj_bij.
PJ_HD uint64_t pj_perm(uint64_t i, uint64_t n, uint64_t key)
{
    const unsigned k = pj_ceil_log2(n);
    uint64_t x = i;
    do {
        x = pj_bij(x, k, key);
    } while (x >= n);
    return x;
}

// ---------------------------------------------------------------------------
// Uniform primary-key / foreign-key workload (queryhashjoin shape, and the
// 16M x 256M uniform join).
//   build tuple i : key = perm_domain(i), payload = rand(seed, 1, i)
//   probe tuple j : a miss iff perm_{n_probe}(j) < n_miss, with key
//                   perm_domain(n_build + that rank) (never a build key);
//                   otherwise key = build key of index rand(seed, 2, j) mod
//                   n_build; payload = rand(seed, 3, j).
// domain must be >= n_build + n_miss; domain == n_build gives the dense keys
// 0..n_build-1 of the reference's unit tests.
typedef struct {
    uint64_t seed;
    uint64_t domain;
    uint64_t n_build, n_probe, n_miss;
} pj_pkfk_params;

PJ_HD uint64_t pj_pkfk_domain_key(const pj_pkfk_params* p, uint64_t idx)
{
    return pj_perm(idx, p->domain, p->seed ^ 0x5eedull);
}

PJ_HD void pj_pkfk_build(const pj_pkfk_params* p, uint64_t i, uint64_t* key, uint64_t* pay)
{
    *key = pj_pkfk_domain_key(p, i);
    *pay = pj_rand64(p->seed, 1, i);
}

// Returns the build index matched by probe tuple j, or ~0 for a miss.
PJ_HD uint64_t pj_pkfk_probe(const pj_pkfk_params* p, uint64_t j, uint64_t* key, uint64_t* pay)
{
    uint64_t idx = ~0ull;
    uint64_t m = p->n_miss ? pj_perm(j, p->n_probe, p->seed ^ 0x4d155ull) : p->n_probe;
    if (m < p->n_miss) {
        *key = pj_pkfk_domain_key(p, p->n_build + m);
    } else {
        idx = pj_rand64(p->seed, 2, j) % p->n_build;
        *key = pj_pkfk_domain_key(p, idx);
    }
    *pay = pj_rand64(p->seed, 3, j);
    return idx;
}

// Batch workload: join j of a batch is the PK-FK workload above with seed
// pj_batch_seed(seed, j), domain 2^40, no misses, and 32-bit payloads (the
// low words of the 64-bit ones).
PJ_HD uint64_t pj_batch_seed(uint64_t seed, uint32_t j)
{
    return pj_fin64(seed ^ (0xba7c4ull + (uint64_t)j * 0x9e3779b97f4a7c15ull));
}

// ---------------------------------------------------------------------------
// Zipf foreign keys over build ranks (generator_zipf semantics,
// operators/generator_zipf.cpp:73-152): u is a 62-bit uniform / 2^62,
// values >= the last CDF entry are redrawn, and rank = the first CDF entry
// > u. cdf[r] = sum_{i<=r} (i+1)^-z / H. Rank r maps to build index
// perm_{n_build}(r) (the rank -> key shuffle of generator_zipf.cpp:85-89).
PJ_HD uint64_t pj_zipf_rank(const double* cdf, uint64_t n, uint64_t seed, uint64_t j)
{
    const double last = cdf[n - 1];
    for (uint64_t attempt = 0;; ++attempt) {
        uint64_t r = pj_rand64(seed, 5 + attempt, j) >> 2;
        double u = (double)r / 4611686018427387904.0;  // 2^62
        if (u >= last)
            continue;
        uint64_t lo = 0, hi = n - 1;                   // first cdf[i] > u
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

PJ_HD uint64_t pj_zipf_build_index(uint64_t rank, uint64_t n_build, uint64_t seed)
{
    return pj_perm(rank, n_build, seed ^ 0x21bfull);
}

// Composite-key build relation (queryhashjoincomposite shape): unique (a, b)
// pairs uniform in [0, 2^31)^2 from a bijection on 62 bits.
PJ_HD void pj_pair_build(uint64_t seed, uint64_t i, uint32_t* a, uint32_t* b, uint64_t* pay)
{
    uint64_t v = pj_bij(i, 62, seed ^ 0xab5eedull);
    *a = (uint32_t)(v & 0x7fffffffu);
    *b = (uint32_t)(v >> 31);
    *pay = pj_rand64(seed, 1, i);
}

// Zipf probe relation over the composite-key build relation: probe tuple j
// is (a, b, payload) of build tuple pj_zipf_build_index(rank) with rank drawn
// from the CDF; its payload is rand(seed, 3, j). Returns the build index.
PJ_HD uint64_t pj_pair_zipf_probe(const double* cdf, uint64_t n_build, uint64_t seed, uint64_t j, uint32_t* a,
                                  uint32_t* b, uint64_t* pay)
{
    const uint64_t rank = pj_zipf_rank(cdf, n_build, seed, j);
    const uint64_t idx = pj_zipf_build_index(rank, n_build, seed);
    uint64_t bp;
    pj_pair_build(seed, idx, a, b, &bp);
    *pay = pj_rand64(seed, 3, j);
    return idx;
}

// ---------------------------------------------------------------------------
// Clustered build keys: a table of dense clusters [centre, centre + width),
// sorted by centre and disjoint; cluster c holds count_c distinct keys
// centre + perm_width(j), j < count_c. Global key index g lies in the cluster
// whose [start_c, start_{c+1}) contains it. Build tuple i has global index
// perm_n(i); probe tuples draw a build index through the Zipf CDF.
#define PJ_MAX_CLUSTERS 1024
typedef struct {
    uint32_t nclusters;
    uint64_t seed;
    uint64_t n_keys;
    uint64_t centre[PJ_MAX_CLUSTERS];
    uint64_t width[PJ_MAX_CLUSTERS];
    uint64_t start[PJ_MAX_CLUSTERS + 1];
} pj_cluster_table;

PJ_HD uint64_t pj_cluster_key(const pj_cluster_table* t, uint64_t g)
{
    uint32_t lo = 0, hi = t->nclusters;   // last c with start[c] <= g
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (t->start[mid] <= g)
            lo = mid;
        else
            hi = mid;
    }
    const uint64_t j = g - t->start[lo];
    return t->centre[lo] + pj_perm(j, t->width[lo], t->seed ^ (0xc1u + (uint64_t)lo));
}

PJ_HD void pj_clustered_build(const pj_cluster_table* t, uint64_t i, uint64_t* key, uint64_t* pay)
{
    *key = pj_cluster_key(t, pj_perm(i, t->n_keys, t->seed ^ 0x9e11ull));
    *pay = pj_rand64(t->seed, 1, i);
}

// Build index of the rank-th most frequent probe key (a seeded shuffle).
PJ_HD uint64_t pj_clustered_probe_with_rank(const pj_cluster_table* t, uint64_t rank, uint64_t j, uint64_t* key,
                                            uint64_t* pay)
{
    const uint64_t idx = pj_zipf_build_index(rank, t->n_keys, t->seed);
    uint64_t bp;
    pj_clustered_build(t, idx, key, &bp);
    *pay = pj_rand64(t->seed, 3, j);
    return idx;
}
