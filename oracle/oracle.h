/*
 * oracle.h -- CPU restatement of Pythia's hash-join path, used ONLY as a
 * checker.
 *
 * This is test infrastructure. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it. The product (libpjoin.so and the
 * C++ operators) never links, loads or calls anything here.
 *
 * What is restated, and from where (paths relative to the reference tree):
 *   - getlogarithm, HashFunction(buckets) ............ hash.cpp:49-69
 *   - the TupleHasher value/byte hash functions ....... hash.h:104-487
 *   - TupleHasher::create (fn selection, offset/size) . hash.cpp:101-248
 *   - ParameterizedModuloValueHasher::generate ........ hash.cpp:251-281
 *   - findIndexOfMaxContiguousAttribute ............... operators/join.cpp:340-359
 *   - Schema packing (no alignment), writeData ........ schema.cpp:82-117, schema.inl:47-93
 *   - HashTable (bucket headers, overflow chains,
 *     iterator) ...................................... util/hashtable.h:44-299, util/hashtable.cpp:57-297
 *   - HashJoinOp init/build/probe/output .............. operators/join.cpp:224-326, 367-518,
 *                                                       648-800, 838-884
 *   - key equality (Int/Long cross-width) ............. rawcompfns.cpp:48-180
 *   - PartitionOp histogram/prefix/scatter ............ operators/partition.cpp:44-404
 *   - ConsumeOp XOR of 32-bit words ................... operators/consume.cpp:68-100
 *
 * How each part is pinned is listed in tests/test_oracle.py (known-answer
 * tests from the reference's unit_tests, closed forms, brute force, and the
 * reference itself compiled into oracle/_ref where that build is present).
 */
#ifndef PYTHIA_ORACLE_H
#define PYTHIA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Column types, numbered as the reference's ColumnType (schema.h:53-60). */
enum { OR_CT_INTEGER = 0, OR_CT_LONG = 1, OR_CT_DECIMAL = 2, OR_CT_CHAR = 3,
       OR_CT_DATE = 4, OR_CT_POINTER = 5 };

#define OR_MAX_COLS 32

typedef struct {
    int ncols;
    int type[OR_MAX_COLS];
    int width[OR_MAX_COLS];
    int offset[OR_MAX_COLS];
    int size;
} or_schema;

/* Packs columns with no alignment (schema.cpp:82-117). widths[] is read only
 * for OR_CT_CHAR. Returns 0, or -1 on a bad declaration. */
int or_schema_init(or_schema* s, int ncols, const int* types, const int* widths);

/* ------------------------------------------------------------------ hashes */
enum { OR_FN_MODULO = 1, OR_FN_PARAMMODULO, OR_FN_KNUTH, OR_FN_RANGE,
       OR_FN_EXACTRANGE, OR_FN_BYTES, OR_FN_TPCHORDERKEY, OR_FN_WILLIS,
       OR_FN_ALWAYSZERO, OR_FN_TPCHQ1MAGIC };

typedef struct {
    int fn;
    int64_t k;          /* the reference's _k (log2 buckets, or a mask) */
    int64_t min, max;   /* range / exactrange / parammodulo offset */
    int64_t bucketrange;
    unsigned skipbits;
    unsigned offset;    /* byte offset of the hashed field(s) in the tuple */
    unsigned size;      /* bytes hashed */
} or_hasher;

unsigned or_getlogarithm(unsigned k);
int or_fn_from_name(const char* name);
/* Builds the function the way TupleHasher::create does for the given fn,
 * bucket count and optional range/offset/skipbits. Returns 0 or -1
 * (UnknownHash / zero buckets). */
int or_hasher_init(or_hasher* h, int fn, unsigned buckets, int64_t range_min,
                   int64_t range_max, unsigned offset_param, unsigned skipbits);
unsigned or_hasher_buckets(const or_hasher* h);
unsigned or_hasher_hash(const or_hasher* h, const void* start, size_t size);
/* Value entry (ValueHasher::hash(CtLong)) for value-based functions. */
unsigned or_hasher_hash_value(const or_hasher* h, int64_t v);
int64_t or_exactrange_minimumforbucket(const or_hasher* h, unsigned bucket);
/* generate(passes) of a parammodulo function: writes per-pass bucket counts
 * and skipbits. Returns the number of functions. */
unsigned or_parammodulo_generate(int64_t offset, unsigned buckets, unsigned skipbits,
                                 unsigned passes, unsigned* out_buckets,
                                 unsigned* out_skipbits);
unsigned or_find_max_contiguous(const uint16_t* v1, const uint16_t* v2, unsigned n);

/* --------------------------------------------------------------- hash join */
enum { OR_BUILD = 0, OR_PROBE = 1 };

typedef struct {
    or_schema build, probe;
    int nkeys;
    uint16_t buildjattr[8], probejattr[8];
    int nproj;
    int proj_side[OR_MAX_COLS];
    int proj_attr[OR_MAX_COLS];
    int hash_fn;
    unsigned buckets;
    int64_t range_min, range_max;
    unsigned hash_offset, skipbits;
    unsigned tuplesperbucket;
    unsigned buffsize;      /* output page size in bytes (Operator::buffsize) */
} or_join_desc;

typedef struct {
    uint64_t out_tuples;
    uint64_t pages;         /* getNext calls that returned tuples or Finished */
    uint64_t spills;        /* overflow buckets allocated (HashTable::spills) */
    unsigned out_tuple_size;
} or_join_stats;

/* Output tuple width for the descriptor's projection. */
int or_join_out_width(const or_join_desc* d);

/* Runs the restated HashJoinOp single-threaded. Output pages of
 * d->buffsize bytes are produced and concatenated into out. Returns 0, -1 on
 * a bad descriptor, -2 when out_cap tuples are not enough. */
int or_hashjoin(const or_join_desc* d, const void* build, uint64_t nbuild,
                const void* probe, uint64_t nprobe, void* out, uint64_t out_cap,
                or_join_stats* st);

/* Independent brute-force nested-loop join with the same key semantics and
 * projection (for small inputs). */
int or_nested_loop_join(const or_join_desc* d, const void* build, uint64_t nbuild,
                        const void* probe, uint64_t nprobe, void* out,
                        uint64_t out_cap, uint64_t* out_n);

/* -------------------------------------------------------------- partition */
/* Restated PartitionOp over T thread inputs. in holds the T inputs back to
 * back (counts[t] tuples each). Buckets == T (partition.cpp:63-64).
 * out receives partition 0, then 1, ... ; part_sizes[T] gets the sizes.
 * Returns 0, or -1 on a bad argument. */
int or_partition(const or_schema* s, int attr, int64_t range_min, int64_t range_max,
                 unsigned threads, const void* in, const uint64_t* counts,
                 void* out, uint64_t* part_sizes);

/* --------------------------------------------------------------- checksums */
/* Order-independent 64-bit checksum: sum mod 2^64 of or_tuple_mix over the
 * tuples. or_tuple_mix folds the tuple's bytes, 8 at a time little-endian
 * (last chunk zero-padded), through the splitmix64 finaliser. */
uint64_t or_tuple_mix(const void* t, unsigned width);
uint64_t or_checksum(const void* tuples, uint64_t n, unsigned width);
/* ConsumeOp-style XOR of all 32-bit words (consume.cpp:68-100). */
uint32_t or_xor32(const void* tuples, uint64_t n, unsigned width);

/* ------------------------------------------------- seeded-workload checks */
/* Expected count and checksum of the PK-FK join of the seeded uniform
 * workload with projection [B$0, B$1, P$1] (see expected.c). */
int or_expected_pkfk(uint64_t seed, uint64_t domain, uint64_t n_build, uint64_t n_probe,
                     uint64_t n_miss, int threads, uint64_t* sum, uint64_t* count);
/* The same relations generated on the host, (key, payload) int64 pairs. */
int or_gen_pkfk(uint64_t seed, uint64_t domain, uint64_t n_build, uint64_t n_probe,
                uint64_t n_miss, void* build, void* probe);
/* Probe tuples [j0, j1) of the full probe relation of n_probe tuples. */
int or_gen_pkfk_probe_range(uint64_t seed, uint64_t domain, uint64_t n_build, uint64_t n_probe,
                            uint64_t n_miss, uint64_t j0, uint64_t j1, void* probe);
/* Join jidx of the batch workload (seeded as pj_gen_batch_join): expected
 * row count and checksum of rows (key, build payload32, probe payload32). */
int or_expected_batch_join(uint64_t seed, uint32_t jidx, uint64_t n_build, uint64_t n_probe, int threads,
                           uint64_t* sum, uint64_t* count);

/* The composite-key Zipf workload: expected output count and checksum of
 * projection [B$0, B$1, B$2, P$2], and host generation of the relations. */
int or_expected_pairs_zipf(uint64_t seed, uint64_t n_build, uint64_t n_probe, const double* cdf,
                           int threads, uint64_t* sum, uint64_t* count);
int or_gen_pairs_zipf(uint64_t seed, uint64_t n_build, uint64_t n_probe, const double* cdf,
                      void* build, void* probe);
/* The clustered-key workload (table from pj_clusters_make): expected output
 * of projection [B$0, B$1, P$1]; Zipf ranks through a guide table. */
int or_expected_clustered(const void* table, uint64_t n_probe, const double* cdf, int threads,
                          uint64_t* sum, uint64_t* count);
int or_zipf_ranks_guided(const double* cdf, uint64_t n, uint64_t seed, uint64_t m, uint64_t* ranks);

#ifdef __cplusplus
}
#endif
#endif
