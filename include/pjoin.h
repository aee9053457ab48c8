/*
 * pjoin.h -- C ABI of the MI355X hash-join path (libpjoin.so).
 *
 * The library replaces the compute behind Pythia's HashJoinOp (with the
 * HashTable and TupleHasher it uses) and PartitionOp. The caller keeps the
 * Operator protocol (init / threadInit / scanStart / getNext / scanStop /
 * threadClose); the C++ operators in pythia-core_amd/host call these
 * entry points at the matching points of that protocol. Reference citations
 * are relative to the reference tree.
 *
 * Conventions: every function returns PJ_OK (0) or a negative PJ_ERR_*;
 * the calling thread's pj_last_error() then describes the failure. No C++
 * exception crosses this boundary. The caller owns every buffer it passes;
 * the library owns device scratch memory and the handles it returns.
 * Tuples are packed with no alignment, as the reference's Schema packs them
 * (schema.cpp:82-117). Output order is unspecified (the reference's own
 * order depends on thread interleaving); the output multiset is exact.
 */
#ifndef PJOIN_H
#define PJOIN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PJ_OK 0
#define PJ_ERR_INVALID (-1)     /* bad descriptor or argument (InvalidParameter,
                                   MissingParameterException, UnknownHashException) */
#define PJ_ERR_UNSUPPORTED (-2) /* a schema the device path cannot join exactly */
#define PJ_ERR_DEVICE (-3)      /* a HIP runtime failure, e.g. hipMalloc */
#define PJ_ERR_CAPACITY (-4)    /* output buffer too small; *out_n = rows needed */
#define PJ_ERR_STATE (-5)       /* call out of protocol order */

/* Column types, numbered as the reference's ColumnType (schema.h:53-60). */
enum { PJ_CT_INTEGER = 0, PJ_CT_LONG = 1, PJ_CT_DECIMAL = 2, PJ_CT_CHAR = 3,
       PJ_CT_DATE = 4, PJ_CT_POINTER = 5 };

#define PJ_MAX_COLS 32
#define PJ_MAX_KEYS 8

typedef struct {
    uint32_t ncols;
    uint32_t type[PJ_MAX_COLS];
    uint32_t width[PJ_MAX_COLS];  /* read for PJ_CT_CHAR only */
} pj_schema;

enum { PJ_SIDE_BUILD = 0, PJ_SIDE_PROBE = 1 };
enum { PJ_PART_RADIX = 0, PJ_PART_LEARNED = 1, PJ_PART_AUTO = 2 };

/* The plan-node parameters of a `hashjoin` that shape its result
 * (operators/operators.h:2965-3070; parsed by operators/join.cpp:224-326 and
 * 367-518). The hash function and bucket count of the plan are validated by
 * the C++ operator but do not change the output multiset; the device path
 * partitions on its own hash (or the learned model's ranges). */
typedef struct {
    pj_schema build;
    pj_schema probe;
    uint32_t nkeys;                     /* buildjattr / probejattr length */
    uint16_t buildjattr[PJ_MAX_KEYS];
    uint16_t probejattr[PJ_MAX_KEYS];
    uint32_t nproj;                     /* projection: "B$i" / "P$j" entries */
    uint8_t proj_side[PJ_MAX_COLS];
    uint16_t proj_attr[PJ_MAX_COLS];
    uint32_t partitioner;               /* PJ_PART_* */
    int32_t radix_bits;                 /* -1: chosen from |R| */
    uint32_t flags;                     /* PJ_FLAG_* */
} pj_join_desc;

#define PJ_FLAG_TIMING 1u               /* record per-phase HIP event times */
#define PJ_FLAG_NARROW_ROW_HASH 2u      /* schemas joined in the row-id form (keys other than one INTEGER /
                                           LONG / DATE column or two INTEGER columns, wider projections):
                                           keep 10 bits of the key hash, so that unequal keys share hashes
                                           and the key check on the tuples decides (a test of that check) */
#define PJ_FLAG_LAYOUT_EXACT 4u         /* partition the probe side after a histogram pass, into contiguous
                                           partitions, whatever its size */
#define PJ_FLAG_LAYOUT_APPEND 8u        /* partition the probe side without a histogram pass (each block
                                           appends whole chunks, then the chunks are listed by partition),
                                           whatever its size, where the join allows it (radix partitions
                                           read once, fewer than 2^30 probe tuples); by default probe
                                           sides of 2^22 tuples or more take it where allowed. It needs
                                           about twice the probe side's bytes of device memory for the
                                           partitioned copy */
#define PJ_FLAG_LAYOUT_TWO_PASS 16u     /* where the build side has more partitions than one write-combining
                                           pass gives the probe side (2 to 8 build partitions a probe
                                           partition), partition the probe side in two radix passes, high bits
                                           first, so that the join reads each probe record once instead of
                                           once a build partition, whatever its size; by default probe sides
                                           of 2^22 tuples or more with 4 to 8 build partitions a probe
                                           partition take it. It needs another copy of the probe side in
                                           device memory */

typedef struct pj_join_s* pj_join_t;

/* Per-phase device times of the last run, in milliseconds (PJ_FLAG_TIMING). */
typedef struct {
    float build_partition_ms;  /* histogram + scan + scatter of R */
    float probe_partition_ms;  /* histogram + scan + scatter of S */
    float join_ms;             /* LDS build + probe + output (join kernel) */
    float model_ms;            /* learned partitioner: sample, fit, boundaries */
    float total_ms;            /* first launch to last, including host planning */
    uint32_t partitions;       /* partitions of S */
    uint32_t subpartitions;    /* R sub-partitions per S partition */
    uint32_t work_items;       /* join work items */
    uint32_t used_learned;     /* 1 if the learned partitioner was used */
    uint64_t spilled_items;    /* work items whose R part exceeded one LDS table */
    float build_hist_ms;       /* of build_partition_ms: the histogram and its scan */
    float probe_hist_ms;       /* of probe_partition_ms: the histogram and its scan */
    uint32_t probe_appended;   /* 1 if the probe side took the appended layout (no histogram pass) */
    uint32_t probe_passes;     /* radix passes over the probe side (2: PJ_FLAG_LAYOUT_TWO_PASS's split) */
} pj_timings;

/* Resumable probe cursor: the state HashJoinOp keeps between getNext calls
 * (HashJoinState, operators/operators.h:3105-3114). Zero-initialise it;
 * release it with pj_cursor_release. */
typedef struct {
    void* impl;
    uint64_t pending;   /* rows produced and not yet handed out */
} pj_cursor;

const char* pj_last_error(void);
const char* pj_version(void);
/* A limit on the device memory the library's own buffers may hold in this
 * process (0, the default: none). No counterpart in the reference, whose
 * allocations fail only when the machine runs out (numaallocate.cpp). A
 * buffer that would pass it fails as a failed hipMalloc does: PJ_ERR_DEVICE
 * from the call that needed it, which a plan operator turns into Error.
 * Buffers held already are not given back. */
int pj_set_device_limit(uint64_t bytes);
/* The device memory the library's buffers hold now, in bytes (scratch of
 * the no-allocation entry point and caller-owned memory not counted). */
int pj_device_bytes(uint64_t* held);

/* Replaces HashJoinOp::init (operators/join.cpp:367-518): validates the
 * schemas, key columns and projection; computes the output schema. */
int pj_join_create(const pj_join_desc* desc, pj_join_t* out);
/* Output tuple width in bytes (HashJoinOp's output Schema::getTupleSize). */
int pj_join_out_width(pj_join_t h, uint32_t* width);
/* Replaces HashJoinOp::destroy + HashTable::destroy (join.cpp:828-836,
 * util/hashtable.cpp:243-252). */
int pj_join_destroy(pj_join_t h);

/* One whole join on device-resident relations: the compute of
 * HashJoinOp::scanStart (build, join.cpp:548-637) followed by every
 * HashJoinOp::getNext (probe, join.cpp:681-757). d_build/d_probe hold packed
 * build/probe tuples in HBM; d_out (16-byte aligned) receives up to out_cap
 * output tuples. *out_n gets the number of output tuples; if it exceeds out_cap the call
 * returns PJ_ERR_CAPACITY and the buffer holds an unspecified subset.
 * stream: a hipStream_t, or NULL for the default stream. Synchronous. */
int pj_join_run_device(pj_join_t h, const void* d_build, uint64_t nbuild,
                       const void* d_probe, uint64_t nprobe, void* d_out,
                       uint64_t out_cap, uint64_t* out_n, void* stream);

/* Host-page protocol used by the C++ HashJoinOp. */
/* Stages build tuples (HashJoinOp::buildFromPage, join.cpp:838-884).
 * Thread-safe; call from every build thread. */
/* aggregate_count grouped by the build key over the join of nbuild x nprobe
 * device tuples (aggregatecount.cpp:37-52 over HashJoinOp's output,
 * join.cpp:681-757), without writing the join's rows: one group tuple per
 * build key with matches, the key column(s) as in the build schema followed by
 * the count (LONG), the layout pj_agg_run_device gives for that count. The
 * count of a key is (build tuples with the key) x (probe tuples with it).
 * Always on radix partitions; the handle's projection is not used. Output
 * 4-byte aligned, room for groups_cap groups; PJ_ERR_CAPACITY when more.
 * Synchronous. */
int pj_join_count_by_key(pj_join_t h, const void* d_build, uint64_t nbuild, const void* d_probe, uint64_t nprobe,
                         void* d_groups, uint64_t groups_cap, uint64_t* ngroups, void* stream);
/* As pj_join_count_by_key with host buffers (copied to the device and
 * back), for the plan operators. Uses the null stream; synchronous. */
int pj_join_count_by_key_host(pj_join_t h, const void* build, uint64_t nbuild, const void* probe, uint64_t nprobe,
                              void* groups, uint64_t groups_cap, uint64_t* ngroups);
int pj_join_build_append(pj_join_t h, const void* tuples, uint64_t n);
/* Uploads and partitions the staged build side once the group barrier has
 * been passed (join.cpp:618). Exactly once per scan. */
int pj_join_build_finish(pj_join_t h);
/* Probes a batch of host probe tuples and hands out up to out_cap output
 * tuples per call (HashJoinOp::getNext, join.cpp:681-757, which returns
 * Ready when its output page fills and resumes from HashJoinState). Pass the
 * batch once (tuples != NULL); then call again with tuples == NULL and the
 * same cursor while cur->pending > 0. Thread-safe across cursors. */
int pj_join_probe(pj_join_t h, const void* tuples, uint64_t n, pj_cursor* cur,
                  void* out, uint64_t out_cap, uint64_t* out_n);
int pj_cursor_release(pj_cursor* cur);
/* Drops the staged and device-resident build side (HashTable::bucketclear +
 * destroy, join.cpp:802-826) so the handle can take another scan. */
int pj_join_reset(pj_join_t h);

int pj_join_last_timings(pj_join_t h, pj_timings* t);

/* Bytes of the fixed device scratch pj_join_run_device_noalloc works in. */
#define PJ_NOALLOC_SCRATCH_BYTES (128u << 20)

/* pj_join_run_device without device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under stream capture, tight serving loops, hosts that forbid hipMalloc).
 * The partitioned copies, histograms, plan and counters the join needs come
 * from one fixed array of PJ_NOALLOC_SCRATCH_BYTES in device memory, which
 * each device the library runs on holds until the process ends; the caller
 * still passes the inputs and the output. Always radix partitions (the same
 * rows as pj_join_run_device). A join whose partitioned copies do not fit
 * the scratch (about 1M x 5M 16-byte tuples) is refused with
 * PJ_ERR_UNSUPPORTED before any device work, as is a handle holding a staged
 * build side (PJ_ERR_STATE). Calls on one device share its scratch and are
 * serialised inside the process; overlapping use of one device's scratch
 * from elsewhere races. Synchronous; *out_n and PJ_ERR_CAPACITY as
 * pj_join_run_device. */
int pj_join_run_device_noalloc(pj_join_t h, const void* d_build, uint64_t nbuild, const void* d_probe,
                               uint64_t nprobe, void* d_out, uint64_t out_cap, uint64_t* out_n, void* stream);

/* PartitionOp (operators/partition.cpp:44-404) on device-resident tuples:
 * the exactrange hash on column attr (an INTEGER, LONG or DATE column) over
 * [range_min, range_max] into `buckets` partitions (hash.h:452-487).
 * d_out receives partition 0, then 1, ...; part_sizes (host, `buckets`
 * entries) gets the sizes. Within a partition the order is unspecified. */
int pj_partition_exactrange_device(const pj_schema* schema, uint32_t attr,
                                   int64_t range_min, int64_t range_max,
                                   uint32_t buckets, const void* d_in, uint64_t n,
                                   void* d_out, uint64_t* part_sizes, void* stream);

/* PartitionOp on host buffers: uploads, partitions on the device as
 * pj_partition_exactrange_device does, and copies the partitions back. */
int pj_partition_exactrange(const pj_schema* schema, uint32_t attr, int64_t range_min,
                            int64_t range_max, uint32_t buckets, const void* in, uint64_t n,
                            void* out, uint64_t* part_sizes);

/* Runs a query plan (libconfig text, as the reference's executequery reads,
 * drivers/executequery.cpp:105-154) through the C++ operators of this
 * library: node types merge, hashjoin, partition, threadidprepend, scan /
 * parallelscan (raw binary tables) and memscan (the buffers passed here,
 * by node name). The tree root must be a merge node. Output tuples of every
 * thread are concatenated into out (in any order); *out_bytes gets the bytes
 * produced, PJ_ERR_CAPACITY if they exceed out_cap_bytes. */
int pj_query_run(const char* plan_text, uint32_t nsources, const char* const* names,
                 const void* const* data, const uint64_t* tuples, void* out,
                 uint64_t out_cap_bytes, uint64_t* out_bytes, uint32_t* out_tuple_width);

/* Sets the calling thread's pj_last_error() text (used by the C++ layer). */
void pj_set_error(const char* msg);

/* Order-independent 64-bit checksum of device-resident tuples: the sum mod
 * 2^64 of a splitmix64-folded mix of each tuple's bytes, plus the XOR of all
 * 32-bit words that ConsumeOp computes (operators/consume.cpp:68-100). */
int pj_checksum_device(const void* d_tuples, uint64_t n, uint32_t width,
                       uint64_t* sum, uint32_t* xor32, void* stream);

/* Workload generators (pythia-core_amd/csrc/pj_datagen.h), on the device
 * (dst in HBM) or on host threads (dst in host memory). */
int pj_gen_pkfk(uint64_t seed, uint64_t domain, uint64_t n_build, uint64_t n_probe,
                uint64_t n_miss, void* build_dst, void* probe_dst, int on_device,
                void* stream);

/* Zipf CDF over ranks 1..n with exponent z, normalised, in double precision
 * (operators/generator_zipf.cpp:91-113). */
int pj_zipf_cdf(uint64_t n, double z, double* cdf);
/* Composite-key workload: build tuples (int32 a, int32 b, int64 payload) with
 * unique (a, b) in [0, 2^31)^2; probe tuples of the same shape whose key is
 * the key of a build tuple drawn Zipf(z) over ranks through cdf (a host
 * array of n_build doubles from pj_zipf_cdf), rank -> build tuple by a seeded
 * permutation (generator_zipf.cpp:85-89). */
int pj_gen_pairs_zipf(uint64_t seed, uint64_t n_build, uint64_t n_probe, const double* cdf,
                      void* build_dst, void* probe_dst, int on_device, void* stream);

/* Clustered-key workload: nclusters dense clusters (centres uniform in
 * [0, 2^62), widths 2^u for u uniform in [log2_wmin, log2_wmax], cluster
 * weights Zipf(weight_z) capped by width) holding n_keys distinct int64
 * keys; build tuples (key, payload) in a seeded order, probe tuples whose key
 * is that of a build tuple drawn Zipf through cdf (n_keys doubles). */
int pj_clusters_make(uint64_t seed, uint32_t nclusters, uint64_t n_keys, uint32_t log2_wmin,
                     uint32_t log2_wmax, double weight_z, void* table_out);
int pj_gen_clustered(const void* table, uint64_t n_probe, const double* cdf, void* build_dst,
                     void* probe_dst, int on_device, void* stream);
/* Bytes of the cluster table pj_clusters_make fills (at most 1024 clusters). */
uint64_t pj_cluster_table_size(void);

/* ---------------------------------------------------------------- learned
 * The learned partitioner's model, a two-stage piecewise-linear CDF over
 * ordered keys (the key with its sign bit flipped, so unsigned order is the
 * signed order of the reference's IntLongEqual keys). No reference
 * counterpart: it replaces the hash bits that choose a partition in
 * HashJoinOp's TupleHasher (hash.h:109-197) and PartitionOp's ranges
 * (partition.cpp:166-200) with ranges fitted to the key distribution.
 *
 * Root: the sorted sample is cut into segments wherever two neighbouring
 * sample keys are more than 2^32 apart; segment s owns leaves
 * seg_leaf0[s] .. seg_leaf0[s+1]-1 of equal key width seg_lw[s] starting at
 * seg_lo[s]. Leaf j: rank(key) ~ leaf_start[j] + a[j] + b[j] * x with
 * x = (key - leaf_lo(j)) / seg_lw in [0, 1), fitted by least squares over
 * the samples of leaves j-window .. j+window of its segment, then refined by
 * gradient steps on that loss plus lambda * (jump at each inner leaf edge)^2. */
#define PJ_MODEL_LEAVES 4096
#define PJ_MODEL_MAX_SEGS 1024
typedef struct {
    uint32_t nleaves;     /* leaves in use */
    uint32_t nseg;        /* root segments */
    uint64_t n;           /* sample keys fitted */
    uint64_t n_keys;      /* keys of the relation sampled */
    uint32_t window;      /* least-squares window, leaves either side */
    uint32_t iters;       /* gradient steps run */
    float lambda;         /* weight of the continuity term */
    float loss_lsq;       /* loss after the least-squares fit */
    float loss;           /* loss after the gradient steps */
    uint32_t reserved;
    uint64_t seg_lo[PJ_MODEL_MAX_SEGS];
    uint64_t seg_hi[PJ_MODEL_MAX_SEGS];
    uint64_t seg_lw[PJ_MODEL_MAX_SEGS];
    uint32_t seg_leaf0[PJ_MODEL_MAX_SEGS + 1];
    uint16_t leaf_seg[PJ_MODEL_LEAVES];
    uint32_t leaf_start[PJ_MODEL_LEAVES + 1];
    float a[PJ_MODEL_LEAVES];
    float b[PJ_MODEL_LEAVES];
} pj_model;

uint64_t pj_model_size(void);
/* Sorts n 64-bit unsigned keys in device memory, in place (LSD radix sort). */
int pj_sort_u64(uint64_t* d_keys, uint64_t n, void* stream);
/* Fits the model to n ordered sample keys in device memory (any order; they
 * are sorted into a copy) drawn from a relation of n_keys keys, running
 * iters gradient steps after the least-squares fit. */
int pj_model_fit(const uint64_t* d_keys, uint64_t n, uint64_t n_keys, uint32_t iters, pj_model* model,
                 void* stream);
/* Loss and gradient at the model's parameters over its (sorted) sample:
 * grad[j] = dL/da[j], grad[nleaves + j] = dL/db[j]. The forward and backward
 * kernels of one training step. */
int pj_model_grad(const pj_model* model, const uint64_t* d_sorted, uint64_t n, float* grad, double* loss,
                  void* stream);
/* Range boundaries (ordered keys) the model places for build partitions of
 * at most max_coarse probe partitions; *nbounds = partitions - 1. */
int pj_model_bounds(const pj_model* model, uint32_t max_coarse, uint64_t* bounds, uint32_t cap,
                    uint32_t* nbounds);

/* ---------------------------------------------------------------- batch
 * Many independent joins of one descriptor, run as one captured pipeline
 * (one hipGraph launch: a histogram, a scan and a scatter over every side of
 * every join, each join's work items counted on the device, and one
 * persistent join kernel over all of them). No reference counterpart: the
 * reference runs one HashJoinOp per plan (query.cpp:301-334); each join of a
 * batch gives the multiset of rows that operator gives for its two inputs. */
typedef struct pj_batch_s* pj_batch_t;
typedef struct {
    const void* build;   /* device tuples of the build side */
    uint64_t nbuild;
    const void* probe;   /* device tuples of the probe side */
    uint64_t nprobe;
    void* out;           /* device rows, 16-byte aligned */
    uint64_t out_cap;    /* rows out holds */
} pj_batch_join;
/* Per size class (floor(log2(build tuples))): the radix bits of both sides
 * (1..12; 0: the fewest whose mean partition fills at most load_factor of an
 * LDS table, or, with load_factor 0, the single join's choice), partition
 * passes (1: every side is partitioned into its 2^bits partitions in one
 * pass; 2 is refused), and the probe records of a work item in units of 4096
 * (piece; 0: about 1/8 of a CU's share of the join's probe side, at least
 * 32768). */
typedef struct {
    uint8_t bits;
    uint8_t passes;
    uint16_t piece;
    float load_factor;
} pj_batch_cfg;

int pj_batch_create(const pj_join_desc* desc, pj_batch_t* out);
int pj_batch_destroy(pj_batch_t h);
int pj_batch_set_config(pj_batch_t h, uint32_t size_class, const pj_batch_cfg* cfg);
int pj_batch_get_config(pj_batch_t h, uint32_t size_class, pj_batch_cfg* cfg);
/* Times the candidate configurations of every size class present on up to
 * `trial` joins of that class and keeps the fastest (pj_batch_prepare must
 * follow). */
int pj_batch_tune(pj_batch_t h, uint32_t njoins, const pj_batch_join* joins, uint32_t trial, void* stream);
/* Plans the batch and captures its pipeline into a graph. */
int pj_batch_prepare(pj_batch_t h, uint32_t njoins, const pj_batch_join* joins, void* stream);
/* Runs the prepared batch on stream: the pipeline's kernels launched one by
 * one (PJ_BATCH_PLAIN), as one graph launch (PJ_BATCH_GRAPH), every join as
 * a whole radix join in turn with the radix bits of its size class
 * (PJ_BATCH_SEQUENTIAL, the one-launch-per-join baseline), or the same
 * joins spread round-robin over four internal streams that start when
 * stream reaches the batch and that stream waits for (PJ_BATCH_CONCURRENT).
 * PJ_BATCH_CONCURRENT_GRAPH captures the concurrent mode's work (four
 * parallel branches) into a graph at its first run after pj_batch_prepare
 * or a change of radix bits, and then runs the batch as one graph launch;
 * done_us then holds the whole batch's time for every join, and
 * pj_batch_run_times zeros. The configurations of pj_batch_set_config and
 * pj_batch_tune (timed through the pipeline) set the radix bits of every
 * mode. */
#define PJ_BATCH_PLAIN 0
#define PJ_BATCH_GRAPH 1
#define PJ_BATCH_SEQUENTIAL 2
#define PJ_BATCH_CONCURRENT 3   /* as SEQUENTIAL, joins spread over 4 streams */
#define PJ_BATCH_CONCURRENT_GRAPH 4   /* CONCURRENT captured into one graph, one launch */
int pj_batch_run(pj_batch_t h, int mode, void* stream);
/* After a run: rows each join produced, and (done_us, optional) the time
 * from the batch's start until each join's last work item finished (device
 * clock for the pipeline, HIP events for PJ_BATCH_SEQUENTIAL).
 * PJ_ERR_CAPACITY when a join's output overflowed. */
int pj_batch_results(pj_batch_t h, uint64_t* out_counts, double* done_us, void* stream);
/* After a PJ_BATCH_SEQUENTIAL or PJ_BATCH_CONCURRENT run: each join's own
 * device time in microseconds, from its start on its stream to its end. */
int pj_batch_run_times(pj_batch_t h, double* run_us);

/* Batch workload: join j of a batch of njoins has build tuples (int64 key,
 * int32 payload) with unique keys (a seeded permutation of a uniform
 * 2^40 sample) and 4x as many probe tuples of the same shape whose key is a
 * uniformly drawn build key; build sizes are log-uniform in [nmin, nmax].
 * pj_batch_sizes fills the sizes; pj_gen_batch_join writes join j's tuples. */
int pj_batch_sizes(uint64_t seed, uint32_t njoins, uint64_t nmin, uint64_t nmax, uint64_t* nbuild);
int pj_gen_batch_join(uint64_t seed, uint32_t j, uint64_t nbuild, uint64_t nprobe, void* build_dst,
                      void* probe_dst, int on_device, void* stream);

/* ------------------------------------------------------------------------
 * Device aggregation: the aggregate_sum / aggregate_count plan nodes
 * (query.cpp:206-209) in global mode over packed tuples in HBM, for
 * instance the output of pj_join_run_device, so that only the groups leave
 * the device. Replaces GenericAggregate::scanStart / remember / getNext
 * (operators/genericaggregate.cpp:247-515) with AggregateSum's fold
 * (aggregatesum.cpp:37-107) or AggregateCount's (aggregatecount.cpp:37-52):
 * one output tuple per distinct value of the grouping columns (one for the
 * whole input when there are none, none for an empty input), the grouping
 * columns followed by the sum (the sum column's type: INTEGER sums wrap at
 * 32 bits, LONG sums at 64) or the count (LONG). Output order unspecified.
 * ---------------------------------------------------------------------- */
enum { PJ_AGG_SUM = 0, PJ_AGG_COUNT = 1 };

typedef struct {
    pj_schema in;           /* schema of the input tuples */
    uint32_t op;            /* PJ_AGG_SUM or PJ_AGG_COUNT */
    uint32_t nfields;       /* grouping columns: 0, 1 (INTEGER, LONG or DATE) or 2 (INTEGER, INTEGER) */
    uint16_t fields[2];     /* their column indexes ("field" / "fields") */
    uint16_t sumfield;      /* PJ_AGG_SUM: an INTEGER or LONG column ("sumfield") */
    uint16_t reserved;
    uint64_t max_groups;    /* distinct groups the device table holds (the plan's hash buckets) */
} pj_agg_desc;

typedef struct pj_agg_s* pj_agg_t;

/* Validates desc as GenericAggregate::init does (operators/genericaggregate.cpp:51-95,
 * aggregatesum.cpp:37-51); PJ_ERR_UNSUPPORTED for DECIMAL sums (their value
 * depends on the order of the additions), other grouping types, or input
 * columns the device reads off 4-byte boundaries. */
int pj_agg_create(const pj_agg_desc* desc, pj_agg_t* out);
int pj_agg_out_width(pj_agg_t h, uint32_t* width);
/* Aggregates n tuples at d_tuples (4-byte aligned) into d_groups (device
 * memory, room for groups_cap output tuples). *ngroups gets the number of
 * groups; PJ_ERR_CAPACITY when they exceed groups_cap or the table's
 * twice max_groups slots. With max_groups >= 2^16 and 2^20 <= n < 2^31 the
 * tuples are radix-partitioned by group first (device memory of about 16
 * bytes a tuple, kept by the handle). Synchronous; calls on one handle are
 * serialised by its lock. */
int pj_agg_run_device(pj_agg_t h, const void* d_tuples, uint64_t n, void* d_groups, uint64_t groups_cap,
                      uint64_t* ngroups, void* stream);
/* As pj_agg_run_device with host buffers: tuples are copied to the device,
 * aggregated there, and the groups copied back (the path of the plan
 * operators, whose pages live in host memory). Synchronous. */
int pj_agg_run_host(pj_agg_t h, const void* tuples, uint64_t n, void* groups, uint64_t groups_cap,
                    uint64_t* ngroups);
int pj_agg_destroy(pj_agg_t h);

/* ------------------------------------------------------------------------
 * Sort-merge join: the sortmergejoin / mpsmjoin plan nodes (query.cpp:167-184,
 * SortMergeJoinOp and MPSMJoinOp, operators/join.cpp:905-1791) on the
 * device. Same descriptor, output rows and limits as pj_join_*, and the same
 * output multiset; it exists to be timed against the hash join. Unlike the
 * reference's SortMergeJoinOp it also takes two INTEGER key columns.
 * ---------------------------------------------------------------------- */
typedef struct pj_smj_s* pj_smj_t;

typedef struct {
    float extract_ms;         /* tuples -> (key, payload) arrays, key ranges */
    float sort_ms;            /* both radix sorts (with one host read of the key ranges) */
    float merge_ms;           /* run search, scan, output rows */
    float total_ms;
    uint32_t build_key_bits;  /* low key bits sorted on each side */
    uint32_t probe_key_bits;
} pj_smj_timings;

/* Validates desc as pj_join_create does (SortMergeJoinOp::init, join.cpp:905-1020). */
int pj_smj_create(const pj_join_desc* desc, pj_smj_t* out);
int pj_smj_out_width(pj_smj_t h, uint32_t* width);
/* As pj_join_run_device: device tuples in, output rows into d_out (16-byte
 * aligned, room for out_cap rows); *out_n gets the row count, and
 * PJ_ERR_CAPACITY (nothing written) when it exceeds out_cap. Synchronous. */
int pj_smj_run_device(pj_smj_t h, const void* d_build, uint64_t nbuild, const void* d_probe, uint64_t nprobe,
                      void* d_out, uint64_t out_cap, uint64_t* out_n, void* stream);
/* As pj_smj_run_device with host buffers (copied to the device and back),
 * for the plan operators. Uses the null stream; synchronous. */
int pj_smj_run_host(pj_smj_t h, const void* build, uint64_t nbuild, const void* probe, uint64_t nprobe, void* out,
                    uint64_t out_cap, uint64_t* out_n);
int pj_smj_last_timings(pj_smj_t h, pj_smj_timings* t);
int pj_smj_destroy(pj_smj_t h);

/* ------------------------------------------------------------------------
 * TupleHasher on the device: the bucket id of every tuple under one of the
 * reference's hash functions (TupleHasher::create, hash.cpp:101-248, and
 * the hashers of hash.h:104-487), the ids a PartitionOp or a hash exchange
 * exposes. The join itself partitions on its own mix.
 * ---------------------------------------------------------------------- */
enum { PJ_HASH_MODULO = 0, PJ_HASH_PARAMMODULO = 1, PJ_HASH_KNUTH = 2, PJ_HASH_RANGE = 3, PJ_HASH_EXACTRANGE = 4,
       PJ_HASH_BYTES = 5, PJ_HASH_TPCHORDERKEY = 6, PJ_HASH_WILLIS = 7, PJ_HASH_ALWAYSZERO = 8,
       PJ_HASH_TPCHQ1MAGIC = 9 };

typedef struct {
    uint32_t fn;                   /* PJ_HASH_* (a plan's hash.fn) */
    uint32_t buckets;              /* hash.buckets; rounded up to a power of two except by exactrange */
    int64_t range_min, range_max;  /* hash.range, for range and exactrange */
    uint32_t offset, skipbits;     /* hash.offset, hash.skipbits (parammodulo, knuth) */
} pj_hash_spec;

/* HashFunction::buckets() of the function spec builds. */
int pj_hash_buckets(const pj_hash_spec* spec, uint32_t* buckets);
/* ParameterizedModuloValueHasher::generate (hash.cpp:251-281): the passes
 * parammodulo functions of a multi-pass radix split of spec (a parammodulo
 * function of 2 buckets or more) into out[0 .. passes-1], high bits first.
 * A tuple's bucket ids under them, most significant first, concatenate to its
 * bucket under spec. PJ_ERR_INVALID for passes = 0 or another function. */
int pj_hash_generate(const pj_hash_spec* spec, uint32_t passes, pj_hash_spec* out);
/* d_buckets[i] = the bucket of tuple i (n tuples of tuple_width bytes at
 * d_tuples, device memory). Value hashers read the key_size (4: INTEGER,
 * sign-extended; 8: LONG or DATE) bytes at key_offset; the byte hasher
 * folds key_size bytes there; tpchq1magic reads 4. Asynchronous on stream. */
int pj_tuple_hash_device(const pj_hash_spec* spec, const void* d_tuples, uint64_t n, uint32_t tuple_width,
                         uint32_t key_offset, uint32_t key_size, uint32_t* d_buckets, void* stream);

/* ------------------------------------------------------------------------
 * Raw binary tables into HBM: the binary scan source of the reference,
 * MemMappedTable::load (operators/loaders/table.cpp:170-347), which globs
 * the pattern and maps every non-empty regular file it names, in glob's
 * sorted order, as packed tuples with no header. Here the files land back
 * to back in one device buffer, read through pinned staging buffers by host
 * threads while earlier chunks are copied to the device.
 * ---------------------------------------------------------------------- */
/* Tuples of tuple_width bytes in the files pattern names (for sizing). */
int pj_binary_tuples(const char* pattern, uint32_t tuple_width, uint64_t* ntuples);
/* Loads them into d_dst (device memory, room for cap tuples); *ntuples gets
 * their number. PJ_ERR_CAPACITY when they exceed cap, PJ_ERR_INVALID for a
 * file whose size is not a whole number of tuples (the reference would read
 * a partial last tuple). No match is an empty table. The copies are on
 * stream; returns after they have completed. */
int pj_load_binary(const char* pattern, uint32_t tuple_width, void* d_dst, uint64_t cap, uint64_t* ntuples,
                   void* stream);

/* ------------------------------------------------------------------------
 * Measurement kernels (no counterpart in the reference): the rates the
 * machine gives the access patterns the join is made of, to set a kernel's
 * time against. n 16-byte records in d_a (and d_b); each call times reps
 * launches after one warm-up, on stream, and keeps the fastest.
 *   READ       n records of d_a streamed in; d_b gets grid x 256 words.
 *   WRITE      n records streamed out to d_b.
 *   COPY       d_a to d_b (bytes counts both directions).
 *   GATHER     d_b[i] = d_a[d_idx[i]] (bytes: record in + index + record out);
 *              d_idx holds n indices below n (a permutation, say).
 *   SCATTER    d_b[d_idx[i]] = d_a[i].
 *   HISTOGRAM  LDS counts of 2^param bins (1..15) over the top bits of the
 *              mixed first 8 bytes of every record of d_a; d_b gets the
 *              2048 x 2^param block counts.
 *   ATOMIC     64 device atomicAdds a lane, results used, over param (a
 *              power of two) counters at d_a; d_b gets a word a lane.
 *   LOOKUP     n random 16-byte reads at hashed positions of a table of
 *              2^param records at the front of d_a (2^param <= n); d_b gets
 *              a word a lane. The lookup rate of a table that size.
 *   PROBE      n records of d_a streamed in, each looked up at a hashed
 *              position of a table of 2^param records at the front of d_a,
 *              a 16-byte result a record streamed out to d_b (bytes: 32 a
 *              record, the lookups not counted).
 * param is the grid of READ, WRITE and COPY (0: 8192). Buffers are the
 * caller's; the call allocates nothing on the device.
 * ---------------------------------------------------------------------- */
enum { PJ_MEASURE_READ = 0, PJ_MEASURE_WRITE = 1, PJ_MEASURE_COPY = 2, PJ_MEASURE_GATHER = 3,
       PJ_MEASURE_SCATTER = 4, PJ_MEASURE_HISTOGRAM = 5, PJ_MEASURE_ATOMIC = 6, PJ_MEASURE_LOOKUP = 7,
       PJ_MEASURE_PROBE = 8 };

typedef struct {
    float ms;         /* fastest launch */
    uint64_t bytes;   /* bytes the pattern moves a launch (0 for ATOMIC) */
    uint64_t ops;     /* records (or atomics) a launch */
    double gbps;      /* bytes / ms, GB/s */
    double gops;      /* ops / ms, G/s */
} pj_measure_result;

int pj_measure(uint32_t kind, void* d_a, void* d_b, const uint32_t* d_idx, uint64_t n, uint32_t param, uint32_t reps,
               pj_measure_result* res, void* stream);

#ifdef __cplusplus
}
#endif
#endif
