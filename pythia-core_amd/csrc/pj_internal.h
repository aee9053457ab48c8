// pj_internal.h -- shared declarations of libpjoin: tuple layouts, the
// canonical 16-byte device record, error plumbing and small device helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/pjoin.h"
#include "pj_datagen.h"

namespace pj {

// ---------------------------------------------------------------- errors
void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);

#define PJ_HIP(call)                                                              \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess)                                                     \
            return ::pj::fail(PJ_ERR_DEVICE, "%s failed: %s (%s:%d)", #call,       \
                              hipGetErrorString(e_), __FILE__, __LINE__);          \
    } while (0)

// ---------------------------------------------------------------- layouts
// Where the canonical record's words come from in a packed tuple of one
// side. A canonical record is four 32-bit words {key_lo, key_hi, pay0, pay1}:
// the join key normalised to 64 bits (an INTEGER key is sign-extended, so
// INTEGER/LONG keys compare as the reference's IntLongEqual does; two
// INTEGER key columns are packed lo/hi), and up to two words of the
// side's projected non-key columns.
//
// Schemas the record cannot hold (keys of two LONG columns or of three or
// more columns, more than 8 payload bytes a side, CHAR or wide projections,
// columns off 4-byte boundaries) take the row-id form instead: the record is
// {h_lo, h_hi, row, 0}, h a 64-bit hash of the key columns normalised to
// int64 and row the tuple's index in its relation. The join then pairs rows
// whose hashes are equal, and the materialising pass checks every key column
// on the tuples themselves and copies the projection (pj_rows.hip).
struct SideLayout {
    uint32_t tw;   // tuple width in 32-bit words
    int32_t k0;    // word index of key_lo
    int32_t k1;    // word index of key_hi, or -1: sign-extend key_lo
    int32_t p0;    // word index of pay0, or -1: zero
    int32_t p1;    // word index of pay1, or -1: zero
    uint32_t ident;  // 1 when the tuple is exactly the record (tw 4, 0,1,2,3)
    uint32_t tbytes = 0;   // tuple width in bytes (both forms)
    // row-id form (hashed = 1): the key columns
    uint32_t hashed = 0;
    uint32_t nk = 0;
    uint16_t koff[PJ_MAX_KEYS] = {};   // byte offset of key column k
    uint8_t kint[PJ_MAX_KEYS] = {};    // 1: INTEGER (4 bytes, sign-extended), 0: LONG or DATE
    uint64_t hmask = ~0ull;            // bits of the key hash kept (all but under PJ_FLAG_NARROW_ROW_HASH)
};

// Output tuple assembly: word w of an output tuple is word src[w] of the
// 8-word concatenation {build record, probe record}.
enum { SRC_BK0 = 0, SRC_BK1, SRC_BP0, SRC_BP1, SRC_PK0, SRC_PK1, SRC_PP0, SRC_PP1 };
struct OutFormat {
    uint32_t words;      // output tuple width in words (<= 8)
    uint8_t src[8];
};

// The row-id form's output: key columns compared on the tuples (the
// reference's conjunction of IntInt / IntLong / LongInt / LongLong
// equalities, rawcompfns.cpp:48-180) and the projection copied byte for byte
// (Schema::writeData of each projected column, schema.inl:47-93).
struct RowKey {
    uint16_t boff, poff;    // byte offsets in the build / probe tuple
    uint8_t bint, pint;     // 1: INTEGER, 0: LONG or DATE
};
struct RowCopy {
    uint8_t side;           // PJ_SIDE_BUILD or PJ_SIDE_PROBE
    uint8_t chr;            // 1: a CHAR column, copied as strncpy copies it (zeros after the first NUL)
    uint16_t src, dst, width;   // byte offset in the tuple, in the output row; bytes
};
struct RowSpec {
    uint32_t nk, ncopy;
    uint32_t bw, pw, ow;    // build, probe and output tuple widths in bytes
    uint32_t words;         // 1: every width and offset is a multiple of 4 (copies by words)
    RowKey key[PJ_MAX_KEYS];
    RowCopy copy[PJ_MAX_COLS];
};

struct Layout {
    SideLayout build, probe;
    OutFormat out;
    uint32_t out_width;   // bytes
    uint32_t rowid = 0;   // 1: the row-id form (out is then the (build row, probe row) pair)
    RowSpec rows = {};
};

// Validates a descriptor and derives the layouts (see pj_join.hip).
int derive_layout(const pj_join_desc* d, Layout* L);

// Byte offsets of a packed schema's columns (the reference's Schema layout,
// no alignment) and the tuple size; PJ_ERR_INVALID for a bad schema.
int schema_offsets(const pj_schema& s, int* off, int* size);

// ---------------------------------------------------------------- device
constexpr uint32_t kJoinThreads = 1024;
constexpr uint32_t kTableCap = 8944;       // records per LDS table (~140 KiB)
constexpr uint32_t kTableBuckets = 4096;   // bucket directory entries
constexpr uint32_t kLdsBytes = 163840;     // gfx950 LDS per CU
constexpr uint32_t kScatterThreads = 512;

// LDS table formats of range-partitioned build partitions.
enum : uint32_t {
    kFmtCodes = 0,   // 32-bit key codes (key - base) in hash buckets, any multiplicity
    kFmtDense = 1,   // the partition holds every key of [base, base + r_cnt) once:
                     // payloads indexed by key - base, no keys stored
};
constexpr uint32_t kCodeCap = 12928;    // records per code table (12 B each + bucket directory)
constexpr uint32_t kDenseCap = 20096;   // records per dense table (8 B each + presence bitmap)

// A range build partition's LDS table, for work items derived on the device.
struct PlanPart {
    uint32_t fmt, cap;      // kFmtCodes or kFmtDense; records a table holds
    uint64_t lo, hi, base;  // ordered-key range and smallest ordered key
};

// The work items of a join, derived on the device (no item list): offs[p]
// is the first item of build partition p and offs[FR] the item count. Item
// q of partition p joins table split q / pieces of its build records with
// piece q % pieces of its probe partition.
struct PlanView {
    const uint32_t* offs;        // FR + 1
    const uint32_t* rst;         // FR + 1 build partition starts
    const uint32_t* sst;         // probe partition starts
    const uint32_t* coarse_of;   // range plans: probe partition of each build partition
    const PlanPart* pp;          // range plans: table of each build partition (null: radix)
    const uint32_t* cl;          // appended probe layout: chunk numbers, partition by partition (null: contiguous)
    uint64_t chunk;              // probe records a piece, about
    uint32_t FR, shift;          // build partitions; radix: probe partition = p >> shift
};

struct Item {           // one join work item
    uint32_t r_off, r_cnt;   // build records [r_off, r_off + r_cnt) of R_part
    uint32_t s_off, s_cnt;   // probe records of S_part
    uint32_t part;           // build partition id; records with other ids are skipped
    uint32_t fmt;            // range partitions: kFmtCodes or kFmtDense
    uint64_t lo, hi;         // range partitions: ordered-key range [lo, hi] of the build partition
    uint64_t base;           // range partitions: smallest ordered build key of the partition
};

// The learned partition function (pj_learned.hip plans it, pj_device.h
// evaluates it): the model's root cuts the ordered key range into segments,
// each segment into leaves of equal key width, and leaf j of segment s
// carries its fitted line in partition units, so that ordered key u in leaf
// j at fraction x of the leaf lies in build partition
//     first[s] + clamp(floor(c[j] + d[j] * x), 0, np[s] - 1).
// Segment s owns the keys from the middle of the gap before it (cut[s]) to
// the middle of the gap after it, so keys just outside the sampled ends of
// a segment stay with it (before its first leaf: x < 0, its first
// partition; past its last: its last partition). A directory of kMapDir
// equal slices of [cut[0], cut[last]] gives the segment of u without a
// search in almost every slice. All of it sits in one blob of kMapBytes at
// fixed offsets, which the kernels that evaluate the function copy into LDS.
constexpr uint32_t kMapSegs = PJ_MODEL_MAX_SEGS;
constexpr uint32_t kMapLeaves = PJ_MODEL_LEAVES;
constexpr uint32_t kMapDir = 4096;
struct MapHdr {
    uint32_t nseg, nleaves, nfine, dshift;   // dshift: slice width log2
    uint64_t lo0;                            // cut[0], the first segment's first sampled key
    uint64_t reserved;
};
constexpr uint32_t kMapOffSegLo = 32;                                   // uint64 [kMapSegs]: first sampled key
constexpr uint32_t kMapOffSegCut = kMapOffSegLo + kMapSegs * 8;         // uint64 [kMapSegs]: first key owned
constexpr uint32_t kMapOffSegInv = kMapOffSegCut + kMapSegs * 8;        // float [kMapSegs]: 1 / leaf width
constexpr uint32_t kMapOffSegMeta = kMapOffSegInv + kMapSegs * 4;       // uint2 [kMapSegs]: (leaf0 | nl << 16, first | np << 16)
constexpr uint32_t kMapOffDir = kMapOffSegMeta + kMapSegs * 8;          // uint16 [kMapDir + 1]
constexpr uint32_t kMapOffLeaf = kMapOffDir + ((kMapDir + 1) * 2 + 15) / 16 * 16;   // float2 [kMapLeaves]: (c, d)
constexpr uint32_t kMapBytes = kMapOffLeaf + kMapLeaves * 8;
static_assert(kMapBytes % 16 == 0, "the blob is copied in 16-byte words");

// A partition function: radix (the top `bits` bits of the key mix, F =
// 2^bits) or the learned function in the blob at map (device memory), its
// build partition divided by div (probe partitions pair up div build
// partitions; div = 1 for the build side). divm = ceil(2^32 / div) for div
// > 1: the high word of fine * divm is fine / div exactly for fine and div
// below 2^16 (fine = q div + r: fine * divm / 2^32 = q + (r + fine e / 2^32)
// / div with e = divm div - 2^32 < div, and fine e / 2^32 < 1).
struct PartFn {
    uint32_t F;
    uint32_t bits;
    const uint8_t* map;
    uint32_t div;
    uint32_t divm;
};

inline PartFn radix_fn(uint32_t bits) { return PartFn{1u << bits, bits, nullptr, 1, 0}; }
inline PartFn map_fn(uint32_t F, const uint8_t* map, uint32_t div)
{
    return PartFn{F, 0, map, div, div > 1 ? (uint32_t)(((1ull << 32) + div - 1) / div) : 0u};
}

// ---------------------------------------------------------------- host
// A fixed device region that buffers are carved from in order (the
// no-allocation entry point's scratch): nothing is freed, and a buffer that
// must grow takes a new piece.
struct Arena {
    uint8_t* base;
    size_t size, used;
};

// A device buffer that grows on demand: from hipMalloc, or, when arena is
// set, from that region (release then only forgets it).
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    Arena* arena = nullptr;
    int ensure(size_t bytes);
    void release();
    ~DevBuf() { release(); }
};

// Radix partitioning for consumers other than the join: n tuples of layout L
// become 16-byte records {key, payload} in 2^bits partitions by the top bits
// of the mixed key, partition p at records [starts[p], starts[p + 1]) of
// recs (device words). A partition holds padding records whose key belongs
// to another partition; consumers skip them by recomputing the partition.
// Probe sides of at least this many tuples take the appended layout by
// default (PJ_FLAG_LAYOUT_EXACT / _APPEND force either).
constexpr uint64_t kAppendMin = 1ull << 22;

struct RadixRecords {
    DevBuf recs, starts, hist, bsum, pids;
};
int radix_records(RadixRecords& r, const void* d_in, uint64_t n, const SideLayout& L, uint32_t bits, hipStream_t st);

// The range partitioning chosen by the learned model for one build side.
struct RangePlan {
    bool ok = false;                 // false: the model declined (radix is used)
    std::string why;                 // reason when declined
    uint32_t nfine = 0, ncoarse = 0; // build partitions; probe partitions = ceil(nfine / div)
    uint32_t div = 1;                // build partitions a probe partition serves
    std::vector<uint8_t> map;        // kMapBytes: the learned partition function's blob
    std::vector<uint32_t> coarse_of; // nfine: probe partition of each build partition
    uint32_t nseg = 0, nleaves = 0;
    float loss_lsq = 0.f, loss = 0.f;
};

// Radix bits of the build side for nr tuples (forced >= 0 overrides): the
// fewest bits whose mean partition, plus five standard deviations, fits an
// LDS table (pj_join.hip).
uint32_t choose_build_bits(uint64_t nr, int32_t forced);

// Enqueues a whole radix join (handle h, PJ_PART_RADIX) on st without
// waiting for it. Its counters land in slot, 8 words of pinned host memory,
// when the stream gets there: rows produced in words 2-3 (low, high), work
// items in word 6, split tables in word 7 (pj_join.hip). The handle's device
// buffers are reused by the next join enqueued on it, in stream order; they
// grow (which waits for the device) when a join needs more than any before.
int join_enqueue(pj_join_s* h, const void* d_build, uint64_t nb, const void* d_probe, uint64_t np, void* d_out,
                 uint64_t out_cap, uint32_t* slot, hipStream_t st);

// The row-id form's second pass (pj_rows.hip): pairs[0 .. min(*ncand,
// pair_cap)) are (build row, probe row) pairs whose key hashes are equal.
// Pair i whose key columns are equal becomes output row i; a pair whose keys
// differ adds one to *nfail and leaves row i unwritten. compact = true
// instead writes the rows whose keys are equal densely, reserving them on
// *rows (the pass the host runs again when *nfail is not zero). Rows at or
// past out_cap are counted, not written.
int rows_materialize(const uint2* pairs, const unsigned long long* ncand, uint64_t pair_cap, const void* d_build,
                     const void* d_probe, const RowSpec& spec, void* d_out, uint64_t out_cap,
                     unsigned long long* rows, uint32_t* nfail, bool compact, hipStream_t st);

// ---------------------------------------------------------------- batched joins
// Many radix joins of one descriptor run as one pipeline of kernels, each
// launched once for the whole batch (pj_join.hip; pj_batch.hip drives it):
// per-tile histograms of every side of every join, one scan over all of
// them, a tile-sorted scatter into one record buffer, each side's partition
// starts, each join's work items counted and scanned on the device, and one
// persistent join kernel over every join's items.

// One side (build or probe) of one join of a batch.
struct MSide {
    const uint32_t* in;   // raw tuples
    uint32_t n;           // tuples
    uint32_t bits;        // radix bits: 1 << bits partitions
    uint32_t probe;       // 0: build side (layout L0), 1: probe side (L1)
    uint32_t nblk;        // tiles (blocks of the histogram and scatter kernels)
    uint32_t chunk;       // tuples a tile
    uint32_t blk0;        // first block of this side's tiles in those kernels' grids
    uint32_t hbase;       // first histogram counter: (1 << bits) * nblk of them, partition-major
    uint32_t rbase;       // first record of the side in the batch's record buffer
    uint32_t pbase;       // first of its (1 << bits) + 1 partition starts
    uint32_t pad;
};

// One join of a batch: its work items (v: offs, rst and sst point into the
// batch's buffers; rst and sst count records from the record buffer's start)
// and where its rows go.
struct MJoin {
    PlanView v;
    uint32_t* out;
    uint64_t out_cap;
    uint32_t bits;        // radix bits of both sides
    uint32_t pad;
};

// Per join, written by the kernels.
struct MState {
    unsigned long long rows;     // output rows reserved (rows at or past out_cap are not written)
    uint32_t remaining;          // work items not finished
    uint32_t items;              // work items
    unsigned long long t_done;   // device clock (100 MHz) when its last item finished
};

// A prepared batch: descriptors and buffers on the device.
struct MBatch {
    uint32_t njoins = 0, nsides = 0, nblocks = 0, maxF = 1;
    uint64_t nhist = 0, nrec = 0, nstarts = 0, noffs = 0, max_items = 0;
    int rm = 0;                  // record mode of the tuples
    Layout lay;
    DevBuf sides, joins, hist, bsum, recs, starts, offs, ibase, state, misc;
};

// Lets the pipeline's kernels use the LDS a batch of this shape needs (call
// outside a stream capture), then enqueues the whole pipeline on st (no host
// sync, no memset: a caller may capture it into a graph).
int mbatch_setup(const MBatch& b);
int mbatch_enqueue(MBatch& b, hipStream_t st);
// LDS records of the batch scatter's tiles for maxF partitions a side.
uint32_t mbatch_scatter_tile(uint32_t maxF);

// Device state of the learned partitioner (pj_learned.hip).
struct LearnedCtx {
    DevBuf sample, tmp, hist, bsum, model, params, moments, resid, bounds, segs, loss;
    std::vector<uint8_t> host_model;   // pj_model
    ~LearnedCtx() {}
};

// Trains the CDF model on a sample of the build keys and derives the range
// partitioning (pj_learned.hip). Returns PJ_OK with plan->ok false when the
// key distribution does not suit range partitions.
int learned_plan(LearnedCtx& ctx, const void* d_build, uint64_t nb, const SideLayout& L, uint32_t max_coarse,
                 hipStream_t st, RangePlan* plan);

}  // namespace pj
