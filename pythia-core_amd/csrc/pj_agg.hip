// pj_agg.hip -- device aggregation over packed tuples in HBM: the
// aggregate_sum / aggregate_count plan nodes in global mode
// (operators/genericaggregate.cpp, aggregatesum.cpp, aggregatecount.cpp),
// typically over the join's output so that only the groups leave the device.
//
//   k_agg      every block folds its share of the tuples into an LDS hash
//              table of kAggSlots groups (linear probing, at most kAggProbes
//              probes); a tuple whose group finds no slot goes straight to the
//              global table; at the end the block flushes its table. Without
//              grouping each thread folds in registers and the block makes one
//              global atomic.
//   k_agg_emit one output tuple per used slot of the global table.
//
// The global table (open addressing, at least twice max_groups slots) is
// lock-free: a 64-bit CAS claims a slot for a key, a 64-bit atomic add folds
// the value into it, and nothing waits. The key kEmpty marks an empty slot;
// a group whose key happens to be kEmpty is folded into a slot of its own.
//
// Many groups (max_groups >= kPartMinGroups) take the partitioned path
// instead: the join's radix partition pass turns the tuples into 16-byte
// {group key, value} records in 2^bits partitions of about kPartGroups
// groups each, and k_agg_parts aggregates one partition a block in an LDS
// table of kPartSlots groups and writes the groups itself. Partitions hold
// disjoint keys, so only a group that finds no LDS slot (a partition with
// more groups than expected) goes to the global table, and the global table
// is cleared and scanned only after a run that used it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "pj_device.h"
#include "pj_internal.h"

namespace pj {
namespace {

constexpr uint64_t kEmpty = 0xA5A5A5A5A5A5A5A5ull;
constexpr uint32_t kAggThreads = 512;
constexpr uint32_t kAggSlots = 4096;     // LDS groups a block: 64 KiB of keys and sums
constexpr uint32_t kAggProbes = 8;
constexpr uint32_t kPartThreads = 1024;
constexpr uint32_t kPartSlots = 8192;    // LDS groups a partition: 128 KiB of keys and sums
constexpr uint32_t kPartProbes = 32;
constexpr uint32_t kPartUnroll = 8;
constexpr uint32_t kPartGroups = 4096;   // expected groups a partition (half the slots)
constexpr uint64_t kPartMinGroups = 1ull << 16;
constexpr uint64_t kPartMinTuples = 1ull << 20;
constexpr uint32_t kPartMaxBits = 13;
constexpr uint32_t kMiscWords = 8;
constexpr uint32_t kClaimedCap = 1u << 16;   // claimed slots listed for a short emit and clear

// How a tuple's key and value are read: 32-bit words at byte offsets.
struct AggLayout {
    uint32_t tw;            // tuple words
    uint32_t nkey;          // 0, 1 or 2 grouping columns
    uint32_t key_off[2];    // word offsets
    uint32_t key_wide[2];   // 1: 8-byte column (LONG, DATE)
    uint32_t val_off;       // word offset of the summed column
    uint32_t val_wide;      // 1: LONG
    uint32_t count;         // 1: AggregateCount
};

struct AggTable {
    unsigned long long* key;   // G slots
    unsigned long long* val;
    unsigned long long* misc;  // [0] kEmpty group's value, [1] its tuples, [2] overflow, [3] groups emitted,
                               // [4] nonzero once a key went to the table (partitioned path), [5] slots claimed
    uint32_t* claimed;         // the first kClaimedCap slots claimed, in claim order
    uint32_t mask;             // G - 1
};

__device__ __forceinline__ uint64_t agg_key(const uint32_t* t, const AggLayout& L)
{
    if (L.nkey == 0)
        return 0;
    const uint64_t a = L.key_wide[0] ? ((uint64_t)t[L.key_off[0]] | ((uint64_t)t[L.key_off[0] + 1] << 32))
                                     : (uint64_t)t[L.key_off[0]];
    if (L.nkey == 1)
        return a;
    return (a & 0xFFFFFFFFull) | ((uint64_t)t[L.key_off[1]] << 32);   // two INTEGER columns
}

__device__ __forceinline__ uint64_t agg_val(const uint32_t* t, const AggLayout& L)
{
    if (L.count)
        return 1;
    if (L.val_wide)
        return (uint64_t)t[L.val_off] | ((uint64_t)t[L.val_off + 1] << 32);
    return (uint64_t)(int64_t)(int32_t)t[L.val_off];   // INTEGER, sign-extended; its low 32 bits wrap as int32
}

__device__ __forceinline__ void global_fold(const AggTable& G, uint64_t k, uint64_t v)
{
    if (k == kEmpty) {
        atomicAdd(&G.misc[0], (unsigned long long)v);
        atomicAdd(&G.misc[1], 1ull);
        return;
    }
    uint32_t s = (uint32_t)pj_fin64(k) & G.mask;
    for (uint32_t i = 0; i <= G.mask; ++i) {
        const unsigned long long prev = atomicCAS(&G.key[s], (unsigned long long)kEmpty, (unsigned long long)k);
        if (prev == kEmpty || prev == k) {
            atomicAdd(&G.val[s], (unsigned long long)v);
            if (prev == kEmpty) {
                const unsigned long long c = atomicAdd(&G.misc[5], 1ull);
                if (c < kClaimedCap)
                    G.claimed[c] = s;
            }
            return;
        }
        s = (s + 1) & G.mask;
    }
    atomicOr(&G.misc[2], 1ull);   // every slot taken: more groups than the table holds
}

__global__ __launch_bounds__(kAggThreads) void k_agg(const uint32_t* __restrict__ in, uint64_t n, AggLayout L,
                                                      AggTable G)
{
    __shared__ unsigned long long lds[2 * kAggSlots];   // 64 KiB: keys, then sums
    unsigned long long* lkey = lds;
    unsigned long long* lval = lds + kAggSlots;
    unsigned long long* wpart = lds;                      // per-wave sums, without grouping
    // Each block folds one contiguous range of tuples (a multiple of the block
    // size), so that input clustered by group -- the join writes each probe
    // round's matches as one run -- meets the LDS table while its groups are hot.
    const uint64_t per = ((n + gridDim.x - 1) / gridDim.x + kAggThreads - 1) / kAggThreads * kAggThreads;
    const uint64_t lo = (uint64_t)blockIdx.x * per;
    const uint64_t hi = lo + per < n ? lo + per : n;
    if (L.nkey == 0) {   // one group: fold in registers, one atomic a block
        uint64_t acc = 0;
        for (uint64_t i = lo + threadIdx.x; i < hi; i += kAggThreads)
            acc += agg_val(in + i * L.tw, L);
        for (int d = 32; d > 0; d >>= 1)
            acc += __shfl_down(acc, d, 64);
        if ((threadIdx.x & 63) == 0)
            wpart[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t b = 0;
            for (uint32_t w = 0; w < kAggThreads / 64; ++w)
                b += wpart[w];
            if (lo < hi)
                global_fold(G, 0, b);
        }
        return;
    }
    for (uint32_t s = threadIdx.x; s < kAggSlots; s += kAggThreads) {
        lkey[s] = kEmpty;
        lval[s] = 0;
    }
    __syncthreads();
    for (uint64_t i = lo + threadIdx.x; i < hi; i += kAggThreads) {
        const uint32_t* t = in + i * L.tw;
        const uint64_t k = agg_key(t, L), v = agg_val(t, L);
        bool done = false;
        if (k != kEmpty) {
            uint32_t s = (uint32_t)pj_fin64(k) & (kAggSlots - 1);
            for (uint32_t p = 0; p < kAggProbes; ++p) {
                // Read first: only an empty slot needs the CAS that claims it.
                unsigned long long cur = __hip_atomic_load(&lkey[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (cur == kEmpty)
                    cur = atomicCAS(&lkey[s], (unsigned long long)kEmpty, (unsigned long long)k);
                if (cur == kEmpty || cur == k) {
                    atomicAdd(&lval[s], (unsigned long long)v);
                    done = true;
                    break;
                }
                s = (s + 1) & (kAggSlots - 1);
            }
        }
        if (!done)
            global_fold(G, k, v);
    }
    __syncthreads();
    for (uint32_t s = threadIdx.x; s < kAggSlots; s += kAggThreads)
        if (lkey[s] != kEmpty)
            global_fold(G, lkey[s], lval[s]);
}

// Writes a word-aligned 8-byte value as two 4-byte stores (packed output
// tuples put 8-byte columns at any 4-byte offset).
__device__ __forceinline__ void put64(uint32_t* o, uint64_t v)
{
    o[0] = (uint32_t)v;
    o[1] = (uint32_t)(v >> 32);
}

// One output tuple: the grouping columns, then the sum or count.
__device__ __forceinline__ void write_group(uint32_t* __restrict__ o, const AggLayout& L, uint64_t k, uint64_t v)
{
    uint32_t w = 0;
    for (uint32_t c = 0; c < L.nkey; ++c) {
        if (L.nkey == 2) {   // two INTEGER columns: low then high half of the key
            o[w++] = c ? (uint32_t)(k >> 32) : (uint32_t)k;
        } else if (L.key_wide[c]) {
            put64(o + w, k);
            w += 2;
        } else {
            o[w++] = (uint32_t)k;
        }
    }
    if (L.count || L.val_wide)
        put64(o + w, v);
    else
        o[w] = (uint32_t)v;
}

__device__ __forceinline__ void emit_group(uint32_t* __restrict__ out, uint64_t cap, const AggLayout& L,
                                           uint32_t ow, const AggTable& G, uint64_t k, uint64_t v)
{
    const unsigned long long pos = atomicAdd(&G.misc[3], 1ull);
    if (pos < cap)
        write_group(out + pos * ow, L, k, v);
}

__global__ void k_agg_emit(AggTable G, AggLayout L, uint32_t ow, uint32_t* __restrict__ out, uint64_t cap,
                           uint32_t only_if_used)
{
    if (only_if_used && G.misc[4] == 0)
        return;
    const uint32_t first = blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t nclaimed = G.misc[5];
    if (only_if_used && nclaimed <= kClaimedCap) {   // few groups in the table: the claim list names them
        for (uint64_t i = first; i < nclaimed; i += stride) {
            const uint32_t s = G.claimed[i];
            emit_group(out, cap, L, ow, G, G.key[s], G.val[s]);
        }
    } else {
        for (uint64_t s = first; s <= G.mask; s += stride) {
            const uint64_t k = G.key[s];
            if (k != kEmpty)
                emit_group(out, cap, L, ow, G, k, G.val[s]);
        }
    }
    if (first == 0 && G.misc[1])
        emit_group(out, cap, L, ow, G, kEmpty, G.misc[0]);
}

// Clears the whole table (table = 1), or only the nlisted slots of the
// previous run's claim list, and the counters.
__global__ void k_agg_clear(AggTable G, uint32_t table, uint32_t nlisted)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (table && s <= G.mask) {
        G.key[s] = kEmpty;
        G.val[s] = 0;
    }
    if (!table && s < nlisted) {
        const uint32_t t = G.claimed[s];
        G.key[t] = kEmpty;
        G.val[t] = 0;
    }
    if (s < kMiscWords)
        G.misc[s] = 0;
}

// One partition of {key, value} records a block (grid-stride over the
// partitions): fold into an LDS table, then write the table's groups to the
// output, one atomic a wave-chunk of slots for their positions. Records
// whose key belongs to another partition are the partition pass's padding.
__global__ __launch_bounds__(kPartThreads) void k_agg_parts(const uint4* __restrict__ recs,
                                                            const uint32_t* __restrict__ starts, uint32_t bits,
                                                            uint32_t F, uint32_t count, AggTable G, AggLayout L,
                                                            uint32_t ow, uint32_t* __restrict__ out, uint64_t cap)
{
    extern __shared__ unsigned long long pl[];
    unsigned long long* lkey = pl;                 // kPartSlots
    unsigned long long* lval = pl + kPartSlots;    // kPartSlots
    __shared__ uint32_t wcnt[kPartThreads / 64];
    __shared__ unsigned long long cbase;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (uint32_t p = blockIdx.x; p < F; p += gridDim.x) {
        for (uint32_t s = tid; s < kPartSlots; s += kPartThreads) {
            lkey[s] = kEmpty;
            lval[s] = 0;
        }
        __syncthreads();
        const uint32_t lo = starts[p], hi = starts[p + 1];
        bool used_global = false;
        // kPartUnroll records a thread in flight before any of them is folded
        for (uint32_t b0 = lo; b0 < hi; b0 += kPartThreads * kPartUnroll) {
            uint4 r[kPartUnroll];
#pragma unroll
            for (uint32_t u = 0; u < kPartUnroll; ++u) {
                const uint32_t i = b0 + u * kPartThreads + tid;
                if (i < hi)
                    r[u] = ld_stream(recs + i);
            }
#pragma unroll
            for (uint32_t u = 0; u < kPartUnroll; ++u) {
                const uint32_t i = b0 + u * kPartThreads + tid;
                if (i >= hi)
                    break;
                const uint64_t k = rec_key(r[u]);
                const uint64_t h = key_mix(k);
                if (radix_part(h, bits) != p)
                    continue;   // padding
                const uint64_t v = count ? 1ull : ((uint64_t)r[u].z | ((uint64_t)r[u].w << 32));
                bool done = false;
                if (k != kEmpty) {
                    // linear probing over aligned 4-slot buckets: the four keys are
                    // read together, then the first empty slot is claimed
                    uint32_t s = (uint32_t)h & (kPartSlots - 4);
                    uint32_t hs = 0;
                    for (uint32_t q = 0; q < kPartProbes / 4 && !done; ++q) {
                        unsigned long long c[4];
#pragma unroll
                        for (uint32_t j = 0; j < 4; ++j)
                            c[j] = __hip_atomic_load(&lkey[s + j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
                        for (uint32_t j = 0; j < 4; ++j) {
                            if (!done) {
                                unsigned long long cur = c[j];
                                if (cur == kEmpty)
                                    cur = atomicCAS(&lkey[s + j], (unsigned long long)kEmpty, (unsigned long long)k);
                                if (cur == kEmpty || cur == k) {
                                    hs = s + j;
                                    done = true;
                                }
                            }
                        }
                        s = (s + 4) & (kPartSlots - 1);
                    }
                    if (done)
                        atomicAdd(&lval[hs], (unsigned long long)v);
                }
                if (!done) {
                    global_fold(G, k, v);
                    used_global = true;
                }
            }
        }
        if (used_global)
            G.misc[4] = 1;
        __syncthreads();
        // the partition's groups: counted, their rows reserved with one atomic, then written
        uint32_t mine = 0;
        for (uint32_t s = tid; s < kPartSlots; s += kPartThreads)
            mine += lkey[s] != kEmpty;
        for (int d = 32; d > 0; d >>= 1)
            mine += __shfl_xor(mine, d, 64);
        if (lane == 0)
            wcnt[wave] = mine;
        __syncthreads();
        if (tid == 0) {
            uint32_t tot = 0;
            for (uint32_t w = 0; w < kPartThreads / 64; ++w)
                tot += wcnt[w];
            cbase = tot ? atomicAdd(&G.misc[3], (unsigned long long)tot) : 0ull;
        }
        __syncthreads();
        uint64_t next = cbase;
        for (uint32_t c = 0; c < kPartSlots; c += kPartThreads) {
            const uint32_t s = c + tid;
            const uint64_t k = lkey[s];
            const bool has = k != kEmpty;
            const uint64_t m = __ballot(has);
            const uint32_t rank = (uint32_t)__popcll(m & (lane ? (~0ull >> (64 - lane)) : 0ull));
            if (lane == 0)
                wcnt[wave] = (uint32_t)__popcll(m);
            __syncthreads();
            uint32_t before = 0, total = 0;
            for (uint32_t w = 0; w < kPartThreads / 64; ++w) {
                before += w < wave ? wcnt[w] : 0u;
                total += wcnt[w];
            }
            __syncthreads();
            if (has) {
                const uint64_t pos = next + before + rank;
                if (pos < cap)
                    write_group(out + pos * ow, L, k, lval[s]);
            }
            next += total;
        }
    }
}

}  // namespace
}  // namespace pj

using namespace pj;

struct pj_agg_s {
    pj_agg_desc desc;
    AggLayout L;
    SideLayout rec;                  // tuple -> {group key, value} record of the partitioned path
    uint32_t bits = 0;               // radix bits of the partitioned path
    uint32_t out_words = 0;
    uint32_t slots = 0;
    DevBuf table;
    bool table_dirty = true;         // the global table may hold keys from an earlier run
    uint64_t claimed_prev = ~0ull;   // slots the last run claimed (all listed when <= kClaimedCap)
    RadixRecords parts;
    uint64_t* host_misc = nullptr;   // pinned readback of the table's counters
    std::mutex mu;
    std::mutex host_mu;              // guards the host path's device copies
    DevBuf host_in, host_out;
};

namespace {

// Blocks of 256 threads for a grid-stride pass over the global table.
uint32_t emit_blocks(uint64_t slots)
{
    return (uint32_t)std::min<uint64_t>((slots + 255) / 256, 2048);
}

int allow_part_lds()
{
    static std::mutex mu;
    static bool done[64] = {};
    int dev = 0;
    PJ_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(mu);
    if (dev >= 0 && dev < 64 && done[dev])
        return PJ_OK;
    PJ_HIP(hipFuncSetAttribute((const void*)k_agg_parts, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)(2 * kPartSlots * 8)));
    if (dev >= 0 && dev < 64)
        done[dev] = true;
    return PJ_OK;
}

}  // namespace

extern "C" {

int pj_agg_create(const pj_agg_desc* desc, pj_agg_t* out)
{
    if (!desc || !out)
        return fail(PJ_ERR_INVALID, "null argument");
    *out = nullptr;
    int off[PJ_MAX_COLS], size;
    int rc = schema_offsets(desc->in, off, &size);
    if (rc)
        return rc;
    if (desc->op != PJ_AGG_SUM && desc->op != PJ_AGG_COUNT)
        return fail(PJ_ERR_INVALID, "op must be PJ_AGG_SUM or PJ_AGG_COUNT");
    if (desc->nfields > 2)
        return fail(PJ_ERR_UNSUPPORTED, "at most two grouping columns");
    if (desc->max_groups == 0 || desc->max_groups > (1ull << 30))
        return fail(PJ_ERR_INVALID, "max_groups must be 1..2^30");
    if (size % 4)
        return fail(PJ_ERR_UNSUPPORTED, "input tuples of %d bytes: the device reads 4-byte words", size);
    AggLayout L = {};
    L.tw = (uint32_t)size / 4;
    L.nkey = desc->nfields;
    uint32_t ow = 0;
    for (uint32_t k = 0; k < desc->nfields; ++k) {
        const uint32_t f = desc->fields[k];
        if (f >= desc->in.ncols)
            return fail(PJ_ERR_INVALID, "grouping column %u out of range", f);
        const uint32_t t = desc->in.type[f];
        const bool wide = t == PJ_CT_LONG || t == PJ_CT_DATE;
        if (!(t == PJ_CT_INTEGER || wide) || (desc->nfields == 2 && t != PJ_CT_INTEGER))
            return fail(PJ_ERR_UNSUPPORTED, "grouping column %u: one INTEGER, LONG or DATE column, or two INTEGER", f);
        if (off[f] % 4)
            return fail(PJ_ERR_UNSUPPORTED, "grouping column %u is not 4-byte aligned", f);
        L.key_off[k] = (uint32_t)off[f] / 4;
        L.key_wide[k] = wide ? 1u : 0u;
        ow += wide ? 2u : 1u;
    }
    if (desc->op == PJ_AGG_SUM) {
        const uint32_t f = desc->sumfield;
        if (f >= desc->in.ncols)
            return fail(PJ_ERR_INVALID, "sumfield %u out of range", f);
        const uint32_t t = desc->in.type[f];
        if (t == PJ_CT_DECIMAL)
            return fail(PJ_ERR_UNSUPPORTED, "DECIMAL sums depend on the order of the additions");
        if (t != PJ_CT_INTEGER && t != PJ_CT_LONG)
            return fail(PJ_ERR_INVALID, "sumfield %u must be a numeric column", f);   // aggregatesum.cpp:42-45
        if (off[f] % 4)
            return fail(PJ_ERR_UNSUPPORTED, "sumfield %u is not 4-byte aligned", f);
        L.val_off = (uint32_t)off[f] / 4;
        L.val_wide = t == PJ_CT_LONG ? 1u : 0u;
        ow += L.val_wide ? 2u : 1u;
    } else {
        L.count = 1;
        ow += 2;   // LONG count (aggregatecount.cpp:37-42)
    }
    pj_agg_s* h = new (std::nothrow) pj_agg_s();
    if (!h)
        return fail(PJ_ERR_INVALID, "out of host memory");
    h->desc = *desc;
    h->L = L;
    h->out_words = ow;
    SideLayout& R = h->rec;   // key words as the partition pass reads them; value words, or none for a count
    R.tw = L.tw;
    R.k0 = (int32_t)L.key_off[0];
    R.k1 = L.nkey == 2 ? (int32_t)L.key_off[1] : L.key_wide[0] ? (int32_t)L.key_off[0] + 1 : -1;
    R.p0 = L.count ? -1 : (int32_t)L.val_off;
    R.p1 = L.count || !L.val_wide ? -1 : (int32_t)L.val_off + 1;
    R.ident = R.tw == 4 && R.k0 == 0 && R.k1 == 1 && R.p0 == 2 && R.p1 == 3;
    while (h->bits < kPartMaxBits && ((uint64_t)kPartGroups << h->bits) < desc->max_groups)
        ++h->bits;
    h->bits = std::max<uint32_t>(h->bits, 1);
    uint64_t slots = 1024;
    while (slots < 2 * desc->max_groups)
        slots <<= 1;
    h->slots = (uint32_t)slots;
    *out = h;
    return PJ_OK;
}

int pj_agg_out_width(pj_agg_t h, uint32_t* width)
{
    if (!h || !width)
        return fail(PJ_ERR_INVALID, "null argument");
    *width = h->out_words * 4;
    return PJ_OK;
}

int pj_agg_run_device(pj_agg_t h, const void* d_tuples, uint64_t n, void* d_groups, uint64_t groups_cap,
                      uint64_t* ngroups, void* stream)
{
    if (!h || !ngroups || (n && !d_tuples) || (groups_cap && !d_groups))
        return fail(PJ_ERR_INVALID, "null argument");
    if ((uintptr_t)d_tuples % 4 || (uintptr_t)d_groups % 4)
        return fail(PJ_ERR_INVALID, "input and output must be 4-byte aligned");
    std::lock_guard<std::mutex> g(h->mu);
    hipStream_t st = (hipStream_t)stream;
    const uint64_t slots = h->slots;
    int rc;
    if ((rc = h->table.ensure((2 * slots + kMiscWords) * 8 + kClaimedCap * 4)))
        return rc;
    if (!h->host_misc)
        PJ_HIP(hipHostMalloc((void**)&h->host_misc, kMiscWords * 8, hipHostMallocDefault));
    AggTable G;
    G.key = (unsigned long long*)h->table.p;
    G.val = G.key + slots;
    G.misc = G.val + slots;
    G.claimed = (uint32_t*)(G.misc + kMiscWords);
    G.mask = (uint32_t)(slots - 1);
    const bool parts = h->L.nkey > 0 && h->desc.max_groups >= kPartMinGroups && n >= kPartMinTuples &&
                       n < (1ull << 31);
    // the whole table, unless the last run left nothing in it or listed every slot it claimed
    const bool clear_list = parts && h->table_dirty && h->claimed_prev <= kClaimedCap;
    const bool clear_table = !parts || (h->table_dirty && !clear_list);
    const uint32_t nlisted = clear_list ? (uint32_t)h->claimed_prev : 0u;
    const uint64_t clear_threads = clear_table ? slots : std::max<uint64_t>(nlisted, kMiscWords);
    h->table_dirty = true;   // until a finished run shows otherwise
    h->claimed_prev = ~0ull;
    hipLaunchKernelGGL(k_agg_clear, dim3((uint32_t)((clear_threads + 255) / 256)), dim3(256), 0, st, G,
                       clear_table ? 1u : 0u, nlisted);
    if (parts) {
        if ((rc = radix_records(h->parts, d_tuples, n, h->rec, h->bits, st)) || (rc = allow_part_lds()))
            return rc;
        const uint32_t F = 1u << h->bits;
        hipLaunchKernelGGL(k_agg_parts, dim3(F), dim3(kPartThreads), 2 * kPartSlots * 8, st,
                           (const uint4*)h->parts.recs.p, (const uint32_t*)h->parts.starts.p, h->bits, F, h->L.count,
                           G, h->L, h->out_words, (uint32_t*)d_groups, groups_cap);
        hipLaunchKernelGGL(k_agg_emit, dim3(emit_blocks(slots)), dim3(256), 0, st, G, h->L, h->out_words,
                           (uint32_t*)d_groups, groups_cap, 1u);
    } else if (n) {
        int dev = 0, cus = 256;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
            cus = 256;
        const uint64_t want = (n + 8191) / 8192;
        const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, (uint64_t)cus * 4));
        hipLaunchKernelGGL(k_agg, dim3(grid), dim3(kAggThreads), 0, st, (const uint32_t*)d_tuples, n, h->L, G);
        hipLaunchKernelGGL(k_agg_emit, dim3(emit_blocks(slots)), dim3(256), 0, st, G, h->L, h->out_words,
                           (uint32_t*)d_groups, groups_cap, 0u);
    }
    PJ_HIP(hipGetLastError());
    PJ_HIP(hipMemcpyAsync(h->host_misc, G.misc, kMiscWords * 8, hipMemcpyDeviceToHost, st));
    PJ_HIP(hipStreamSynchronize(st));
    h->table_dirty = !parts || h->host_misc[4] != 0;
    h->claimed_prev = h->host_misc[5];
    *ngroups = h->host_misc[3];
    if (h->host_misc[2])
        return fail(PJ_ERR_CAPACITY, "more distinct groups than the table's %llu slots (max_groups %llu)",
                    (unsigned long long)slots, (unsigned long long)h->desc.max_groups);
    if (*ngroups > groups_cap)
        return fail(PJ_ERR_CAPACITY, "%llu groups, the output holds %llu", (unsigned long long)*ngroups,
                    (unsigned long long)groups_cap);
    return PJ_OK;
}

int pj_agg_run_host(pj_agg_t h, const void* tuples, uint64_t n, void* groups, uint64_t groups_cap,
                    uint64_t* ngroups)
{
    if (!h || !ngroups || (n && !tuples) || (groups_cap && !groups))
        return fail(PJ_ERR_INVALID, "null argument");
    *ngroups = 0;
    std::lock_guard<std::mutex> g(h->host_mu);
    int in_size = 0, off[PJ_MAX_COLS];
    int rc = schema_offsets(h->desc.in, off, &in_size);
    if (rc)
        return rc;
    const uint64_t ow = (uint64_t)h->out_words * 4;
    // room on the device for every group the input can have
    const uint64_t dcap = std::max<uint64_t>(std::min<uint64_t>(n, std::max<uint64_t>(groups_cap, 1)), 1);
    if ((rc = h->host_in.ensure(std::max<uint64_t>(n * (uint64_t)in_size, 16))) || (rc = h->host_out.ensure(dcap * ow)))
        return rc;
    if (n)
        PJ_HIP(hipMemcpy(h->host_in.p, tuples, n * (uint64_t)in_size, hipMemcpyHostToDevice));
    rc = pj_agg_run_device(h, h->host_in.p, n, h->host_out.p, dcap, ngroups, nullptr);
    if (rc == PJ_ERR_CAPACITY && *ngroups > dcap && *ngroups > groups_cap)
        return fail(PJ_ERR_CAPACITY, "%llu groups, the output holds %llu", (unsigned long long)*ngroups,
                    (unsigned long long)groups_cap);
    if (rc)
        return rc;
    if (*ngroups)
        PJ_HIP(hipMemcpy(groups, h->host_out.p, *ngroups * ow, hipMemcpyDeviceToHost));
    return PJ_OK;
}

int pj_agg_destroy(pj_agg_t h)
{
    if (!h)
        return PJ_OK;
    h->table.release();
    if (h->host_misc)
        (void)hipHostFree(h->host_misc);
    delete h;
    return PJ_OK;
}

}  // extern "C"
