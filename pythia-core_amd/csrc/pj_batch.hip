// pj_batch.hip -- many independent joins in one captured pipeline.
//
// No counterpart in the reference, which runs one HashJoinOp per plan
// (query.cpp:301-334); this is the batched form of the same join for
// callers with hundreds of small and medium joins. Every phase runs once
// for the whole batch:
//   k_bhist / scan / k_bscatter   radix partition of every join (pass 1),
//                                 one block per tile of one side of one join
//   k_bbounds                     partition starts of pass 1
//   k_bpass2                      second pass for joins tuned to two passes:
//                                 one block per first-pass partition
//   k_bitem_count / scan / emit   the work list, planned on the device
//   k_bjoin                       a persistent kernel over the work list:
//                                 per item an LDS open-addressing table
//                                 (linear probing, load factor from the
//                                 tuner) and a streaming probe
// All of it is captured once into a hipGraph and replayed with one launch.
// The last item of each join stamps the device clock, which gives every
// join's completion time within the batch.
//
// The tuner picks, per size class (floor(log2(build tuples))), radix bits
// 4..12, one or two passes and a load factor 0.5..0.9 by timing candidate
// configurations on joins of that class from the batch itself.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "pj_device.h"
#include "pj_internal.h"
#include "pj_probe.h"
#include "pj_scan.h"

namespace pj {
namespace {

constexpr uint32_t kBThreads = 512;       // batch join and partition blocks
constexpr uint32_t kBSlotsLog2 = 12;      // largest LDS table: 4096 slots
constexpr uint32_t kBTile = 16384;        // tuples per pass-1 tile
constexpr uint32_t kBMinBits = 4, kBMaxBits = 12;
constexpr uint32_t kBChunk = 65536;       // probe records per work item at most
constexpr uint32_t kBScatterTile = 4096;  // records per LDS-sorted scatter round
constexpr uint32_t kClasses = 32;
constexpr uint32_t kLanes = 4;            // streams of the concurrent mode (4 hardware queues a process)

struct BJoin {                 // one join of the batch (device copy)
    const uint32_t* in[2];     // raw build / probe tuples
    uint64_t n[2];
    uint32_t* out;
    uint64_t out_cap;
    uint32_t bits, b1, b2;     // partition bits = b1 (pass 1) + b2 (pass 2)
    uint32_t slots_log2;       // LDS table slots of this join's partitions
    uint32_t cap;              // build records a table holds (load factor * slots)
    uint32_t nt[2];            // pass-1 tiles per side
    uint64_t hbase[2];         // first pass-1 histogram entry per side
    uint32_t pbase[2];         // first pass-1 partition start entry per side (F1 + 1 entries)
    uint32_t fbase;            // first global build partition index
    uint32_t p2base;           // first pass-2 block of this join (two-pass joins)
};

struct BTile {
    uint32_t join, side, t, count;
    uint64_t start;
};

struct BItem {
    uint32_t r_off, r_cnt, s_off, s_cnt;
    uint32_t join, slots_log2, shift, pad;
};

struct BState {                // per join, written by the kernels
    unsigned long long cursor; // output rows reserved
    uint32_t remaining;        // work items not finished
    uint32_t pad;
    unsigned long long t_done; // device clock when the last item finished
};

__device__ __forceinline__ uint64_t dev_clock() { return __builtin_amdgcn_s_memrealtime(); }

// ---------------------------------------------------------------- pass 1

__global__ __launch_bounds__(kBThreads) void k_bhist(const BJoin* __restrict__ J, const BTile* __restrict__ tiles,
                                                      SideLayout L0, SideLayout L1, uint32_t* __restrict__ hist)
{
    extern __shared__ uint32_t bcnt[];
    const BTile T = tiles[blockIdx.x];
    const BJoin& j = J[T.join];
    const uint32_t F = 1u << j.b1;
    const SideLayout L = T.side ? L1 : L0;
    for (uint32_t p = threadIdx.x; p < F; p += kBThreads)
        bcnt[p] = 0;
    __syncthreads();
    const uint32_t* in = j.in[T.side];
    for (uint32_t i = threadIdx.x; i < T.count; i += kBThreads) {
        const uint4 r = load_rec<false>(in, T.start + i, L);
        atomicAdd(&bcnt[radix_part(key_mix(rec_key(r)), j.b1)], 1u);
    }
    __syncthreads();
    const uint32_t nt = j.nt[T.side];
    for (uint32_t p = threadIdx.x; p < F; p += kBThreads)
        hist[j.hbase[T.side] + (uint64_t)p * nt + T.t] = bcnt[p];
}

// LDS-sorted scatter of one tile into the pass-1 buffer A.
__global__ __launch_bounds__(kBThreads) void k_bscatter(const BJoin* __restrict__ J, const BTile* __restrict__ tiles,
                                                         SideLayout L0, SideLayout L1,
                                                         const uint32_t* __restrict__ hist, uint4* __restrict__ A)
{
    extern __shared__ uint4 bsm[];
    const BTile T = tiles[blockIdx.x];
    const BJoin& j = J[T.join];
    const uint32_t F = 1u << j.b1;
    const SideLayout L = T.side ? L1 : L0;
    uint4* tile = bsm;                                                    // kBScatterTile
    uint32_t* cnt = reinterpret_cast<uint32_t*>(tile + kBScatterTile);    // F + 1
    uint32_t* goff = cnt + (1u << kBMaxBits) + 1;                         // F
    uint32_t* wsum = goff + (1u << kBMaxBits);                            // 16
    const uint32_t nt = j.nt[T.side];
    for (uint32_t p = threadIdx.x; p < F; p += kBThreads)
        goff[p] = hist[j.hbase[T.side] + (uint64_t)p * nt + T.t];
    const uint32_t* in = j.in[T.side];
    constexpr int K = kBScatterTile / kBThreads;
    for (uint32_t t0 = 0; t0 < T.count; t0 += kBScatterTile) {
        const uint32_t tn = T.count - t0 < kBScatterTile ? T.count - t0 : kBScatterTile;
        for (uint32_t p = threadIdx.x; p <= F; p += kBThreads)
            cnt[p] = 0;
        __syncthreads();
        uint4 r[K];
        uint32_t part[K], rank[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t i = k * kBThreads + threadIdx.x;
            if (i < tn) {
                r[k] = load_rec<false>(in, T.start + t0 + i, L);
                part[k] = radix_part(key_mix(rec_key(r[k])), j.b1);
                rank[k] = atomicAdd(&cnt[part[k]], 1u);
            }
        }
        __syncthreads();
        block_scan_array<kBThreads>(cnt, F, wsum);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t i = k * kBThreads + threadIdx.x;
            if (i < tn)
                tile[cnt[part[k]] + rank[k]] = r[k];
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < tn; i += kBThreads) {
            const uint4 v = tile[i];
            const uint32_t p = radix_part(key_mix(rec_key(v)), j.b1);
            A[goff[p] + (i - cnt[p])] = v;
        }
        __syncthreads();
        for (uint32_t p = threadIdx.x; p < F; p += kBThreads)
            goff[p] += cnt[p + 1] - cnt[p];
        __syncthreads();
    }
}

// pstart[pbase + p] = first record of pass-1 partition p of (join, side);
// pstart[pbase + F1] = the end of the side's region. For one-pass joins the
// pass-1 partitions are the build partitions: fstart / fcount are written.
__global__ void k_bbounds(const BJoin* __restrict__ J, uint32_t njoins, const uint32_t* __restrict__ hist,
                          uint32_t* __restrict__ pstart, uint32_t* __restrict__ fstart0,
                          uint32_t* __restrict__ fcount0, uint32_t* __restrict__ fstart1,
                          uint32_t* __restrict__ fcount1)
{
    const uint32_t ji = blockIdx.x, side = blockIdx.y;
    if (ji >= njoins)
        return;
    const BJoin& j = J[ji];
    const uint32_t F = 1u << j.b1, nt = j.nt[side];
    const uint32_t base = hist[j.hbase[side]];
    uint32_t* fs = side ? fstart1 : fstart0;
    uint32_t* fc = side ? fcount1 : fcount0;
    for (uint32_t p = threadIdx.x; p <= F; p += blockDim.x) {
        const uint32_t s = p < F ? hist[j.hbase[side] + (uint64_t)p * nt] : base + (uint32_t)j.n[side];
        pstart[j.pbase[side] + p] = s;
        if (j.b2 == 0 && p < F) {
            const uint32_t e = p + 1 < F ? hist[j.hbase[side] + (uint64_t)(p + 1) * nt] : base + (uint32_t)j.n[side];
            fs[j.fbase + p] = s;
            fc[j.fbase + p] = e - s;
        }
    }
}

// ---------------------------------------------------------------- pass 2
// One block per (two-pass join, side, pass-1 partition): counts the next b2
// radix bits of its records, then scatters them into buffer B at the same
// place, each sub-partition after the previous one.
struct P2Block {
    uint32_t join, side, p1, pad;
};

__global__ __launch_bounds__(kBThreads) void k_bpass2(const BJoin* __restrict__ J, const P2Block* __restrict__ blocks,
                                                       const uint32_t* __restrict__ pstart,
                                                       const uint4* __restrict__ A, uint4* __restrict__ B,
                                                       uint32_t* __restrict__ fstart0, uint32_t* __restrict__ fcount0,
                                                       uint32_t* __restrict__ fstart1, uint32_t* __restrict__ fcount1)
{
    __shared__ uint32_t cnt[(1u << kBMaxBits) + 1];
    __shared__ uint32_t wsum[kBThreads / 64 + 1];
    const P2Block P = blocks[blockIdx.x];
    const BJoin& j = J[P.join];
    const uint32_t F2 = 1u << j.b2;
    const uint32_t s0 = pstart[j.pbase[P.side] + P.p1], s1 = pstart[j.pbase[P.side] + P.p1 + 1];
    for (uint32_t p = threadIdx.x; p <= F2; p += kBThreads)
        cnt[p] = 0;
    __syncthreads();
    const uint32_t sh = 64 - j.b1 - j.b2;
    const uint32_t m2 = F2 - 1;
    for (uint32_t i = s0 + threadIdx.x; i < s1; i += kBThreads)
        atomicAdd(&cnt[(uint32_t)(key_mix(rec_key(A[i])) >> sh) & m2], 1u);
    __syncthreads();
    block_scan_array<kBThreads>(cnt, F2, wsum);   // cnt[p] = start of sub-partition p, cnt[F2] = total
    uint32_t* fs = P.side ? fstart1 : fstart0;
    uint32_t* fc = P.side ? fcount1 : fcount0;
    for (uint32_t p = threadIdx.x; p < F2; p += kBThreads) {
        fs[j.fbase + P.p1 * F2 + p] = s0 + cnt[p];
        fc[j.fbase + P.p1 * F2 + p] = cnt[p + 1] - cnt[p];
    }
    __syncthreads();
    for (uint32_t i = s0 + threadIdx.x; i < s1; i += kBThreads) {
        const uint4 v = A[i];
        const uint32_t p = (uint32_t)(key_mix(rec_key(v)) >> sh) & m2;
        B[s0 + atomicAdd(&cnt[p], 1u)] = v;
    }
}

// ---------------------------------------------------------------- work list

__device__ __forceinline__ uint32_t join_of(const uint32_t* __restrict__ fbase, uint32_t njoins, uint32_t g)
{
    uint32_t lo = 0, hi = njoins;   // last join with fbase <= g
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (fbase[mid] <= g)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t items_of(uint32_t rc, uint32_t sc, uint32_t cap)
{
    return rc && sc ? ((rc + cap - 1) / cap) * ((sc + kBChunk - 1) / kBChunk) : 0u;
}

__global__ void k_bitem_count(const BJoin* __restrict__ J, const uint32_t* __restrict__ fbase, uint32_t njoins,
                              uint32_t G, const uint32_t* __restrict__ fcount0,
                              const uint32_t* __restrict__ fcount1, uint32_t* __restrict__ icnt,
                              BState* __restrict__ state)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= G)
        return;
    const uint32_t ji = join_of(fbase, njoins, g);
    const uint32_t c = items_of(fcount0[g], fcount1[g], J[ji].cap);
    icnt[g] = c;
    if (c)
        atomicAdd(&state[ji].remaining, c);
}

__global__ void k_bitem_emit(const BJoin* __restrict__ J, const uint32_t* __restrict__ fbase, uint32_t njoins,
                             uint32_t G, const uint32_t* __restrict__ fstart0, const uint32_t* __restrict__ fcount0,
                             const uint32_t* __restrict__ fstart1, const uint32_t* __restrict__ fcount1,
                             const uint32_t* __restrict__ ioff, BItem* __restrict__ items)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= G)
        return;
    const uint32_t ji = join_of(fbase, njoins, g);
    const BJoin& j = J[ji];
    const uint32_t rc = fcount0[g], sc = fcount1[g];
    if (!items_of(rc, sc, j.cap))
        return;
    uint32_t o = ioff[g];
    for (uint32_t r = 0; r < rc; r += j.cap)
        for (uint32_t s = 0; s < sc; s += kBChunk) {
            BItem it;
            it.r_off = fstart0[g] + r;
            it.r_cnt = rc - r < j.cap ? rc - r : j.cap;
            it.s_off = fstart1[g] + s;
            it.s_cnt = sc - s < kBChunk ? sc - s : kBChunk;
            it.join = ji;
            it.slots_log2 = j.slots_log2;
            it.shift = 64 - j.bits - j.slots_log2;
            it.pad = j.b2 ? 1u : 0u;   // records in buffer B (two passes) or A
            items[o++] = it;
        }
}

// Zeroes the per-join state and stamps the batch's start.
__global__ void k_bstart(BState* __restrict__ state, uint32_t njoins, uint32_t* __restrict__ counters,
                         unsigned long long* __restrict__ t0)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < njoins) {
        state[i].cursor = 0;
        state[i].remaining = 0;
        state[i].t_done = 0;
    }
    if (i < 4)
        counters[i] = 0;
    if (i == 0)
        *t0 = dev_clock();
}

// ---------------------------------------------------------------- join

// LDS open-addressing table with linear probing: slot j holds key[j] and
// pay[j] when bit j of occ is set. A probe walks from its home slot to the
// first empty one, so the table is never full (load factor < 1).
struct OATab {
    const uint64_t* key;
    const uint2* pay;
    const uint32_t* occ;
    uint32_t mask, shift;
    __device__ __forceinline__ bool used(uint32_t j) const { return (occ[j >> 5] >> (j & 31u)) & 1u; }
    __device__ __forceinline__ void lookup(const uint4& s, uint32_t& nm, uint32_t& first) const
    {
        const uint64_t k = rec_key(s);
        uint32_t j = (uint32_t)(key_mix(k) >> shift) & mask;
        while (used(j)) {
            if (key[j] == k) {
                if (nm == 0)
                    first = j;
                nm++;
            }
            j = (j + 1) & mask;
        }
    }
    __device__ __forceinline__ uint64_t mkey(const uint4& s) const { return rec_key(s); }
    __device__ __forceinline__ bool match(uint32_t j, uint64_t mk) const { return key[j] == mk; }
    __device__ __forceinline__ uint32_t next(uint32_t j) const { return (j + 1) & mask; }
    __device__ __forceinline__ uint64_t bkey(uint32_t j, const uint4&) const { return key[j]; }
    __device__ __forceinline__ uint2 bpay(uint32_t j) const { return pay[j]; }
};

__global__ __launch_bounds__(kBThreads) void k_bjoin(const uint4* __restrict__ A, const uint4* __restrict__ B,
                                                      const BItem* __restrict__ items,
                                                      const uint32_t* __restrict__ nitems_ptr,
                                                      uint32_t* __restrict__ item_ctr, const BJoin* __restrict__ J,
                                                      BState* __restrict__ state, OutFormat fmt)
{
    extern __shared__ uint64_t bl64[];
    const uint32_t tid = threadIdx.x;
    uint32_t* misc = reinterpret_cast<uint32_t*>(bl64);   // [0] item, [2..17] wave bases
    uint32_t* wsum = misc + 40;
    uint64_t* key = bl64 + 32;
    uint2* pay = reinterpret_cast<uint2*>(key + (1u << kBSlotsLog2));
    uint32_t* occ = reinterpret_cast<uint32_t*>(pay + (1u << kBSlotsLog2));
    const bool kbp = fmt.words == 6 && fmt.src[0] == SRC_BK0 && fmt.src[1] == SRC_BK1 && fmt.src[2] == SRC_BP0 &&
                     fmt.src[3] == SRC_BP1 && fmt.src[4] == SRC_PP0 && fmt.src[5] == SRC_PP1;
    const uint32_t nitems = *nitems_ptr;
    for (;;) {
        if (tid == 0)
            misc[0] = atomicAdd(item_ctr, 1u);
        __syncthreads();
        const uint32_t it = misc[0];
        if (it >= nitems)
            break;
        const BItem I = items[it];
        const uint4* src = I.pad ? B : A;
        const uint32_t slots = 1u << I.slots_log2, mask = slots - 1;
        for (uint32_t w = tid; w < (slots + 31) / 32; w += kBThreads)
            occ[w] = 0;
        __syncthreads();
        for (uint32_t i = tid; i < I.r_cnt; i += kBThreads) {
            const uint4 r = src[I.r_off + i];
            const uint64_t k = rec_key(r);
            uint32_t j = (uint32_t)(key_mix(k) >> I.shift) & mask;
            for (;;) {
                const uint32_t bit = 1u << (j & 31u);
                if (!(atomicOr(&occ[j >> 5], bit) & bit)) {
                    key[j] = k;
                    pay[j] = make_uint2(r.z, r.w);
                    break;
                }
                j = (j + 1) & mask;
            }
        }
        __syncthreads();
        const BJoin& jn = J[I.join];
        const OATab tab{key, pay, occ, mask, I.shift};
        probe_rounds<kBThreads>(tab, src + I.s_off, I.s_cnt, jn.out, jn.out_cap, &state[I.join].cursor, fmt, kbp,
                                wsum, misc + 2);
        __syncthreads();
        if (tid == 0 && atomicSub(&state[I.join].remaining, 1u) == 1u)
            state[I.join].t_done = dev_clock();
    }
}

constexpr size_t kBJoinLds = 32 * 8 + (size_t)(1u << kBSlotsLog2) * 16 + (1u << kBSlotsLog2) / 8;
constexpr size_t kBScatterLds = (size_t)kBScatterTile * 16 + ((1u << kBMaxBits) * 2 + 1 + 16) * 4;
static_assert(2 * kBJoinLds <= kLdsBytes, "two join blocks a CU");
static_assert(kBScatterLds <= kLdsBytes, "scatter LDS budget");

uint32_t size_class(uint64_t n)
{
    uint32_t c = 0;
    while (c + 1 < kClasses && (2ull << c) <= n)
        ++c;
    return c;
}

}  // namespace
}  // namespace pj

using namespace pj;

// ======================================================================
// handle
// ======================================================================

struct pj_batch_s {
    pj_join_desc desc;
    Layout lay;
    std::mutex mu;
    pj_batch_cfg cfg[kClasses];
    bool tuned[kClasses];
    // the prepared batch
    std::vector<pj_batch_join> joins;
    std::vector<BJoin> hj;
    uint32_t G = 0, ntiles = 0, np2 = 0, nhist = 0, npstart = 0;
    uint64_t nrec = 0;
    DevBuf dJ, dtiles, dp2, hist, hsum, pstart, fstart0, fcount0, fstart1, fcount1, fbase, icnt, isum, items,
        state, counters, t0, A, B;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool prepared = false;
    bool enabled = false;
    // sequential mode: the joins run one by one through pj_join_run_device,
    // one join handle per radix-bits value
    // one join handle per (radix bits, lane): the lanes of the concurrent
    // mode each run on a stream of their own, with buffers of their own
    std::map<std::pair<uint32_t, uint32_t>, pj_join_t> seq;
    std::map<std::pair<uint32_t, uint32_t>, std::pair<uint64_t, uint64_t>> seq_warm;   // largest (build, probe) run
    std::vector<hipStream_t> lanes;   // created for the concurrent mode, default priority
    hipGraphExec_t lane_exec = nullptr;     // the concurrent mode captured (PJ_BATCH_CONCURRENT_GRAPH)
    std::vector<uint32_t> lane_graph_bits;  // radix bits of each join in that capture
    std::vector<hipEvent_t> seq_ev;   // two a join, reused across runs
    uint32_t* seq_slots = nullptr;    // pinned: 8 counter words a join
    size_t seq_nslots = 0;
    bool last_sequential = false;
    std::vector<uint64_t> seq_counts;
    std::vector<double> seq_done_us, seq_run_us;
};

namespace {

// The partition plan of a join of nb build tuples under a configuration.
void plan_join(const pj_batch_cfg& c, uint64_t nb, BJoin* j)
{
    j->bits = c.bits;
    j->b2 = c.passes == 2 ? c.bits / 2 : 0;
    j->b1 = c.bits - j->b2;
    // the largest partition expected (mean + 5 sigma + a little), over the load factor
    const double mean = (double)nb / (double)(1u << c.bits);
    const double want = (mean + 5.0 * std::sqrt(mean) + 8.0) / c.load_factor;
    uint32_t sl = 5;
    while (sl < kBSlotsLog2 && (double)(1u << sl) < want)
        ++sl;
    j->slots_log2 = sl;
    j->cap = std::max<uint32_t>(1, (uint32_t)std::floor(c.load_factor * (double)(1u << sl)));
}

// Default configuration of a size class: the fewest bits whose expected
// largest partition fits a 4096-slot table at load factor 0.7.
pj_batch_cfg default_cfg(uint32_t cls)
{
    const double n = std::ldexp(1.5, (int)cls);
    pj_batch_cfg c;
    c.load_factor = 0.7f;
    c.passes = 1;
    c.bits = kBMinBits;
    while (c.bits < kBMaxBits) {
        const double mean = n / (double)(1u << c.bits);
        if ((mean + 5.0 * std::sqrt(mean) + 8.0) / c.load_factor <= (double)(1u << kBSlotsLog2))
            break;
        ++c.bits;
    }
    c.reserved = 0;
    return c;
}

int cfg_valid(const pj_batch_cfg& c)
{
    return c.bits >= kBMinBits && c.bits <= kBMaxBits && (c.passes == 1 || c.passes == 2) && c.load_factor >= 0.5f &&
           c.load_factor <= 0.9f;
}

// Plans the batch (per-join layouts, tiles, buffers) and captures the
// pipeline into a graph.
int prepare(pj_batch_s* h, const std::vector<pj_batch_join>& joins, const std::vector<pj_batch_cfg>& cfgs,
            hipStream_t st)
{
    int rc;
    const uint32_t nj = (uint32_t)joins.size();
    if (h->exec) {
        (void)hipGraphExecDestroy(h->exec);
        h->exec = nullptr;
    }
    if (h->graph) {
        (void)hipGraphDestroy(h->graph);
        h->graph = nullptr;
    }
    h->prepared = false;
    h->joins = joins;
    h->hj.assign(nj, BJoin());
    std::vector<BTile> tiles;
    std::vector<P2Block> p2;
    uint64_t hb = 0, nrec = 0;
    uint32_t pb = 0, fb = 0;
    std::vector<uint32_t> fbase(nj);
    for (uint32_t i = 0; i < nj; ++i) {
        BJoin& j = h->hj[i];
        const pj_batch_join& q = joins[i];
        plan_join(cfgs[i], q.nbuild, &j);
        j.in[0] = (const uint32_t*)q.build;
        j.in[1] = (const uint32_t*)q.probe;
        j.n[0] = q.nbuild;
        j.n[1] = q.nprobe;
        j.out = (uint32_t*)q.out;
        j.out_cap = q.out_cap;
        const uint32_t F1 = 1u << j.b1;
        for (uint32_t side = 0; side < 2; ++side) {
            const uint64_t n = j.n[side];
            j.nt[side] = (uint32_t)std::max<uint64_t>(1, (n + kBTile - 1) / kBTile);
            j.hbase[side] = hb;
            hb += (uint64_t)F1 * j.nt[side];
            j.pbase[side] = pb;
            pb += F1 + 1;
            for (uint32_t t = 0; t < j.nt[side]; ++t) {
                BTile T;
                T.join = i;
                T.side = side;
                T.t = t;
                T.start = (uint64_t)t * kBTile;
                T.count = (uint32_t)std::min<uint64_t>(kBTile, n - std::min<uint64_t>(n, T.start));
                tiles.push_back(T);
            }
            nrec += n;
            if (j.b2)
                for (uint32_t p = 0; p < F1; ++p)
                    p2.push_back(P2Block{i, side, p, 0});
        }
        j.fbase = fb;
        fbase[i] = fb;
        fb += 1u << j.bits;
    }
    if (nrec >= (1ull << 32) || hb >= (1ull << 32))
        return fail(PJ_ERR_UNSUPPORTED, "a batch holds fewer than 2^32 tuples");
    h->G = fb;
    h->ntiles = (uint32_t)tiles.size();
    h->np2 = (uint32_t)p2.size();
    h->nhist = (uint32_t)hb;
    h->npstart = pb;
    h->nrec = nrec;
    const uint32_t hsb = (uint32_t)((hb + kScanSeg - 1) / kScanSeg), isb = (uint32_t)((fb + kScanSeg - 1) / kScanSeg);
    // the largest possible work list: every build partition split by probe chunks and table capacity
    uint64_t maxitems = 0;
    for (uint32_t i = 0; i < nj; ++i) {
        const BJoin& j = h->hj[i];
        maxitems += (uint64_t)(1u << j.bits) + (j.n[0] + j.cap - 1) / j.cap + (j.n[1] + kBChunk - 1) / kBChunk +
                    ((j.n[0] + j.cap - 1) / j.cap) * ((j.n[1] + kBChunk - 1) / kBChunk);
    }
    if ((rc = h->dJ.ensure(nj * sizeof(BJoin))) || (rc = h->dtiles.ensure(tiles.size() * sizeof(BTile))) ||
        (rc = h->dp2.ensure(std::max<size_t>(p2.size(), 1) * sizeof(P2Block))) ||
        (rc = h->hist.ensure(std::max<uint64_t>(hb, 1) * 4)) || (rc = h->hsum.ensure((hsb + 2) * 4)) ||
        (rc = h->pstart.ensure((size_t)pb * 4)) || (rc = h->fstart0.ensure((size_t)fb * 4)) ||
        (rc = h->fcount0.ensure((size_t)fb * 4)) || (rc = h->fstart1.ensure((size_t)fb * 4)) ||
        (rc = h->fcount1.ensure((size_t)fb * 4)) || (rc = h->fbase.ensure((size_t)nj * 4)) ||
        (rc = h->icnt.ensure((size_t)fb * 4)) || (rc = h->isum.ensure((isb + 2) * 4)) ||
        (rc = h->items.ensure(std::max<uint64_t>(maxitems, 1) * sizeof(BItem))) ||
        (rc = h->state.ensure(nj * sizeof(BState))) || (rc = h->counters.ensure(64)) || (rc = h->t0.ensure(64)) ||
        (rc = h->A.ensure(std::max<uint64_t>(nrec, 1) * 16)) || (rc = h->B.ensure(std::max<uint64_t>(nrec, 1) * 16)))
        return rc;
    PJ_HIP(hipMemcpyAsync(h->dJ.p, h->hj.data(), nj * sizeof(BJoin), hipMemcpyHostToDevice, st));
    PJ_HIP(hipMemcpyAsync(h->dtiles.p, tiles.data(), tiles.size() * sizeof(BTile), hipMemcpyHostToDevice, st));
    if (!p2.empty())
        PJ_HIP(hipMemcpyAsync(h->dp2.p, p2.data(), p2.size() * sizeof(P2Block), hipMemcpyHostToDevice, st));
    PJ_HIP(hipMemcpyAsync(h->fbase.p, fbase.data(), nj * 4, hipMemcpyHostToDevice, st));
    PJ_HIP(hipStreamSynchronize(st));   // host arrays above are locals
    PJ_HIP(hipFuncSetAttribute((const void*)k_bscatter, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)kBScatterLds));
    PJ_HIP(hipFuncSetAttribute((const void*)k_bjoin, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kBJoinLds));
    h->prepared = true;
    return PJ_OK;
}

int device_cus()
{
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            cus = prop.multiProcessorCount;
        if (cus <= 0)
            cus = 256;
    }
    return cus;
}

// Enqueues the whole pipeline on st (captured or not).
int enqueue(pj_batch_s* h, hipStream_t st)
{
    const uint32_t nj = (uint32_t)h->joins.size();
    const BJoin* J = (const BJoin*)h->dJ.p;
    BState* state = (BState*)h->state.p;
    uint32_t* counters = (uint32_t*)h->counters.p;
    hipLaunchKernelGGL(k_bstart, dim3((std::max<uint32_t>(nj, 4) + 255) / 256), dim3(256), 0, st, state, nj, counters,
                       (unsigned long long*)h->t0.p);
    const uint32_t maxF1 = 1u << kBMaxBits;
    hipLaunchKernelGGL(k_bhist, dim3(h->ntiles), dim3(kBThreads), maxF1 * 4, st, J, (const BTile*)h->dtiles.p,
                       h->lay.build, h->lay.probe, (uint32_t*)h->hist.p);
    const uint32_t hsb = (uint32_t)((h->nhist + kScanSeg - 1) / kScanSeg);
    uint32_t* hist = (uint32_t*)h->hist.p;
    uint32_t* hsum = (uint32_t*)h->hsum.p;
    hipLaunchKernelGGL(k_scan_reduce, dim3(hsb), dim3(kScanThreads), 0, st, hist, (uint64_t)h->nhist, hsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, hsum, hsb);
    hipLaunchKernelGGL(k_scan_down, dim3(hsb), dim3(kScanThreads), 0, st, hist, (uint64_t)h->nhist,
                       (const uint32_t*)hsum);
    hipLaunchKernelGGL(k_bscatter, dim3(h->ntiles), dim3(kBThreads), kBScatterLds, st, J, (const BTile*)h->dtiles.p,
                       h->lay.build, h->lay.probe, (const uint32_t*)hist, (uint4*)h->A.p);
    uint32_t* fs0 = (uint32_t*)h->fstart0.p;
    uint32_t* fc0 = (uint32_t*)h->fcount0.p;
    uint32_t* fs1 = (uint32_t*)h->fstart1.p;
    uint32_t* fc1 = (uint32_t*)h->fcount1.p;
    hipLaunchKernelGGL(k_bbounds, dim3(nj, 2), dim3(256), 0, st, J, nj, (const uint32_t*)hist,
                       (uint32_t*)h->pstart.p, fs0, fc0, fs1, fc1);
    if (h->np2)
        hipLaunchKernelGGL(k_bpass2, dim3(h->np2), dim3(kBThreads), 0, st, J, (const P2Block*)h->dp2.p,
                           (const uint32_t*)h->pstart.p, (const uint4*)h->A.p, (uint4*)h->B.p, fs0, fc0, fs1, fc1);
    const uint32_t G = h->G;
    uint32_t* icnt = (uint32_t*)h->icnt.p;
    uint32_t* isum = (uint32_t*)h->isum.p;
    hipLaunchKernelGGL(k_bitem_count, dim3((G + 255) / 256), dim3(256), 0, st, J, (const uint32_t*)h->fbase.p, nj, G,
                       (const uint32_t*)fc0, (const uint32_t*)fc1, icnt, state);
    const uint32_t isb = (uint32_t)((G + kScanSeg - 1) / kScanSeg);
    hipLaunchKernelGGL(k_scan_reduce, dim3(isb), dim3(kScanThreads), 0, st, icnt, (uint64_t)G, isum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, isum, isb);
    hipLaunchKernelGGL(k_scan_down, dim3(isb), dim3(kScanThreads), 0, st, icnt, (uint64_t)G, (const uint32_t*)isum);
    hipLaunchKernelGGL(k_bitem_emit, dim3((G + 255) / 256), dim3(256), 0, st, J, (const uint32_t*)h->fbase.p, nj, G,
                       (const uint32_t*)fs0, (const uint32_t*)fc0, (const uint32_t*)fs1, (const uint32_t*)fc1,
                       (const uint32_t*)icnt, (BItem*)h->items.p);
    hipLaunchKernelGGL(k_bjoin, dim3(2 * device_cus()), dim3(kBThreads), kBJoinLds, st, (const uint4*)h->A.p,
                       (const uint4*)h->B.p, (const BItem*)h->items.p, (const uint32_t*)(isum + isb), counters, J,
                       state, h->lay.out);
    PJ_HIP(hipGetLastError());
    return PJ_OK;
}

int capture(pj_batch_s* h, hipStream_t st)
{
    hipStream_t cs;
    PJ_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    int rc = PJ_OK;
    if (e == hipSuccess) {
        rc = enqueue(h, cs);
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(cs, &g);
        if (rc == PJ_OK && e == hipSuccess) {
            h->graph = g;
            e = hipGraphInstantiate(&h->exec, g, nullptr, nullptr, 0);
        }
    }
    (void)hipStreamDestroy(cs);
    (void)st;
    if (rc)
        return rc;
    if (e != hipSuccess)
        return fail(PJ_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
    return PJ_OK;
}

std::vector<pj_batch_cfg> cfgs_for(pj_batch_s* h, const std::vector<pj_batch_join>& joins)
{
    std::vector<pj_batch_cfg> c(joins.size());
    for (size_t i = 0; i < joins.size(); ++i) {
        const uint32_t cls = size_class(joins[i].nbuild);
        c[i] = h->tuned[cls] ? h->cfg[cls] : default_cfg(cls);
    }
    return c;
}

int seq_handle(pj_batch_s* h, uint32_t bits, uint32_t lane, pj_join_t* out)
{
    auto it = h->seq.find({bits, lane});
    if (it != h->seq.end()) {
        *out = it->second;
        return PJ_OK;
    }
    pj_join_desc d = h->desc;
    d.partitioner = PJ_PART_RADIX;
    d.radix_bits = (int32_t)bits;
    d.flags = 0;
    int rc = pj_join_create(&d, out);
    if (rc)
        return rc;
    h->seq[{bits, lane}] = *out;
    return PJ_OK;
}

// Radix bits of a sequential-mode join: its size class's tuned bits, else
// the bits the single-join path picks by itself.
uint32_t seq_bits(const pj_batch_s* h, uint64_t nb)
{
    const uint32_t c = size_class(nb);
    return h->tuned[c] ? h->cfg[c].bits : choose_build_bits(nb, -1);
}

// Drops the captured lane graph. Its nodes hold the addresses of the join
// handles' device buffers, which any other run or the tuner may grow.
void drop_lane_graph(pj_batch_s* h)
{
    if (h->lane_exec) {
        (void)hipGraphExecDestroy(h->lane_exec);
        h->lane_exec = nullptr;
    }
    h->lane_graph_bits.clear();
}

// Sets up the per-join state of the lane modes: lane streams, events, one
// pinned counter slot a join; then every handle (radix bits, lane) runs the
// largest build and probe sides it will meet, so that no buffer grows (and
// waits for the device, or allocates inside a graph capture) later.
int prepare_lanes(pj_batch_s* h, const std::vector<pj_batch_join>& joins, const std::vector<uint32_t>& bits,
                  uint32_t nlanes, hipStream_t st)
{
    const size_t nj = joins.size();
    while (h->lanes.size() < (nlanes > 1 ? nlanes : 0)) {
        hipStream_t ls;
        PJ_HIP(hipStreamCreateWithFlags(&ls, hipStreamNonBlocking));
        h->lanes.push_back(ls);
    }
    while (h->seq_ev.size() < 2 * nj + 1 + nlanes) {
        hipEvent_t e;
        PJ_HIP(hipEventCreate(&e));
        h->seq_ev.push_back(e);
    }
    if (h->seq_nslots < nj) {
        if (h->seq_slots)
            PJ_HIP(hipHostFree(h->seq_slots));
        h->seq_slots = nullptr;
        h->seq_nslots = 0;
        PJ_HIP(hipHostMalloc((void**)&h->seq_slots, nj * 8 * sizeof(uint32_t), hipHostMallocDefault));
        h->seq_nslots = nj;
    }
    int rc;
    std::map<std::pair<uint32_t, uint32_t>, std::pair<size_t, size_t>> big;   // most build, most probe tuples
    for (size_t i = 0; i < nj; ++i) {
        const std::pair<uint32_t, uint32_t> key{bits[i], (uint32_t)(i % nlanes)};
        auto it = big.find(key);
        if (it == big.end()) {
            big[key] = {i, i};
            continue;
        }
        if (joins[i].nbuild > joins[it->second.first].nbuild)
            it->second.first = i;
        if (joins[i].nprobe > joins[it->second.second].nprobe)
            it->second.second = i;
    }
    bool waited = false;
    for (const auto& kv : big) {
        pj_join_t jh;
        if ((rc = seq_handle(h, kv.first.first, kv.first.second, &jh)))
            return rc;
        auto& w = h->seq_warm[kv.first];
        const hipStream_t ls = nlanes > 1 ? h->lanes[kv.first.second] : st;
        for (size_t i : {kv.second.first, kv.second.second}) {
            const pj_batch_join& q = joins[i];
            if (q.nbuild <= w.first && q.nprobe <= w.second)
                continue;
            if (!waited) {   // the inputs may still be in flight on st
                PJ_HIP(hipStreamSynchronize(st));
                waited = true;
            }
            uint64_t n = 0;
            rc = pj_join_run_device(jh, q.build, q.nbuild, q.probe, q.nprobe, q.out, q.out_cap, &n, ls);
            if (rc && rc != PJ_ERR_CAPACITY)
                return rc;
            w.first = std::max<uint64_t>(w.first, q.nbuild);
            w.second = std::max<uint64_t>(w.second, q.nprobe);
        }
    }
    return PJ_OK;
}

// Enqueues every join as a whole radix join (partition both sides, join)
// with `bits` radix bits, nothing waiting in between: with one lane one after
// another on st; with more, join i on lane i % nlanes, each lane a stream of
// its own that starts where st is and that st waits for at the end. Each
// join's counters land in its own pinned slot. With `timed`, HIP events
// around each join give its own device time and its completion time within
// the batch (not inside a graph capture).
int enqueue_lanes(pj_batch_s* h, const std::vector<pj_batch_join>& joins, const std::vector<uint32_t>& bits,
                  uint32_t nlanes, hipStream_t st, bool timed, size_t* ran)
{
    const size_t nj = joins.size();
    hipEvent_t* ev = h->seq_ev.data();
    hipEvent_t* lane_end = ev + 2 * nj + 1;
    PJ_HIP(hipEventRecord(ev[0], st));
    if (nlanes > 1)
        for (uint32_t l = 0; l < nlanes; ++l)
            PJ_HIP(hipStreamWaitEvent(h->lanes[l], ev[0], 0));
    int rc = PJ_OK;
    *ran = 0;
    for (size_t i = 0; i < nj && rc == PJ_OK; ++i) {
        const pj_batch_join& q = joins[i];
        const uint32_t l = (uint32_t)(i % nlanes);
        const hipStream_t ls = nlanes > 1 ? h->lanes[l] : st;
        pj_join_t jh;
        if ((rc = seq_handle(h, bits[i], l, &jh)))
            break;
        if (timed)
            (void)hipEventRecord(ev[2 * i + 1], ls);
        rc = join_enqueue(jh, q.build, q.nbuild, q.probe, q.nprobe, q.out, q.out_cap, h->seq_slots + 8 * i, ls);
        if (timed)
            (void)hipEventRecord(ev[2 * i + 2], ls);
        *ran = i + 1;
    }
    if (nlanes > 1)
        for (uint32_t l = 0; l < nlanes; ++l) {
            PJ_HIP(hipEventRecord(lane_end[l], h->lanes[l]));
            PJ_HIP(hipStreamWaitEvent(st, lane_end[l], 0));
        }
    return rc;
}

// Counts from the slots, and with `timed` the events' times.
void collect_lanes(pj_batch_s* h, size_t nj, size_t ran, bool timed, std::vector<uint64_t>& counts,
                   std::vector<double>& done_us, std::vector<double>& run_us)
{
    counts.assign(nj, 0);
    done_us.assign(nj, 0.0);
    run_us.assign(nj, 0.0);
    hipEvent_t* ev = h->seq_ev.data();
    for (size_t i = 0; i < ran; ++i) {
        const uint32_t* c = h->seq_slots + 8 * i;
        counts[i] = (uint64_t)c[2] | ((uint64_t)c[3] << 32);
        if (!timed)
            continue;
        float ms = 0.f, t = 0.f;
        (void)hipEventElapsedTime(&ms, ev[2 * i + 1], ev[2 * i + 2]);
        (void)hipEventElapsedTime(&t, ev[0], ev[2 * i + 2]);
        run_us[i] = ms * 1e3;
        done_us[i] = t * 1e3;
    }
}

// The sequential and concurrent modes: prepare, enqueue, wait once, collect.
int run_lanes(pj_batch_s* h, const std::vector<pj_batch_join>& joins, const std::vector<uint32_t>& bits,
              uint32_t nlanes, hipStream_t st, std::vector<uint64_t>& counts, std::vector<double>& done_us,
              std::vector<double>& run_us)
{
    drop_lane_graph(h);
    int rc = prepare_lanes(h, joins, bits, nlanes, st);
    if (rc)
        return rc;
    size_t ran = 0;
    rc = enqueue_lanes(h, joins, bits, nlanes, st, true, &ran);
    PJ_HIP(hipStreamSynchronize(st));
    collect_lanes(h, joins.size(), ran, true, counts, done_us, run_us);
    return rc;
}

// The concurrent mode captured into one graph (lanes as parallel branches),
// captured again only when the joins or their radix bits change; one
// hipGraphLaunch runs the batch. No per-join times: done_us is the whole
// batch's time for every join.
int run_lanes_graph(pj_batch_s* h, const std::vector<pj_batch_join>& joins, const std::vector<uint32_t>& bits,
                    hipStream_t st, std::vector<uint64_t>& counts, std::vector<double>& done_us,
                    std::vector<double>& run_us)
{
    int rc;
    if (!h->lane_exec || h->lane_graph_bits != bits) {
        drop_lane_graph(h);
        if ((rc = prepare_lanes(h, joins, bits, kLanes, st)))
            return rc;
        PJ_HIP(hipStreamSynchronize(st));
        hipStream_t cs;
        PJ_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        size_t ran = 0;
        rc = PJ_OK;
        if (e == hipSuccess) {
            rc = enqueue_lanes(h, joins, bits, kLanes, cs, false, &ran);
            hipGraph_t g = nullptr;
            e = hipStreamEndCapture(cs, &g);
            if (rc == PJ_OK && e == hipSuccess)
                e = hipGraphInstantiate(&h->lane_exec, g, nullptr, nullptr, 0);
            if (g)
                (void)hipGraphDestroy(g);
        }
        (void)hipStreamDestroy(cs);
        if (rc)
            return rc;
        if (e != hipSuccess) {
            h->lane_exec = nullptr;
            return fail(PJ_ERR_DEVICE, "graph capture of the batch failed: %s", hipGetErrorString(e));
        }
        if (ran != joins.size())
            return fail(PJ_ERR_DEVICE, "graph capture stopped after %zu joins", ran);
        h->lane_graph_bits = bits;
    }
    hipEvent_t* ev = h->seq_ev.data();
    PJ_HIP(hipEventRecord(ev[0], st));
    PJ_HIP(hipGraphLaunch(h->lane_exec, st));
    PJ_HIP(hipEventRecord(ev[1], st));
    PJ_HIP(hipStreamSynchronize(st));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
    collect_lanes(h, joins.size(), joins.size(), false, counts, done_us, run_us);
    for (auto& d : done_us)
        d = ms * 1e3;
    return PJ_OK;
}

}  // namespace

extern "C" {

int pj_batch_create(const pj_join_desc* desc, pj_batch_t* out)
{
    if (!desc || !out)
        return fail(PJ_ERR_INVALID, "null argument");
    *out = nullptr;
    Layout L;
    int rc = derive_layout(desc, &L);
    if (rc)
        return rc;
    if (L.rowid)
        return fail(PJ_ERR_UNSUPPORTED,
                    "batched joins take one INTEGER / LONG / DATE key or two INTEGER keys, at most 8 payload bytes "
                    "a side and output tuples of at most 32 bytes without CHAR columns");
    pj_batch_s* h = new (std::nothrow) pj_batch_s();
    if (!h)
        return fail(PJ_ERR_INVALID, "out of host memory");
    h->desc = *desc;
    h->lay = L;
    for (uint32_t c = 0; c < kClasses; ++c) {
        h->cfg[c] = default_cfg(c);
        h->tuned[c] = false;
    }
    *out = h;
    return PJ_OK;
}

int pj_batch_destroy(pj_batch_t h)
{
    if (!h)
        return PJ_OK;
    for (auto& kv : h->seq)
        pj_join_destroy(kv.second);
    if (h->lane_exec)
        (void)hipGraphExecDestroy(h->lane_exec);
    for (hipStream_t ls : h->lanes)
        (void)hipStreamDestroy(ls);
    for (hipEvent_t e : h->seq_ev)
        (void)hipEventDestroy(e);
    if (h->seq_slots)
        (void)hipHostFree(h->seq_slots);
    if (h->exec)
        (void)hipGraphExecDestroy(h->exec);
    if (h->graph)
        (void)hipGraphDestroy(h->graph);
    delete h;
    return PJ_OK;
}

int pj_batch_set_config(pj_batch_t h, uint32_t size_class, const pj_batch_cfg* cfg)
{
    if (!h || !cfg || size_class >= kClasses || !cfg_valid(*cfg))
        return fail(PJ_ERR_INVALID, "bad size class or configuration");
    std::lock_guard<std::mutex> g(h->mu);
    h->cfg[size_class] = *cfg;
    h->tuned[size_class] = true;
    return PJ_OK;
}

int pj_batch_get_config(pj_batch_t h, uint32_t size_class, pj_batch_cfg* cfg)
{
    if (!h || !cfg || size_class >= kClasses)
        return fail(PJ_ERR_INVALID, "bad size class");
    *cfg = h->cfg[size_class];
    return PJ_OK;
}

int pj_batch_prepare(pj_batch_t h, uint32_t njoins, const pj_batch_join* joins, void* stream)
{
    if (!h || (njoins && !joins) || njoins == 0)
        return fail(PJ_ERR_INVALID, "need at least one join");
    for (uint32_t i = 0; i < njoins; ++i) {
        const pj_batch_join& q = joins[i];
        if ((q.nbuild && !q.build) || (q.nprobe && !q.probe) || (q.out_cap && !q.out) || ((uintptr_t)q.out % 16))
            return fail(PJ_ERR_INVALID, "join %u: null or misaligned buffer", i);
    }
    std::lock_guard<std::mutex> g(h->mu);
    hipStream_t st = (hipStream_t)stream;
    std::vector<pj_batch_join> js(joins, joins + njoins);
    h->last_sequential = false;
    h->seq_counts.clear();
    drop_lane_graph(h);   // captured for the previous list
    if (!h->enabled) {   // only the lane modes run: keep the list
        h->joins = js;
        h->prepared = true;
        return PJ_OK;
    }
    int rc = prepare(h, js, cfgs_for(h, js), st);
    if (rc)
        return rc;
    return capture(h, st);
}

int pj_batch_run(pj_batch_t h, int mode, void* stream)
{
    if (!h || mode < PJ_BATCH_PLAIN || mode > PJ_BATCH_CONCURRENT_GRAPH)
        return fail(PJ_ERR_INVALID, "null handle or unknown mode");
    std::lock_guard<std::mutex> g(h->mu);
    if (!h->prepared)
        return fail(PJ_ERR_STATE, "pj_batch_run before pj_batch_prepare");
    hipStream_t st = (hipStream_t)stream;
    if (mode == PJ_BATCH_SEQUENTIAL || mode == PJ_BATCH_CONCURRENT || mode == PJ_BATCH_CONCURRENT_GRAPH) {
        std::vector<uint32_t> bits(h->joins.size());
        for (size_t i = 0; i < bits.size(); ++i)
            bits[i] = seq_bits(h, h->joins[i].nbuild);
        const uint32_t nlanes = mode == PJ_BATCH_SEQUENTIAL ? 1u : kLanes;
        const int rc = mode == PJ_BATCH_CONCURRENT_GRAPH
                           ? run_lanes_graph(h, h->joins, bits, st, h->seq_counts, h->seq_done_us, h->seq_run_us)
                           : run_lanes(h, h->joins, bits, nlanes, st, h->seq_counts, h->seq_done_us, h->seq_run_us);
        h->last_sequential = rc == PJ_OK;
        return rc;
    }
    // The first device run of the batched pipeline (graph replay of a
    // 40-join batch) stopped making progress and its cause has not been
    // found; the pipeline stays off until it is.
    if (!h->enabled)
        return fail(PJ_ERR_UNSUPPORTED, "the batched pipeline is disabled: its first device run hung, cause not "
                                         "found (PJ_BATCH_SEQUENTIAL runs)");
    h->last_sequential = false;
    if (mode == PJ_BATCH_GRAPH) {
        if (!h->exec)
            return fail(PJ_ERR_STATE, "no captured graph");
        PJ_HIP(hipGraphLaunch(h->exec, st));
        return PJ_OK;
    }
    return enqueue(h, st);
}

int pj_batch_results(pj_batch_t h, uint64_t* out_counts, double* done_us, void* stream)
{
    if (!h || !out_counts)
        return fail(PJ_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(h->mu);
    if (!h->prepared)
        return fail(PJ_ERR_STATE, "no prepared batch");
    hipStream_t st = (hipStream_t)stream;
    const uint32_t nj = (uint32_t)h->joins.size();
    if (h->last_sequential || !h->enabled) {
        if (!h->last_sequential || h->seq_counts.size() != nj)
            return fail(PJ_ERR_STATE, "no completed run");
        int rc = PJ_OK;
        for (uint32_t i = 0; i < nj; ++i) {
            out_counts[i] = h->seq_counts[i];
            if (done_us)
                done_us[i] = h->seq_done_us[i];
            if (h->seq_counts[i] > h->joins[i].out_cap && rc == PJ_OK)
                rc = fail(PJ_ERR_CAPACITY, "join %u produced %llu rows, its output holds %llu", i,
                          (unsigned long long)h->seq_counts[i], (unsigned long long)h->joins[i].out_cap);
        }
        return rc;
    }
    std::vector<BState> s(nj);
    unsigned long long t0 = 0;
    PJ_HIP(hipMemcpyAsync(s.data(), h->state.p, nj * sizeof(BState), hipMemcpyDeviceToHost, st));
    PJ_HIP(hipMemcpyAsync(&t0, h->t0.p, 8, hipMemcpyDeviceToHost, st));
    PJ_HIP(hipStreamSynchronize(st));
    int rc = PJ_OK;
    for (uint32_t i = 0; i < nj; ++i) {
        out_counts[i] = s[i].cursor;
        if (done_us)   // the device clock of s_memrealtime runs at 100 MHz
            done_us[i] = s[i].t_done ? (double)(s[i].t_done - t0) / 100.0 : 0.0;
        if (s[i].cursor > h->joins[i].out_cap && rc == PJ_OK)
            rc = fail(PJ_ERR_CAPACITY, "join %u produced %llu rows, its output holds %llu", i,
                      (unsigned long long)s[i].cursor, (unsigned long long)h->joins[i].out_cap);
    }
    return rc;
}

int pj_batch_run_times(pj_batch_t h, double* run_us)
{
    if (!h || !run_us)
        return fail(PJ_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(h->mu);
    if (!h->last_sequential || h->seq_run_us.size() != h->joins.size())
        return fail(PJ_ERR_STATE, "no completed sequential or concurrent run");
    for (size_t i = 0; i < h->seq_run_us.size(); ++i)
        run_us[i] = h->seq_run_us[i];
    return PJ_OK;
}

// Times each candidate configuration of every size class present in the
// batch on up to `trial` of its joins (the batch's own buffers) and keeps
// the fastest.
int pj_batch_tune(pj_batch_t h, uint32_t njoins, const pj_batch_join* joins, uint32_t trial, void* stream)
{
    if (!h || !joins || njoins == 0 || trial == 0)
        return fail(PJ_ERR_INVALID, "bad arguments");
    hipStream_t st = (hipStream_t)stream;
    std::lock_guard<std::mutex> g(h->mu);
    drop_lane_graph(h);   // the trial runs may grow the handles' buffers
    std::vector<std::vector<pj_batch_join>> by(kClasses);
    for (uint32_t i = 0; i < njoins; ++i) {
        const uint32_t c = size_class(joins[i].nbuild);
        if (by[c].size() < trial)
            by[c].push_back(joins[i]);
    }
    if (!h->enabled) {
        // the pipeline is off: tune the radix bits of the sequential mode,
        // around the bits the single-join path picks by itself
        for (uint32_t c = 0; c < kClasses; ++c) {
            if (by[c].empty())
                continue;
            uint64_t nmax = 0;
            for (const auto& q : by[c])
                nmax = std::max<uint64_t>(nmax, q.nbuild);
            const uint32_t b0 = std::max<uint32_t>(kBMinBits, std::min<uint32_t>(kBMaxBits, choose_build_bits(nmax, -1)));
            double best = 1e300;
            pj_batch_cfg bestc = default_cfg(c);
            for (uint32_t b = b0; b <= std::min<uint32_t>(kBMaxBits, b0 + 3); ++b) {
                std::vector<uint32_t> bits(by[c].size(), b);
                std::vector<uint64_t> cnt;
                std::vector<double> done, run;
                double t = 1e300;
                for (int rep = 0; rep < 3; ++rep) {   // the first run warms up
                    int rc = run_lanes(h, by[c], bits, 1, st, cnt, done, run);
                    if (rc)
                        return rc;
                    if (rep)
                        t = std::min(t, done.back());
                }
                if (t < best) {
                    best = t;
                    bestc.bits = (uint8_t)b;
                    bestc.passes = 1;
                }
            }
            h->cfg[c] = bestc;
            h->tuned[c] = true;
        }
        return PJ_OK;
    }
    const float lfs[3] = {0.5f, 0.7f, 0.9f};
    for (uint32_t c = 0; c < kClasses; ++c) {
        if (by[c].empty())
            continue;
        uint64_t nmax = 0;
        for (const auto& q : by[c])
            nmax = std::max<uint64_t>(nmax, q.nbuild);
        double best = 1e300;
        pj_batch_cfg bestc = default_cfg(c);
        for (float lf : lfs) {
            // the fewest bits that fit, and two more
            uint32_t b0 = kBMinBits;
            while (b0 < kBMaxBits) {
                const double mean = (double)nmax / (double)(1u << b0);
                if ((mean + 5.0 * std::sqrt(mean) + 8.0) / lf <= (double)(1u << kBSlotsLog2))
                    break;
                ++b0;
            }
            for (uint32_t b = b0; b <= std::min<uint32_t>(b0 + 2, kBMaxBits); ++b)
                for (uint32_t passes = 1; passes <= (b >= 8 ? 2u : 1u); ++passes) {
                    pj_batch_cfg cand;
                    cand.bits = (uint8_t)b;
                    cand.passes = (uint8_t)passes;
                    cand.load_factor = lf;
                    cand.reserved = 0;
                    std::vector<pj_batch_cfg> cc(by[c].size(), cand);
                    int rc = prepare(h, by[c], cc, st);
                    if (rc)
                        return rc;
                    double t = 1e300;
                    for (int rep = 0; rep < 3; ++rep) {   // the first run warms up
                        PJ_HIP(hipStreamSynchronize(st));
                        const auto a = std::chrono::steady_clock::now();
                        if ((rc = enqueue(h, st)))
                            return rc;
                        PJ_HIP(hipStreamSynchronize(st));
                        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
                        if (rep)
                            t = std::min(t, dt);
                    }
                    if (t < best) {
                        best = t;
                        bestc = cand;
                    }
                }
        }
        h->cfg[c] = bestc;
        h->tuned[c] = true;
    }
    h->prepared = false;   // the trial batches are gone: prepare the real one again
    return PJ_OK;
}

}  // extern "C"
