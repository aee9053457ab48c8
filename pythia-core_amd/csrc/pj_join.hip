// pj_join.hip -- the partitioned hash join for gfx950.
//
// What it replaces in the reference (paths relative to its tree):
//   HashJoinOp::scanStart / buildFromPage  operators/join.cpp:548-637, 838-884
//   HashJoinOp::getNext / constructOutput  operators/join.cpp:648-757
//   HashTable allocate / iterator          util/hashtable.cpp:254-297, util/hashtable.h:126-171
//   ConjunctionEqualsEvaluator on keys     conjunctionevaluator.h:65-73, rawcompfns.cpp:48-180
// The reference shares one chained hash table among T threads. Here both
// relations are partitioned (histogram, exclusive scan, LDS-buffered
// scatter), either on the top bits of a 64-bit mix of the key (radix) or on
// ranges of the key chosen by the learned CDF model (pj_learned.hip). Every
// build partition is loaded into an LDS-resident table and the probe
// partition streams past it; a whole workgroup reserves output space with one
// atomic per round and lanes write adjacent output tuples.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

#include "pj_device.h"
#include "pj_internal.h"
#include "pj_probe.h"
#include "pj_scan.h"

namespace pj {

// ======================================================================
// partition pass: histogram, scan, scatter
// ======================================================================
//
// A partition's records from one block form a segment padded to a whole
// number of B-record chunks, so the write-combining scatter below only ever
// writes whole, aligned chunks. Padding records carry a key of another
// partition; the join drops any record outside the work item's partition.

constexpr int kHistThreads = 1024;
constexpr int kHistItems = 8;

// hist[p * nblk + blk] = tuples of block blk's chunk in partition p, rounded
// up to a multiple of pad. RANGE: partitions of the learned function (its
// build partition divided by pf.div); with pmin/pmax (a build side) it also
// lowers pmin[p] / raises pmax[p] to the partition's ordered keys, and it
// stores every tuple's partition in pid for the scatter.
template <int RM, bool RANGE>
__global__ __launch_bounds__(kHistThreads) void k_hist(const uint32_t* __restrict__ in, uint64_t n,
                                                        SideLayout L, PartFn pf, uint64_t chunk,
                                                        uint32_t nblk, uint32_t pad, uint32_t* __restrict__ hist,
                                                        unsigned long long* __restrict__ pmin,
                                                        unsigned long long* __restrict__ pmax,
                                                        uint16_t* __restrict__ pid)
{
    extern __shared__ uint64_t smh[];
    const uint32_t F = pf.F;
    const bool mm = RANGE && pmin != nullptr;
    uint8_t* mapl = reinterpret_cast<uint8_t*>(smh);                                          // RANGE: kMapBytes
    unsigned long long* lmin = reinterpret_cast<unsigned long long*>(mapl + (RANGE ? kMapBytes : 0));
    unsigned long long* lmax = lmin + (mm ? F : 0);
    uint32_t* cnt = reinterpret_cast<uint32_t*>(lmax + (mm ? F : 0));
    MapView mv = {nullptr};
    if (RANGE)
        mv = load_map(mapl, pf.map);
    for (uint32_t p = threadIdx.x; p < F; p += kHistThreads) {
        cnt[p] = 0;
        if (mm) {
            lmin[p] = ~0ull;
            lmax[p] = 0ull;
        }
    }
    __syncthreads();
    const uint64_t beg = (uint64_t)blockIdx.x * chunk;
    const uint64_t end = beg + chunk < n ? beg + chunk : n;
    for (uint64_t base = beg; base < end; base += (uint64_t)kHistThreads * kHistItems) {
        uint4 r[kHistItems];
#pragma unroll
        for (int k = 0; k < kHistItems; ++k) {
            uint64_t i = base + (uint64_t)k * kHistThreads + threadIdx.x;
            r[k] = make_uint4(0, 0, 0, 0);
            if (i < end)
                r[k] = load_rec<RM>(in, i, L);
        }
        uint32_t part[kHistItems];
        if (RANGE) {
#pragma unroll
            for (int k = 0; k < kHistItems; ++k)   // records past the end are looked up harmlessly
                part[k] = coarse_of_fine(mv.fine(ord_key(rec_key(r[k]))), pf);
        }
#pragma unroll
        for (int k = 0; k < kHistItems; ++k) {
            uint64_t i = base + (uint64_t)k * kHistThreads + threadIdx.x;
            if (i < end) {
                const uint64_t key = rec_key(r[k]);
                const uint32_t p = RANGE ? part[k] : part_of(key, pf);
                atomicAdd(&cnt[p], 1u);
                if (RANGE)
                    pid[i] = (uint16_t)p;
                if (mm) {
                    atomicMin(&lmin[p], (unsigned long long)ord_key(key));
                    atomicMax(&lmax[p], (unsigned long long)ord_key(key));
                }
            }
        }
    }
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < F; p += kHistThreads) {
        hist[(uint64_t)p * nblk + blockIdx.x] = (cnt[p] + pad - 1) / pad * pad;
        if (mm && cnt[p]) {
            atomicMin(&pmin[p], lmin[p]);
            atomicMax(&pmax[p], lmax[p]);
        }
    }
}

// a[0..n) = v. A kernel rather than hipMemsetAsync, so that streams a
// caller captures into a graph hold no memset nodes.
__global__ void k_fill_u32(uint32_t* __restrict__ a, uint32_t n, uint32_t v)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        a[i] = v;
}

// part_start[p] = first record of partition p; part_start[F] = total.
__global__ void k_part_bounds(const uint32_t* __restrict__ offs, uint32_t F, uint32_t nblk,
                              const uint32_t* __restrict__ total, uint32_t* __restrict__ part_start)
{
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < F)
        part_start[p] = offs[(uint64_t)p * nblk];
    if (p == F)
        part_start[F] = *total;
}

constexpr int kWcThreads = 1024;
// Appended 2048-way scatter: a round of a block whose partitions fill fewer
// whole chunks of overflow than this (uniform keys: about five a round)
// stores them nontemporal.
constexpr uint32_t kWcHotChunks = 16;

// Software write-combining scatter: every partition owns a B-record buffer
// in LDS and leaves HBM only in aligned chunks of B records. A round claims
// buffer slots with one LDS atomic per record; a claim index s >= B ranks the
// record among the partition's overflow. After the claims, a full buffer
// leaves as one chunk (a B-lane group writes it), whole chunks of overflow
// records are written straight after it, and the overflow remainder refills
// the buffer. Three barriers a round, whatever the skew.
//
// SPLIT = 2 halves the LDS a block needs, so that 2048 partitions still get
// 8-record (128-byte, whole cache line) chunks: blocks b and b + 8 share
// input chunk `part` (under round-robin dispatch both run on the same XCD, so
// the second read of a tile mostly hits its L2) and block half h buffers and
// writes only its half of the partitions: [h * FL, (h + 1) * FL) for hash
// partitions, the partitions p with p % 2 == h for range partitions. On
// MI355X a 2048-way scatter of 256M 16-byte records takes 2.5 ms with 64-byte
// chunks and 1.8-1.9 ms split with 128-byte chunks. Alternate range
// partitions took 5.3 ms for a 2044-way scatter of 512M clustered records
// where halves of the key order took 6.1 ms (hash partitions: 4.6 ms in
// halves, 5.2 ms alternating).
//
// APP (the appended layout, radix partitions): no histogram pass before it.
// Slot k of block b is chunk k * gridDim.x + b of out: the blocks append
// their chunks side by side at about the same pace. Every chunk a block
// writes takes its next slot and is tagged with its partition in
// app.cmap[b * app.region + k] (region: the input chunk's records in whole
// chunks plus one partial chunk a partition). Lane l of wave w owns
// partition w * 64 + l: it claims the slots of its partition's chunks each
// round (a wave-wide scan of the owners' counts, one LDS atomic a wave) and
// keeps the round's first slot and overflow records in LDS, from which the
// records that leave as overflow or refill the buffer find their place. At
// the end the block stores its chunks a partition, app.bhist[p * gridDim.x
// + b], and its slots used, app.nchunks[b]; k_chunk_place then lists the
// chunks partition by partition.
struct AppendOut {
    uint16_t* cmap;      // partition of each chunk slot
    uint32_t* bhist;     // F x gridDim.x chunk counts, partition-major
    uint32_t* nchunks;   // slots each block used
    uint64_t region;     // chunk slots a block
};

template <int RM, int B, bool RANGE, int SPLIT, bool APP>
__global__ __launch_bounds__(kWcThreads) void k_scatter_wc(const uint32_t* __restrict__ in, uint64_t n,
                                                            SideLayout L, PartFn pf, uint64_t chunk,
                                                            uint32_t nblk, const uint32_t* __restrict__ offs,
                                                            uint4* __restrict__ out, uint4 pad_other, uint4 pad_zero,
                                                            const uint16_t* __restrict__ pid, AppendOut app)
{
    static_assert(SPLIT == 1 || SPLIT == 2, "one or two blocks an input chunk");
    static_assert(!APP || (!RANGE && 64 % B == 0 && kWcThreads == 1024), "appended: radix, a lane a partition");
    // (APP: at most 1024 partitions a block, so a partition fits 11 bits)
    constexpr int IT = SPLIT == 1 ? 4 : 8;   // records a thread per round
    extern __shared__ uint4 sm4[];
    const uint32_t F = pf.F;
    const uint32_t FL = SPLIT == 1 ? F : (F + 1) / 2;   // partitions this block buffers
    const uint32_t part = SPLIT == 1 ? blockIdx.x : (blockIdx.x / 16) * 8 + blockIdx.x % 8;
    const uint32_t half = SPLIT == 1 ? 0u : (blockIdx.x / 8) & 1u;
    constexpr bool ALT = SPLIT == 2 && RANGE;   // alternate partitions: local q is global 2q + half
    const uint32_t p0 = ALT ? 0u : half * FL;
    const uint32_t pn = SPLIT == 1 ? F : ALT ? (half ? F / 2 : (F + 1) / 2) : (half ? F - FL : FL);
    auto global_part = [&](uint32_t q) { return ALT ? 2 * q + half : p0 + q; };
    uint4* buf = sm4;                                                          // FL * B
    uint32_t* cnt = reinterpret_cast<uint32_t*>(buf + FL * B);                 // FL: buffered records
    uint32_t* goff = cnt + FL;   // FL: next chunk's record index (APP: the round's first chunk slot)
    uint32_t* ndir = goff + FL;  // APP, FL: the round's overflow records that leave in whole chunks
    uint32_t* ccur = ndir + FL;  // APP: chunk slots the block has used
    uint32_t* hot = ccur + 1;    // APP: the round's whole chunks of overflow (all partitions)
    uint32_t nch = 0;            // APP: chunks written of the thread's own partition
    const uint64_t rbase = APP ? (uint64_t)blockIdx.x * app.region : 0;   // APP: the block's first tag
    auto slot_rec = [&](uint64_t slot) -> uint64_t { return (rbase + slot) * B; };
    for (uint32_t p = threadIdx.x; p < pn; p += kWcThreads) {
        if (!APP)
            goff[p] = offs[(uint64_t)global_part(p) * nblk + part];
        cnt[p] = 0;
    }
    if (APP && threadIdx.x == 0) {
        *ccur = 0;
        *hot = 0;
    }
    __syncthreads();
    const uint64_t beg = (uint64_t)part * chunk;
    const uint64_t end = beg + chunk < n ? beg + chunk : n;
    const uint32_t gl = threadIdx.x % B, grp = threadIdx.x / B;
    constexpr uint32_t ngrp = kWcThreads / B;
    // APP: thread t owns partition t of the block (t < pn). own_slots(n),
    // called by every lane of a wave with its own partition's chunk count,
    // claims the wave's slots with one LDS atomic and returns the lane's
    // first slot (the wave's lanes get consecutive runs, so the owners' tag
    // stores coalesce); group g of the wave then writes the wave's partitions
    // g, g + 64 / B, ... with the owners' slots and counts.
    const uint32_t wl = threadIdx.x & 63u, wq0 = threadIdx.x & ~63u, gi = wl / B;
    // The wave's exclusive scan of the counts runs bit by bit over ballots
    // (a few bits: counts are small), without LDS round trips.
    const uint64_t below = wl ? ~0ull >> (64 - wl) : 0ull;
    auto own_slots = [&](uint32_t nq) -> uint32_t {
        uint32_t pre = 0, tot = 0;
        for (uint32_t b = 0; __ballot(nq >> b) != 0; ++b) {
            const uint64_t m = __ballot((nq >> b) & 1u);
            pre += (uint32_t)__popcll(m & below) << b;
            tot += (uint32_t)__popcll(m) << b;
        }
        uint32_t base = 0;
        if (wl == 0 && tot)
            base = atomicAdd(ccur, tot);
        return __builtin_amdgcn_readfirstlane(base) + pre;
    };
    const bool have_pid = RANGE;   // range partitions come from the histogram pass
    // A split pair reads its input chunk twice: ordinary loads keep the
    // lines in the XCD's L2 for the partner, where streaming loads would
    // mark them to be evicted first.
    auto ld = [&](uint64_t i) -> uint4 {
        if (RM == REC_IDENT && SPLIT == 2)
            return reinterpret_cast<const uint4*>(in)[i];
        return load_rec<RM>(in, i, L);
    };
    constexpr uint64_t R = (uint64_t)kWcThreads * IT;   // records a round
    // One round over the records in cur (loaded a round ahead), loading the
    // next round's into nxt. FULL: every record of this round and of the
    // next lies inside the chunk, so no record needs a bounds check.
    auto round = [&](uint64_t t0, uint4(&cur)[IT], uint32_t(&curp)[IT], uint4(&nxt)[IT], uint32_t(&nxtp)[IT],
                     auto full_tag) {
        constexpr bool FULL = decltype(full_tag)::value;
        uint32_t p[IT], s[IT];
        bool v[IT];
#pragma unroll
        for (int k = 0; k < IT; ++k) {
            const uint64_t i = t0 + (uint64_t)k * kWcThreads + threadIdx.x;
            p[k] = 0;
            v[k] = false;
            if (FULL || i < end) {   // partition (a split block keeps only its half's partitions)
                const uint32_t pp = have_pid ? curp[k] : part_of(rec_key(cur[k]), pf);
                v[k] = SPLIT == 1 || (ALT ? (pp & 1u) == half : (pp >= FL) == (half != 0));
                p[k] = ALT ? pp >> 1 : pp - p0;
            }
            const uint64_t i2 = i + R;
            if (FULL || i2 < end) {
                nxt[k] = ld(i2);
                if (have_pid)
                    nxtp[k] = pid[i2];
            }
        }
        // claim: the round's LDS atomics leave back to back, the buffer
        // writes wait for their results after. (Range partitions, whose
        // loop also carries the partition ids, claim and write record by
        // record: the batched form spills registers there.)
        if (RANGE) {
#pragma unroll
            for (int k = 0; k < IT; ++k)
                if (v[k]) {
                    s[k] = atomicAdd(&cnt[p[k]], 1u);
                    if (s[k] < (uint32_t)B)
                        buf[p[k] * B + s[k]] = cur[k];
                }
        } else {
#pragma unroll
            for (int k = 0; k < IT; ++k)
                if (v[k])
                    s[k] = atomicAdd(&cnt[p[k]], 1u);
#pragma unroll
            for (int k = 0; k < IT; ++k)
                if (v[k] && s[k] < (uint32_t)B)
                    buf[p[k] * B + s[k]] = cur[k];
        }
        // APP: record k's partition and claim in one word (~0: none), which
        // frees the registers of both for the slot claims below
        uint32_t ps[IT];
        if (APP) {
#pragma unroll
            for (int k = 0; k < IT; ++k)
                ps[k] = v[k] ? p[k] | (s[k] << 11) : ~0u;
            // (the previous round's reads of the count ended at its last barrier)
            if (SPLIT == 2 && threadIdx.x == 0)
                *hot = 0;
        }
        __syncthreads();
        // full buffers leave; whole chunks of overflow go straight out
        if (APP) {
            // the owner of each partition claims its chunks' slots (the
            // buffer's chunk, then the overflow's), keeps them in LDS and
            // tags them
            {
                const uint32_t qo = threadIdx.x;
                const uint32_t co = qo < pn ? cnt[qo] : 0u;
                const uint32_t no = co >= (uint32_t)B ? 1 + (co - B) / B : 0u;
                const uint32_t so = own_slots(no);
                if (no) {
                    goff[qo] = so;
                    ndir[qo] = (no - 1) * B;
                    nch += no;
                }
                if (SPLIT == 2 && no > 1)
                    atomicAdd(hot, no - 1);
                for (uint32_t t = 0; t < no; ++t)
                    app.cmap[rbase + so + t] = (uint16_t)global_part(qo);
            }
            // the groups read the slots their wave's owners wrote to LDS
            // (one wave's LDS accesses complete in order)
            __builtin_amdgcn_wave_barrier();
            for (uint32_t j = gi; j < 64; j += 64 / B) {
                const uint32_t q = wq0 + j;
                const uint32_t c = q < pn ? cnt[q] : 0u;
                if (c >= (uint32_t)B)
                    out[slot_rec(goff[q]) + gl] = buf[q * B + gl];
            }
            __syncthreads();
            // overflow records: whole chunks leave at their slots, the rest
            // refill the buffer. Two plain stores of one value in one
            // if / else merge into a flat store to either address, so: where
            // few whole chunks leave (uniform keys: a few a round, which a
            // 2048-way split block writes nontemporal, keeping the stores
            // apart and its partner's input lines in L2) one loop stores
            // them nontemporal; where many leave (skewed keys, or fewer
            // partitions: nontemporal stores of many partial lines run up to
            // three times slower) a loop of their own stores them before the
            // refill.
            if (SPLIT == 2 && *hot < kWcHotChunks) {
#pragma unroll
                for (int k = 0; k < IT; ++k)
                    if (ps[k] != ~0u && (ps[k] >> 11) >= (uint32_t)B) {
                        const uint32_t pk = ps[k] & 0x7FFu, rank = (ps[k] >> 11) - B, direct = ndir[pk];
                        if (rank < direct) {
                            const u32x4 r = {cur[k].x, cur[k].y, cur[k].z, cur[k].w};
                            __builtin_nontemporal_store(
                                r, reinterpret_cast<u32x4*>(out + slot_rec(goff[pk] + 1 + rank / B) + rank % B));
                        } else {
                            buf[pk * B + rank - direct] = cur[k];
                        }
                    }
            } else {
#pragma unroll
                for (int k = 0; k < IT; ++k)
                    if (ps[k] != ~0u && (ps[k] >> 11) >= (uint32_t)B) {
                        const uint32_t pk = ps[k] & 0x7FFu, rank = (ps[k] >> 11) - B;
                        if (rank < ndir[pk])
                            out[slot_rec(goff[pk] + 1 + rank / B) + rank % B] = cur[k];
                    }
#pragma unroll
                for (int k = 0; k < IT; ++k)
                    if (ps[k] != ~0u && (ps[k] >> 11) >= (uint32_t)B) {
                        const uint32_t pk = ps[k] & 0x7FFu, rank = (ps[k] >> 11) - B, direct = ndir[pk];
                        if (rank >= direct)
                            buf[pk * B + rank - direct] = cur[k];
                    }
            }
        } else {
            for (uint32_t q = grp; q < pn; q += ngrp)
                if (cnt[q] >= (uint32_t)B)
                    out[goff[q] + gl] = buf[q * B + gl];
            uint32_t keep[IT];
#pragma unroll
            for (int k = 0; k < IT; ++k) {
                keep[k] = ~0u;
                if (v[k] && s[k] >= (uint32_t)B) {
                    const uint32_t rank = s[k] - B;
                    const uint32_t direct = (cnt[p[k]] - B) / B * B;   // overflow records in whole chunks
                    if (rank < direct)
                        out[goff[p[k]] + B + rank] = cur[k];
                    else
                        keep[k] = rank - direct;
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < IT; ++k)
                if (keep[k] != ~0u)
                    buf[p[k] * B + keep[k]] = cur[k];
        }
        for (uint32_t q = threadIdx.x; q < pn; q += kWcThreads) {
            const uint32_t c = cnt[q];
            if (c >= (uint32_t)B) {
                const uint32_t direct = (c - B) / B * B;
                if (!APP)
                    goff[q] += B + direct;
                cnt[q] = c - B - direct;
            }
        }
        __syncthreads();
    };
    uint4 ra[IT], rb[IT];   // records of the current and the next round
    uint32_t pa[IT] = {}, pb[IT] = {};
#pragma unroll
    for (int k = 0; k < IT; ++k) {
        const uint64_t i = beg + (uint64_t)k * kWcThreads + threadIdx.x;
        if (i < end) {
            ra[k] = ld(i);
            if (have_pid)
                pa[k] = pid[i];
        }
    }
    uint64_t t0 = beg;
    // Full rounds without bounds checks: two an iteration with the record
    // registers taking turns (no copies), or, with the partition ids as well
    // (range partitions), one a round with a copy, which needs fewer
    // registers. The chunk's last rounds with bounds checks.
    if (!have_pid && !APP) {
        for (; t0 + 3 * R <= end; t0 += 2 * R) {
            round(t0, ra, pa, rb, pb, std::true_type{});
            round(t0 + R, rb, pb, ra, pa, std::true_type{});
        }
    } else {
        for (; t0 + 2 * R <= end; t0 += R) {
            round(t0, ra, pa, rb, pb, std::true_type{});
#pragma unroll
            for (int k = 0; k < IT; ++k) {
                ra[k] = rb[k];
                pa[k] = pb[k];
            }
        }
    }
    for (; t0 < end; t0 += R) {
        round(t0, ra, pa, rb, pb, std::false_type{});
#pragma unroll
        for (int k = 0; k < IT; ++k) {
            ra[k] = rb[k];
            pa[k] = pb[k];
        }
    }
    // partial buffers leave padded to a whole chunk. The buffer slot is read
    // whatever c is (gl < B keeps it inside the buffer) and the padding record
    // chosen word by word: a select between an LDS record and an argument
    // record compiles to a flat load from either address, with the argument
    // copied to scratch memory.
    auto tail = [&](uint32_t q, uint64_t at) {
        const uint32_t c = cnt[q];
        const uint4 r = buf[q * B + gl];
        const bool keep = gl < c, z = global_part(q) == 0;
        const uint4 pad = make_uint4(z ? pad_zero.x : pad_other.x, z ? pad_zero.y : pad_other.y,
                                     z ? pad_zero.z : pad_other.z, z ? pad_zero.w : pad_other.w);
        out[at + gl] = make_uint4(keep ? r.x : pad.x, keep ? r.y : pad.y, keep ? r.z : pad.z, keep ? r.w : pad.w);
    };
    if (APP) {
        const uint32_t qo = threadIdx.x;
        const uint32_t no = qo < pn && cnt[qo] > 0 ? 1u : 0u;
        const uint32_t so = own_slots(no);
        if (no) {
            nch += 1;
            app.cmap[rbase + so] = (uint16_t)global_part(qo);
        }
        for (uint32_t j = gi; j < 64; j += 64 / B) {
            const uint32_t nq = __shfl(no, j, 64), slot = __shfl(so, j, 64);
            if (nq)
                tail(wq0 + j, slot_rec(slot));
        }
        __syncthreads();
        // chunk counts partition-major: the owners' partitions, then 0 for
        // the partner block's half
        if (qo < pn)
            app.bhist[(uint64_t)(p0 + qo) * gridDim.x + blockIdx.x] = nch;
        if (SPLIT == 2) {
            const uint32_t gp = (half ? 0u : FL) + qo;
            if (gp < (half ? FL : F))
                app.bhist[(uint64_t)gp * gridDim.x + blockIdx.x] = 0;
        }
        if (threadIdx.x == 0)
            app.nchunks[blockIdx.x] = *ccur;
    } else {
        for (uint32_t q = grp; q < pn; q += ngrp)
            if (cnt[q] > 0)
                tail(q, goff[q]);
    }
}

// The appended layout's chunk list: block b of the scatter's grid moves the
// numbers of its chunks (slot k is chunk k * NB + b) to cl, each at the next
// place of its partition in this block's share (offs: the chunk counts
// scanned, partition-major), so that cl lists partition 0's chunks, then
// partition 1's, and so on. A tile of the block's slots is counting-sorted by
// partition in LDS first, so that consecutive lanes write a partition's
// consecutive places.
constexpr uint32_t kPlaceThreads = 1024, kPlaceTile = 18432, kPlaceMaxF = 4096;

__global__ __launch_bounds__(kPlaceThreads) void k_chunk_place(const uint16_t* __restrict__ cmap,
                                                                const uint32_t* __restrict__ nchunks,
                                                                const uint32_t* __restrict__ offs, uint32_t F,
                                                                uint64_t region, uint32_t* __restrict__ cl)
{
    __shared__ uint32_t idx[kPlaceTile];          // the tile's slots, sorted by partition
    __shared__ uint16_t tpart[kPlaceTile];        // their partitions
    __shared__ uint32_t cnt[kPlaceMaxF + 1];      // the tile's chunks a partition, then their starts
    __shared__ uint32_t gpos[kPlaceMaxF];         // next place of each partition in cl
    __shared__ uint32_t wsum[kPlaceThreads / 64 + 1];
    constexpr uint32_t PER = kPlaceTile / kPlaceThreads;
    const uint32_t b = blockIdx.x, NB = gridDim.x, tid = threadIdx.x;
    for (uint32_t p = tid; p < F; p += kPlaceThreads)
        gpos[p] = offs[(uint64_t)p * NB + b];
    const uint16_t* tags = cmap + (uint64_t)b * region;
    const uint32_t nc = nchunks[b];
    for (uint32_t t0 = 0; t0 < nc; t0 += kPlaceTile) {
        const uint32_t tn = nc - t0 < kPlaceTile ? nc - t0 : kPlaceTile;
        for (uint32_t p = tid; p <= F; p += kPlaceThreads)
            cnt[p] = 0;
        __syncthreads();
        // all tag loads first (past tn: the tile's last tag again, unused),
        // then the counting atomics back to back
        uint32_t part[PER], rank[PER];
#pragma unroll
        for (uint32_t j = 0; j < PER; ++j) {
            const uint32_t i = j * kPlaceThreads + tid;
            part[j] = tags[t0 + (i < tn ? i : tn - 1)];
        }
#pragma unroll
        for (uint32_t j = 0; j < PER; ++j)
            if (j * kPlaceThreads + tid < tn)
                rank[j] = atomicAdd(&cnt[part[j]], 1u);
        __syncthreads();
        block_scan_array<kPlaceThreads>(cnt, F, wsum);   // cnt[p]: the tile's first place of p
#pragma unroll
        for (uint32_t j = 0; j < PER; ++j) {
            const uint32_t i = j * kPlaceThreads + tid;
            if (i < tn) {
                idx[cnt[part[j]] + rank[j]] = t0 + i;
                tpart[cnt[part[j]] + rank[j]] = (uint16_t)part[j];
            }
        }
        __syncthreads();
        for (uint32_t i = tid; i < tn; i += kPlaceThreads) {
            const uint32_t p = tpart[i];
            cl[gpos[p] + (i - cnt[p])] = (uint32_t)(b * region) + idx[i];
        }
        __syncthreads();
        for (uint32_t p = tid; p < F; p += kPlaceThreads)
            gpos[p] += cnt[p + 1] - cnt[p];
        __syncthreads();
    }
}

// Partition starts of the appended layout in records: part_start[p] =
// records before partition p's first listed chunk; part_start[F] = all.
__global__ void k_chunk_bounds(const uint32_t* __restrict__ offs, uint32_t F, uint32_t NB,
                               const uint32_t* __restrict__ total, uint32_t* __restrict__ part_start)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < F)
        part_start[p] = offs[(uint64_t)p * NB] * kChunkRecs;
    if (p == F)
        part_start[F] = *total * kChunkRecs;
}

// ---------------------------------------------------------------- two passes
// A probe side whose build partitions are finer (2^bR) than the 2^bS
// partitions one write-combining pass produces is partitioned in two passes,
// high bits first: the write-combining scatter on the top bS bits of the key
// mix, then k_subsplit on the next bR - bS bits within each of those
// partitions. Both passes' offsets come from one histogram over the 2^bR
// fine partitions.

// coarse[p * nblk + b] = the fine counts of partition p's D = 2^dbits fine
// partitions in input chunk b, summed and rounded up to whole chunks of pad.
__global__ void k_coarsen(const uint32_t* __restrict__ fine, uint32_t F, uint32_t nblk, uint32_t dbits, uint32_t pad,
                          uint32_t* __restrict__ coarse)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)F * nblk)
        return;
    const uint32_t p = (uint32_t)(i / nblk), b = (uint32_t)(i % nblk), D = 1u << dbits;
    uint32_t c = 0;
    for (uint32_t q = 0; q < D; ++q)
        c += fine[((uint64_t)p * D + q) * nblk + b];
    coarse[i] = (c + pad - 1) / pad * pad;
}

constexpr uint32_t kSubThreads = 1024, kSubItems = 8;

// The second pass. Block b re-reads the segments that input chunk b's
// scatter wrote: segment p at coff[p * nblk + b] (the scanned coarse counts;
// its end is the next segment's start, or *ctotal), padding included. Each
// record of partition p moves to its fine partition p * D + q (q: the next
// dbits bits of the key mix) at foff[(p * D + q) * nblk + b] (the scanned
// fine counts) plus its rank among the segment's records of q; padding
// records (keys of other partitions) are dropped. Waves take segments
// p = wave, wave + 16, ...; lane q < D keeps fine partition q's next place.
__global__ __launch_bounds__(kSubThreads) void k_subsplit(const uint4* __restrict__ src,
                                                           const uint32_t* __restrict__ coff,
                                                           const uint32_t* __restrict__ ctotal, uint32_t F,
                                                           uint32_t nblk, uint32_t bS, uint32_t dbits,
                                                           const uint32_t* __restrict__ foff, uint4* __restrict__ dst)
{
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, b = blockIdx.x;
    const uint64_t below = lane ? ~0ull >> (64 - lane) : 0ull;
    const uint32_t D = 1u << dbits, bR = bS + dbits;
    const uint64_t m = (uint64_t)F * nblk;
    for (uint32_t p = wave; p < F; p += kSubThreads / 64) {
        const uint64_t idx = (uint64_t)p * nblk + b;
        const uint32_t s0 = coff[idx];
        const uint32_t len = (idx + 1 < m ? coff[idx + 1] : *ctotal) - s0;
        uint32_t next = lane < D ? foff[((uint64_t)p * D + lane) * nblk + b] : 0u;
        for (uint32_t i0 = 0; i0 < len; i0 += 64 * kSubItems) {
            // unconditional loads (past the segment: its last record again)
            uint4 r[kSubItems];
#pragma unroll
            for (uint32_t k = 0; k < kSubItems; ++k) {
                const uint32_t i = i0 + k * 64 + lane;
                r[k] = src[s0 + (i < len ? i : len - 1)];
            }
#pragma unroll
            for (uint32_t k = 0; k < kSubItems; ++k) {
                const uint32_t f = radix_part(key_mix(rec_key(r[k])), bR);
                const bool v = i0 + k * 64 + lane < len && (f >> dbits) == p;
                const uint32_t q = f & (D - 1);
                uint32_t dest = 0;
                for (uint32_t qq = 0; qq < D; ++qq) {
                    const bool mine = v && q == qq;
                    const uint64_t mq = __ballot(mine);
                    const uint32_t base = __builtin_amdgcn_readlane(next, qq);
                    if (mine)
                        dest = base + (uint32_t)__popcll(mq & below);
                    if (lane == qq)
                        next += (uint32_t)__popcll(mq);
                }
                if (v)
                    dst[dest] = r[k];
            }
        }
    }
}

constexpr int kScatterItems = 16;

// Scatter with an LDS-sorted tile, for fan-outs too wide for one LDS buffer
// per partition: each tile is counting-sorted by partition in LDS, then
// written so that lanes of a wave store consecutive records of a partition.
// Range partitions take each record's partition from the histogram pass
// (pid) and keep it in LDS (tpart).
template <int RM, bool RANGE>
__global__ __launch_bounds__(kScatterThreads) void k_scatter(const uint32_t* __restrict__ in, uint64_t n,
                                                              SideLayout L, PartFn pf, uint64_t chunk,
                                                              uint32_t nblk, const uint32_t* __restrict__ offs,
                                                              uint4* __restrict__ out, uint32_t ts,
                                                              const uint16_t* __restrict__ pid)
{
    extern __shared__ uint4 sm4[];
    const uint32_t F = pf.F;
    uint4* tile = sm4;
    uint32_t* cnt = reinterpret_cast<uint32_t*>(tile + ts);                       // F + 1
    uint32_t* goff = cnt + F + 1;                                                 // F
    uint32_t* wsum = goff + F;                                                    // 16
    uint16_t* tpart = reinterpret_cast<uint16_t*>(wsum + 16);                     // ts (range)
    for (uint32_t p = threadIdx.x; p < F; p += kScatterThreads)
        goff[p] = offs[(uint64_t)p * nblk + blockIdx.x];
    const uint64_t beg = (uint64_t)blockIdx.x * chunk;
    const uint64_t end = beg + chunk < n ? beg + chunk : n;
    for (uint64_t t0 = beg; t0 < end; t0 += ts) {
        const uint32_t tn = (uint32_t)(end - t0 < ts ? end - t0 : ts);
        for (uint32_t p = threadIdx.x; p <= F; p += kScatterThreads)
            cnt[p] = 0;
        __syncthreads();
        uint4 r[kScatterItems];
        uint32_t part[kScatterItems], rank[kScatterItems];
#pragma unroll
        for (int k = 0; k < kScatterItems; ++k) {
            uint32_t i = k * kScatterThreads + threadIdx.x;
            if (i < tn) {
                r[k] = load_rec<RM>(in, t0 + i, L);
                if (RANGE)
                    part[k] = pid[t0 + i];
            }
        }
#pragma unroll
        for (int k = 0; k < kScatterItems; ++k) {
            uint32_t i = k * kScatterThreads + threadIdx.x;
            if (i < tn) {
                if (!RANGE)
                    part[k] = part_of(rec_key(r[k]), pf);
                rank[k] = atomicAdd(&cnt[part[k]], 1u);
            }
        }
        __syncthreads();
        block_scan_array<kScatterThreads>(cnt, F, wsum);
#pragma unroll
        for (int k = 0; k < kScatterItems; ++k) {
            uint32_t i = k * kScatterThreads + threadIdx.x;
            if (i < tn) {
                const uint32_t slot = cnt[part[k]] + rank[k];
                tile[slot] = r[k];
                if (RANGE)
                    tpart[slot] = (uint16_t)part[k];
            }
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < tn; i += kScatterThreads) {
            uint4 v = tile[i];
            const uint32_t p = RANGE ? (uint32_t)tpart[i] : part_of(rec_key(v), pf);
            out[goff[p] + (i - cnt[p])] = v;
        }
        __syncthreads();
        for (uint32_t p = threadIdx.x; p < F; p += kScatterThreads)
            goff[p] += cnt[p + 1] - cnt[p];
        __syncthreads();
    }
}

// ======================================================================
// join: LDS-resident table per build partition, streaming probe
// ======================================================================

constexpr uint32_t kMiscWords = 64;   // LDS words ahead of the table: item, flags, wave bases, wave sums

// Bucket of a 32-bit key code (Fibonacci hashing onto 12 bits).
__device__ __forceinline__ uint32_t code_bucket(uint32_t c) { return (c * 0x9E3779B1u) >> 20; }

// Table views. lookup() counts a probe record's matches and returns the
// slot of the first; match() tests a later slot of the same bucket against
// the probe's match key mkey(); bkey() is the build key of a matching slot.
struct RadixTab {
    const uint64_t* tkey;
    const uint2* tpay;
    const uint16_t* dir;
    uint32_t bits, bkt_shift, part;
    __device__ __forceinline__ void lookup(const uint4& s, uint32_t& nm, uint32_t& first) const
    {
        const uint64_t key = rec_key(s);
        const uint64_t h = key_mix(key);
        if (radix_part(h, bits) == part) {
            const uint32_t b = (uint32_t)(h >> bkt_shift) & (kTableBuckets - 1);
            const uint32_t lo = b ? dir[b - 1] : 0u, hi = dir[b];
            for (uint32_t j = lo; j < hi; ++j)
                if (tkey[j] == key) {
                    if (nm == 0)
                        first = j;
                    nm++;
                }
        }
    }
    __device__ __forceinline__ uint64_t mkey(const uint4& s) const { return rec_key(s); }
    __device__ __forceinline__ bool match(uint32_t j, uint64_t mk) const { return tkey[j] == mk; }
    __device__ __forceinline__ uint32_t next(uint32_t j) const { return j + 1; }
    __device__ __forceinline__ uint64_t bkey(uint32_t j, const uint4&) const { return tkey[j]; }
    __device__ __forceinline__ uint2 bpay(uint32_t j) const { return tpay[j]; }
};

struct CodeTab {
    const uint32_t* tcode;
    const uint2* tpay;
    const uint16_t* dir;
    uint64_t lo, hi, base;
    __device__ __forceinline__ void lookup(const uint4& s, uint32_t& nm, uint32_t& first) const
    {
        const uint64_t u = ord_key(rec_key(s));
        const uint64_t d = u - base;
        if (u >= lo && u <= hi && (d >> 32) == 0) {
            const uint32_t c = (uint32_t)d;
            const uint32_t b = code_bucket(c);
            const uint32_t blo = b ? dir[b - 1] : 0u, bhi = dir[b];
            for (uint32_t j = blo; j < bhi; ++j)
                if (tcode[j] == c) {
                    if (nm == 0)
                        first = j;
                    nm++;
                }
        }
    }
    __device__ __forceinline__ uint64_t mkey(const uint4& s) const { return (uint32_t)(ord_key(rec_key(s)) - base); }
    __device__ __forceinline__ bool match(uint32_t j, uint64_t mk) const { return tcode[j] == (uint32_t)mk; }
    __device__ __forceinline__ uint32_t next(uint32_t j) const { return j + 1; }
    __device__ __forceinline__ uint64_t bkey(uint32_t, const uint4& s) const { return rec_key(s); }
    __device__ __forceinline__ uint2 bpay(uint32_t j) const { return tpay[j]; }
};

struct DenseTab {
    const uint2* tpay;
    uint64_t lo, hi, base;
    uint32_t cnt;
    __device__ __forceinline__ void lookup(const uint4& s, uint32_t& nm, uint32_t& first) const
    {
        const uint64_t u = ord_key(rec_key(s));
        const uint64_t d = u - base;
        if (u >= lo && u <= hi && d < cnt) {
            first = (uint32_t)d;
            nm = 1;
        }
    }
    __device__ __forceinline__ uint64_t mkey(const uint4&) const { return 0; }
    __device__ __forceinline__ bool match(uint32_t, uint64_t) const { return true; }
    __device__ __forceinline__ uint32_t next(uint32_t j) const { return j + 1; }
    __device__ __forceinline__ uint64_t bkey(uint32_t, const uint4& s) const { return rec_key(s); }
    __device__ __forceinline__ uint2 bpay(uint32_t j) const { return tpay[j]; }
};

// Two-pass counting sort of build records into an LDS bucket table. Pass 1
// counts per bucket, a scan turns counts into bucket starts, pass 2 places
// each record at its bucket's cursor; afterwards dir entry b holds the end of
// bucket b. Two 16-bit entries share a 32-bit word. bucket_of(r) returns the
// bucket of record r, or ~0 to skip it; put(j, r, i) stores record r (index
// i of R) at slot j. Each pass loads kBuildBatch records a thread before
// using any (one memory latency a batch, not one a record), and pass 2
// issues a batch's cursor atomics back to back before its stores.
constexpr uint32_t kBuildBatch = 8;

template <class BucketOf, class Put>
__device__ __forceinline__ void build_buckets(const uint4* __restrict__ R, uint32_t r_off, uint32_t r_cnt,
                                              uint32_t* dirw, uint32_t* wsum, BucketOf bucket_of, Put put)
{
    const uint32_t tid = threadIdx.x;
    constexpr uint32_t BU = kBuildBatch;
    // batch i0: record i0 + u * kJoinThreads + tid in r[u]; indexes past
    // r_cnt (r_cnt > 0 inside the loops) load the last record again, so that
    // every load is unconditional
    auto load = [&](uint32_t i0, uint4(&r)[BU]) {
#pragma unroll
        for (uint32_t u = 0; u < BU; ++u) {
            const uint32_t i = i0 + u * kJoinThreads + tid;
            r[u] = R[r_off + (i < r_cnt ? i : r_cnt - 1)];
        }
    };
    for (uint32_t w = tid; w <= kTableBuckets / 2; w += kJoinThreads)
        dirw[w] = 0;
    __syncthreads();
    for (uint32_t i0 = 0; i0 < r_cnt; i0 += BU * kJoinThreads) {
        uint4 r[BU];
        load(i0, r);
#pragma unroll
        for (uint32_t u = 0; u < BU; ++u) {
            const uint32_t b = i0 + u * kJoinThreads + tid < r_cnt ? bucket_of(r[u]) : ~0u;
            if (b != ~0u)
                atomicAdd(&dirw[b >> 1], 1u << ((b & 1u) * 16u));
        }
    }
    __syncthreads();
    {
        static_assert(kTableBuckets == 4 * kJoinThreads, "scan layout");
        const uint32_t w0 = dirw[2 * tid], w1 = dirw[2 * tid + 1];
        const uint32_t c0 = w0 & 0xFFFFu, c1 = w0 >> 16, c2 = w1 & 0xFFFFu, c3 = w1 >> 16;
        uint32_t total;
        const uint32_t pre = block_excl_scan<kJoinThreads>(c0 + c1 + c2 + c3, wsum, &total);
        dirw[2 * tid] = pre | ((pre + c0) << 16);
        dirw[2 * tid + 1] = (pre + c0 + c1) | ((pre + c0 + c1 + c2) << 16);
    }
    __syncthreads();
    for (uint32_t i0 = 0; i0 < r_cnt; i0 += BU * kJoinThreads) {
        uint4 r[BU];
        load(i0, r);
        uint32_t b[BU], j[BU];
#pragma unroll
        for (uint32_t u = 0; u < BU; ++u)
            b[u] = i0 + u * kJoinThreads + tid < r_cnt ? bucket_of(r[u]) : ~0u;
#pragma unroll
        for (uint32_t u = 0; u < BU; ++u)
            if (b[u] != ~0u) {
                const uint32_t sh = (b[u] & 1u) * 16u;
                j[u] = (atomicAdd(&dirw[b[u] >> 1], 1u << sh) >> sh) & 0xFFFFu;
            }
#pragma unroll
        for (uint32_t u = 0; u < BU; ++u)
            if (b[u] != ~0u)
                put(j[u], r[u], r_off + i0 + u * kJoinThreads + tid);
    }
    __syncthreads();
}

// ---------------------------------------------------------------- work items
// The work list is not materialised: k_plan counts each build partition's
// items (table splits x probe pieces) and scans the counts, and k_join maps
// an item number back to its partition by binary search over the offsets.

// Pieces the s probe records of one build partition are cut into: about
// `chunk` records each and all of one size, so that no item is a small
// remainder that builds a whole table for a few probes.
__host__ __device__ __forceinline__ uint32_t probe_pieces(uint64_t s, uint64_t chunk)
{
    const uint64_t n = (s + chunk / 2) / chunk;
    return n ? (uint32_t)n : 1u;
}

__device__ __forceinline__ uint32_t plan_probe_part(const PlanView& v, uint32_t p)
{
    return v.coarse_of ? v.coarse_of[p] : p >> v.shift;
}

// Items of build partition p; *splits gets its table count.
__device__ __forceinline__ uint32_t plan_count(const PlanView& v, uint32_t p, uint32_t* splits)
{
    const uint32_t ps = plan_probe_part(v, p);
    const uint32_t r = v.rst[p + 1] - v.rst[p], sn = v.sst[ps + 1] - v.sst[ps];
    const uint32_t cap = v.pp ? v.pp[p].cap : kTableCap;
    *splits = (r + cap - 1) / cap;
    return r && sn ? *splits * probe_pieces(sn, v.chunk) : 0u;
}

// Work item `it` (< offs[FR]) of build partition p (offs[p] <= it < offs[p + 1]).
__device__ __forceinline__ Item plan_item_of(const PlanView& v, uint32_t it, uint32_t p)
{
    const uint32_t q = it - v.offs[p];
    const uint32_t ps = plan_probe_part(v, p);
    const uint32_t r0 = v.rst[p], r = v.rst[p + 1] - r0;
    const uint32_t s0 = v.sst[ps];
    const uint64_t sn = v.sst[ps + 1] - s0;
    const uint32_t cap = v.pp ? v.pp[p].cap : kTableCap;
    const uint32_t np = probe_pieces(sn, v.chunk);
    const uint32_t rs = q / np, k = q % np;
    Item I;
    I.r_off = r0 + rs * cap;
    I.r_cnt = r - rs * cap < cap ? r - rs * cap : cap;
    if (v.cl) {   // listed chunks: pieces of whole chunks (partition starts and sizes are whole chunks)
        const uint64_t nc = sn / kChunkRecs;
        I.s_off = s0 + (uint32_t)(nc * k / np) * kChunkRecs;
        I.s_cnt = (uint32_t)(nc * (k + 1) / np - nc * k / np) * kChunkRecs;
    } else {
        I.s_off = s0 + (uint32_t)(sn * k / np);
        I.s_cnt = (uint32_t)(sn * (k + 1) / np - sn * k / np);
    }
    I.part = p;
    I.fmt = v.pp ? v.pp[p].fmt : 0u;
    I.lo = v.pp ? v.pp[p].lo : 0ull;
    I.hi = v.pp ? v.pp[p].hi : 0ull;
    I.base = v.pp ? v.pp[p].base : 0ull;
    return I;
}

// Work item `it` (< offs[FR]) by binary search over offs.
__device__ __forceinline__ Item plan_item(const PlanView& v, uint32_t it)
{
    uint32_t lo = 0, hi = v.FR;   // the last p with offs[p] <= it
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (v.offs[mid] <= it)
            lo = mid;
        else
            hi = mid;
    }
    return plan_item_of(v, it, lo);
}

// Work item `it` found by the whole block at once (it block-uniform): each
// thread tests its share of the build partitions, the one whose item range
// holds `it` leaves its number in *slot (LDS). One memory latency instead of
// a binary search's chain of dependent loads. Ends with a barrier; *slot
// must not be rewritten before the block's next barrier after this one.
__device__ __forceinline__ Item plan_item_block(const PlanView& v, uint32_t it, uint32_t* slot)
{
    for (uint32_t p = threadIdx.x; p < v.FR; p += blockDim.x)
        if (v.offs[p] <= it && it < v.offs[p + 1])
            *slot = p;
    __syncthreads();
    return plan_item_of(v, it, *slot);
}

constexpr uint32_t kPlanThreads = 1024;
constexpr uint32_t kPlanPer = 16;   // build partitions a thread counts: up to 16384

// One block: offs[p] = first item of build partition p, offs[FR] = items;
// also clears the join's counters (words 0-5 and 8-11) and records the item
// count and the number of split tables in them (words 6 and 7).
__global__ __launch_bounds__(kPlanThreads) void k_plan(PlanView v, uint32_t* __restrict__ offs,
                                                        uint32_t* __restrict__ counters)
{
    __shared__ uint32_t wsum[kPlanThreads / 64 + 1];
    const uint32_t per = (v.FR + kPlanThreads - 1) / kPlanThreads;
    const uint32_t p0 = threadIdx.x * per;
    uint32_t c[kPlanPer];
    uint32_t sum = 0, spilled = 0;
#pragma unroll
    for (uint32_t q = 0; q < kPlanPer; ++q) {
        c[q] = 0;
        if (q < per && p0 + q < v.FR) {
            uint32_t splits;
            c[q] = plan_count(v, p0 + q, &splits);
            sum += c[q];
            if (splits > 1 && c[q])
                spilled += splits;
        }
    }
    uint32_t total, stotal;
    uint32_t pre = block_excl_scan<kPlanThreads>(sum, wsum, &total);
#pragma unroll
    for (uint32_t q = 0; q < kPlanPer; ++q)
        if (q < per && p0 + q < v.FR) {
            offs[p0 + q] = pre;
            pre += c[q];
        }
    block_excl_scan<kPlanThreads>(spilled, wsum, &stotal);
    if (threadIdx.x == 0) {
        offs[v.FR] = total;
        for (int w = 0; w < 12; ++w)
            counters[w] = 0;
        counters[6] = total;
        counters[7] = stotal;
    }
}

// Work items come from a shared counter; per item the block builds the LDS
// table of its build records and streams the probe records past it.
// RANGE = false: radix partitions, tables of 64-bit keys. RANGE = true:
// range partitions, whose tables are dense (payload slot = key - base) or
// hold 32-bit key codes; a dense item whose keys turn out not to be distinct
// is rebuilt as a code table, or, when it is too large for one, counted in
// *fail_ctr (the caller then runs the join on radix partitions).
// Probe records in flight a thread in the range kernel, which holds both
// table formats' probe loops: with the radix kernel's 8 it spills registers
// to scratch inside the probe loop.
constexpr int kRangeProbeItems = 6;
// Probe records in flight a thread when the probe side is listed by chunk
// (the appended layout): the chunk numbers and their addresses take the
// registers of one record (8 spill 6 words to scratch).
constexpr int kListProbeItems = 7;

template <bool RANGE, bool CL>
__global__ __launch_bounds__(kJoinThreads) void k_join(const uint4* __restrict__ R, const uint4* __restrict__ S,
                                                        PlanView plan, uint32_t* __restrict__ item_ctr,
                                                        uint32_t* __restrict__ out, uint64_t out_cap,
                                                        unsigned long long* __restrict__ out_cursor,
                                                        OutFormat fmt, uint32_t bitsR, uint32_t bkt_shift,
                                                        uint32_t* __restrict__ fail_ctr)
{
    extern __shared__ uint64_t lds64[];
    const uint32_t tid = threadIdx.x;
    uint32_t* misc = reinterpret_cast<uint32_t*>(lds64);   // [0] item, [1] flag, [2..33] wave bases, [34] partition
    uint32_t* wsum = misc + 40;                            // 17
    uint64_t* table = lds64 + kMiscWords / 2;
    const bool kbp = fmt.words == 6 && fmt.src[0] == SRC_BK0 && fmt.src[1] == SRC_BK1 && fmt.src[2] == SRC_BP0 &&
                     fmt.src[3] == SRC_BP1 && fmt.src[4] == SRC_PP0 && fmt.src[5] == SRC_PP1;
    const uint32_t nitems = plan.offs[plan.FR];

    for (;;) {
        if (tid == 0) {
            misc[0] = atomicAdd(item_ctr, 1u);
            misc[1] = 0;
        }
        __syncthreads();
        const uint32_t it = misc[0];
        if (it >= nitems)
            break;
        const Item I = plan_item_block(plan, it, misc + 34);
        // CL: the probe records are listed by chunk from plan.cl (whole
        // chunks, so I.s_off is a multiple of the chunk's records)
        const uint4* Sp = CL ? S : S + I.s_off;
        const uint32_t* clp = CL ? plan.cl + I.s_off / kChunkRecs : nullptr;

        if (!RANGE) {
            uint64_t* tkey = table;                                            // kTableCap
            uint2* tpay = reinterpret_cast<uint2*>(tkey + kTableCap);          // kTableCap
            uint32_t* dirw = reinterpret_cast<uint32_t*>(tpay + kTableCap);    // kTableBuckets / 2 + 1
            build_buckets(
                R, I.r_off, I.r_cnt, dirw, wsum,
                [&](const uint4& r) -> uint32_t {
                    const uint64_t h = key_mix(rec_key(r));
                    return radix_part(h, bitsR) == I.part ? (uint32_t)(h >> bkt_shift) & (kTableBuckets - 1) : ~0u;
                },
                [&](uint32_t j, const uint4& r, uint32_t) {
                    tkey[j] = rec_key(r);
                    tpay[j] = make_uint2(r.z, r.w);
                });
            const RadixTab tab{tkey, tpay, reinterpret_cast<const uint16_t*>(dirw), bitsR, bkt_shift, I.part};
            probe_rounds<kJoinThreads, RadixTab, CL ? kListProbeItems : kProbeItems, CL>(
                tab, Sp, I.s_cnt, out, out_cap, out_cursor, fmt, kbp, wsum, misc + 2, clp);
        } else {
            uint2* tpay = reinterpret_cast<uint2*>(table);
            bool dense = I.fmt == kFmtDense;
            if (dense) {
                uint32_t* bits = reinterpret_cast<uint32_t*>(tpay + kDenseCap);   // presence bitmap
                for (uint32_t w = tid; w < (I.r_cnt + 31) / 32; w += kJoinThreads)
                    bits[w] = 0;
                __syncthreads();
                for (uint32_t i = tid; i < I.r_cnt; i += kJoinThreads) {
                    const uint4 r = R[I.r_off + i];
                    const uint64_t d = ord_key(rec_key(r)) - I.base;   // < r_cnt when the plan is right
                    if (d >= I.r_cnt) {
                        misc[1] = 2;
                        continue;
                    }
                    tpay[d] = make_uint2(r.z, r.w);
                    const uint32_t bit = 1u << (d & 31u);
                    if (atomicOr(&bits[d >> 5], bit) & bit)
                        misc[1] = 1;   // a repeated key: the partition is not dense
                }
                __syncthreads();
                if (misc[1]) {
                    if (misc[1] == 2 || I.r_cnt > kCodeCap) {
                        if (tid == 0)
                            atomicAdd(fail_ctr, 1u);
                        __syncthreads();
                        continue;
                    }
                    dense = false;
                }
            }
            if (dense) {
                const DenseTab tab{tpay, I.lo, I.hi, I.base, I.r_cnt};
                probe_rounds<kJoinThreads, DenseTab, kRangeProbeItems, CL>(tab, Sp, I.s_cnt, out, out_cap,
                                                                           out_cursor, fmt, kbp, wsum, misc + 2, clp);
            } else {
                uint32_t* tcode = reinterpret_cast<uint32_t*>(tpay + kCodeCap);   // kCodeCap
                uint32_t* dirw = tcode + kCodeCap;                                 // kTableBuckets / 2 + 1
                const uint64_t base = I.base;
                build_buckets(
                    R, I.r_off, I.r_cnt, dirw, wsum,
                    [&](const uint4& r) -> uint32_t { return code_bucket((uint32_t)(ord_key(rec_key(r)) - base)); },
                    [&](uint32_t j, const uint4& r, uint32_t) {
                        tcode[j] = (uint32_t)(ord_key(rec_key(r)) - base);
                        tpay[j] = make_uint2(r.z, r.w);
                    });
                const CodeTab tab{tcode, tpay, reinterpret_cast<const uint16_t*>(dirw), I.lo, I.hi, I.base};
                probe_rounds<kJoinThreads, CodeTab, kRangeProbeItems, CL>(tab, Sp, I.s_cnt, out, out_cap,
                                                                          out_cursor, fmt, kbp, wsum, misc + 2, clp);
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- group-join
// The count per build key of a radix join's matches (aggregate_count grouped
// by the join key over the join's output) without writing the join's rows.
// Slot j of the build table keeps its record's index in R_part (x) and the
// matches it met in this item (y, an LDS counter); after the item the
// counts are added to rcount[] in HBM, where an item's probe pieces and
// table splits add up. A key that an earlier slot of its bucket already
// holds, or a partition split over several tables, sets *dup: equal keys
// then have several records, which the caller merges.
// CL: the probe side is appended (its records listed by chunk in plan.cl,
// read as probe_rounds reads them).
constexpr uint32_t kGjUnroll = 8;

template <bool CL>
__global__ __launch_bounds__(kJoinThreads) void k_groupjoin(const uint4* __restrict__ R, const uint4* __restrict__ S,
                                                             PlanView plan, uint32_t* __restrict__ item_ctr,
                                                             uint32_t* __restrict__ rcount, uint32_t bitsR,
                                                             uint32_t bkt_shift, uint32_t* __restrict__ dup)
{
    extern __shared__ uint64_t lds64[];
    const uint32_t tid = threadIdx.x;
    uint32_t* misc = reinterpret_cast<uint32_t*>(lds64);
    uint32_t* wsum = misc + 40;
    uint64_t* table = lds64 + kMiscWords / 2;
    const uint32_t nitems = plan.offs[plan.FR];
    for (;;) {
        if (tid == 0)
            misc[0] = atomicAdd(item_ctr, 1u);
        __syncthreads();
        const uint32_t it = misc[0];
        if (it >= nitems)
            break;
        const Item I = plan_item_block(plan, it, misc + 34);
        const uint4* Sp = CL ? S : S + I.s_off;
        const uint32_t* clp = CL ? plan.cl + I.s_off / kChunkRecs : nullptr;
        uint64_t* tkey = table;                                            // kTableCap
        uint2* tpay = reinterpret_cast<uint2*>(tkey + kTableCap);          // kTableCap
        uint32_t* dirw = reinterpret_cast<uint32_t*>(tpay + kTableCap);    // kTableBuckets / 2 + 1
        build_buckets(
            R, I.r_off, I.r_cnt, dirw, wsum,
            [&](const uint4& r) -> uint32_t {
                const uint64_t h = key_mix(rec_key(r));
                return radix_part(h, bitsR) == I.part ? (uint32_t)(h >> bkt_shift) & (kTableBuckets - 1) : ~0u;
            },
            [&](uint32_t j, const uint4& r, uint32_t i) {
                tkey[j] = rec_key(r);
                tpay[j] = make_uint2(i, 0u);
            });
        const uint16_t* dir = reinterpret_cast<const uint16_t*>(dirw);
        const RadixTab tab{tkey, tpay, dir, bitsR, bkt_shift, I.part};
        // Each wave probes a slice of kGjUnroll * 64 records a round; a
        // record's register is refilled with the next round's record as soon
        // as its lookup is done, so loads stay in flight while LDS work runs.
        constexpr uint32_t kSlice = 64 * kGjUnroll, kRound = (kJoinThreads / 64) * kSlice;
        constexpr uint32_t kSliceChunks = kSlice / kChunkRecs;
        static_assert(kSliceChunks <= 64, "a slice's chunks fit a wave");
        const uint32_t lane = tid & 63, wave = tid >> 6;
        // CL: lane l holds the number of chunk l of its wave's slice of round
        // r0 (loaded two rounds ahead) and hands it to the lanes that load it
        auto chunk_at = [&](uint32_t r0) -> uint32_t {
            const uint32_t c = (r0 + wave * kSlice) / kChunkRecs + lane;
            return lane < kSliceChunks && c * kChunkRecs < I.s_cnt ? clp[c] : 0u;
        };
        auto rec_at = [&](uint32_t ch, uint32_t u) -> const uint4* {
            const uint32_t c = __shfl(ch, u * (64 / kChunkRecs) + lane / kChunkRecs, 64);
            return Sp + ((uint64_t)c * kChunkRecs + lane % kChunkRecs);
        };
        uint4 nxt[kGjUnroll];
        uint32_t ch_next = 0;
        if (CL) {   // unconditional first loads (a lane past s_cnt: the list's first chunk)
            const uint32_t ch0 = chunk_at(0);
            ch_next = chunk_at(kRound);
#pragma unroll
            for (uint32_t u = 0; u < kGjUnroll; ++u)
                nxt[u] = ld_stream(rec_at(ch0, u));
        } else if (I.s_cnt) {   // unconditional first loads (past s_cnt: the last record again, never looked up)
#pragma unroll
            for (uint32_t u = 0; u < kGjUnroll; ++u) {
                const uint32_t i = wave * kSlice + u * 64 + lane;
                nxt[u] = ld_stream(Sp + (i < I.s_cnt ? i : I.s_cnt - 1));
            }
        }
        for (uint32_t r0 = 0; r0 < I.s_cnt; r0 += kRound) {
            const uint32_t t0 = r0 + wave * kSlice;
            const uint32_t ch_after = CL ? chunk_at(r0 + 2 * kRound) : 0u;
#pragma unroll
            for (uint32_t u = 0; u < kGjUnroll; ++u) {
                const uint32_t i = t0 + u * 64 + lane;
                const uint4 s = nxt[u];
                uint32_t nm = 0, first = 0;
                if (i < I.s_cnt)
                    tab.lookup(s, nm, first);
                if (CL) {
                    const uint4* a = rec_at(ch_next, u);
                    if (i + kRound < I.s_cnt)
                        nxt[u] = ld_stream(a);
                } else if (i + kRound < I.s_cnt) {
                    nxt[u] = ld_stream(Sp + i + kRound);
                }
                if (nm == 1) {
                    atomicAdd(&tpay[first].y, 1u);
                } else if (nm) {
                    const uint64_t mk = rec_key(s);
                    for (uint32_t j = first, c = 0; c < nm; ++j)   // the nm matches lie in first's bucket
                        if (tkey[j] == mk) {
                            atomicAdd(&tpay[j].y, 1u);
                            ++c;
                        }
                }
            }
            ch_next = ch_after;
        }
        __syncthreads();
        const uint32_t ntab = dir[kTableBuckets - 1];
        const bool split = plan.rst[I.part + 1] - plan.rst[I.part] > I.r_cnt;
        // An item that is its partition's only one (one table, one probe
        // piece) holds the final counts and stores them; otherwise items add.
        const uint32_t ps = plan_probe_part(plan, I.part);
        const bool sole = !split && probe_pieces(plan.sst[ps + 1] - plan.sst[ps], plan.chunk) == 1;
        bool seen = split;
        for (uint32_t j = tid; j < ntab; j += kJoinThreads) {
            const uint2 pv = tpay[j];
            if (pv.y) {
                if (sole)
                    rcount[pv.x] = pv.y;
                else
                    atomicAdd(&rcount[pv.x], pv.y);
            }
            const uint64_t k = tkey[j];
            const uint32_t b = (uint32_t)(key_mix(k) >> bkt_shift) & (kTableBuckets - 1);
            for (uint32_t m = b ? dir[b - 1] : 0u; m < j; ++m)
                if (tkey[m] == k)
                    seen = true;
        }
        if (seen)
            atomicOr(dup, 1u);
        __syncthreads();
    }
}

// One (build key, count) tuple per record of R_part that met matches
// (records [0, rst[FR]); padding records never match). Each block takes one
// contiguous range: it counts its groups, reserves their rows with one
// atomic on *cursor, then writes them in order. Rows past cap are counted,
// not written.
constexpr uint32_t kEmitThreads = 256;

__global__ __launch_bounds__(kEmitThreads) void k_groupjoin_emit(const uint4* __restrict__ R,
                                                                  const uint32_t* __restrict__ rst, uint32_t FR,
                                                                  const uint32_t* __restrict__ rcount,
                                                                  uint32_t key_words, uint32_t* __restrict__ out,
                                                                  uint64_t cap, unsigned long long* __restrict__ cursor)
{
    __shared__ uint32_t wcnt[kEmitThreads / 64];
    __shared__ unsigned long long base_s;
    const uint32_t n = rst[FR];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t per = ((n + gridDim.x - 1) / gridDim.x + kEmitThreads - 1) / kEmitThreads * kEmitThreads;
    const uint32_t lo = blockIdx.x * per < n ? blockIdx.x * per : n;
    const uint32_t hi = n - lo > per ? lo + per : n;
    uint32_t mine = 0;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += kEmitThreads)
        mine += rcount[i] != 0;
    for (int d = 32; d > 0; d >>= 1)
        mine += __shfl_xor(mine, d, 64);
    if (lane == 0)
        wcnt[wave] = mine;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < kEmitThreads / 64; ++w)
            t += wcnt[w];
        base_s = t ? atomicAdd(cursor, (unsigned long long)t) : 0ull;
    }
    __syncthreads();
    uint64_t next = base_s;   // the block's next row
    for (uint32_t i0 = lo; i0 < hi; i0 += kEmitThreads) {
        const uint32_t i = i0 + threadIdx.x;
        const uint32_t c = i < hi ? rcount[i] : 0u;
        const uint64_t m = __ballot(c != 0);
        if (lane == 0)
            wcnt[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (uint32_t w = 0; w < kEmitThreads / 64; ++w) {
            before += w < wave ? wcnt[w] : 0u;
            total += wcnt[w];
        }
        __syncthreads();
        if (c) {
            const uint64_t row = next + before + (uint64_t)__popcll(m & (lane ? (~0ull >> (64 - lane)) : 0ull));
            if (row < cap) {
                const uint64_t k = rec_key(R[i]);
                uint32_t* o = out + row * (key_words + 2);
                o[0] = (uint32_t)k;
                if (key_words == 2)
                    o[1] = (uint32_t)(k >> 32);
                o[key_words] = c;   // the count as a LONG
                o[key_words + 1] = 0u;
            }
        }
        next += total;
    }
}

// ---------------------------------------------------------------- batched joins
// The pipeline of a batch (MBatch, pj_internal.h): every kernel runs once for
// all joins. A block of the histogram and scatter kernels finds its side by
// binary search over the sides' first blocks; a work item of the join kernel
// finds its join by binary search over the joins' first items. The device
// loops are bounded by the descriptors the host wrote and checked.

constexpr uint32_t kMaxBatchBits = 12;

__device__ __forceinline__ uint32_t side_of_block(const MSide* __restrict__ sides, uint32_t nsides, uint32_t b)
{
    uint32_t lo = 0, hi = nsides;   // the last side whose first block is at or below b
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (sides[mid].blk0 <= b)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// hist[S.hbase + p * S.nblk + t] = tuples of tile t of side S in partition p.
template <int RM>
__global__ __launch_bounds__(kHistThreads) void k_mhist(const MSide* __restrict__ sides, uint32_t nsides,
                                                         SideLayout L0, SideLayout L1, uint32_t* __restrict__ hist)
{
    extern __shared__ uint32_t mcnt[];
    const MSide S = sides[side_of_block(sides, nsides, blockIdx.x)];
    const uint32_t t = blockIdx.x - S.blk0, F = 1u << S.bits;
    const SideLayout& L = S.probe ? L1 : L0;
    for (uint32_t p = threadIdx.x; p < F; p += kHistThreads)
        mcnt[p] = 0;
    __syncthreads();
    const uint64_t beg = (uint64_t)t * S.chunk;
    const uint64_t end = beg + S.chunk < S.n ? beg + S.chunk : S.n;
    for (uint64_t base = beg; base < end; base += (uint64_t)kHistThreads * kHistItems) {
        uint4 r[kHistItems];
#pragma unroll
        for (int k = 0; k < kHistItems; ++k) {
            const uint64_t i = base + (uint64_t)k * kHistThreads + threadIdx.x;
            if (i < end)
                r[k] = load_rec<RM>(S.in, i, L);
        }
#pragma unroll
        for (int k = 0; k < kHistItems; ++k) {
            const uint64_t i = base + (uint64_t)k * kHistThreads + threadIdx.x;
            if (i < end)
                atomicAdd(&mcnt[radix_part(key_mix(rec_key(r[k])), S.bits)], 1u);
        }
    }
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < F; p += kHistThreads)
        hist[S.hbase + p * S.nblk + t] = mcnt[p];
}

// Tile-sorted scatter of tile t of side S into recs at the scanned offsets:
// each LDS tile of ts records is counting-sorted by partition, then written
// so that lanes of a wave store consecutive records of a partition.
template <int RM>
__global__ __launch_bounds__(kScatterThreads) void k_mscatter(const MSide* __restrict__ sides, uint32_t nsides,
                                                               SideLayout L0, SideLayout L1,
                                                               const uint32_t* __restrict__ offs, uint32_t ts,
                                                               uint32_t maxF, uint4* __restrict__ recs)
{
    extern __shared__ uint4 sm4[];
    const MSide S = sides[side_of_block(sides, nsides, blockIdx.x)];
    const uint32_t t = blockIdx.x - S.blk0, F = 1u << S.bits;
    const SideLayout& L = S.probe ? L1 : L0;
    uint4* tile = sm4;
    uint32_t* cnt = reinterpret_cast<uint32_t*>(tile + ts);   // maxF + 1
    uint32_t* goff = cnt + maxF + 1;                          // maxF
    uint32_t* wsum = goff + maxF;                             // 16
    for (uint32_t p = threadIdx.x; p < F; p += kScatterThreads)
        goff[p] = offs[S.hbase + p * S.nblk + t];
    const uint64_t beg = (uint64_t)t * S.chunk;
    const uint64_t end = beg + S.chunk < S.n ? beg + S.chunk : S.n;
    for (uint64_t t0 = beg; t0 < end; t0 += ts) {
        const uint32_t tn = (uint32_t)(end - t0 < ts ? end - t0 : ts);
        for (uint32_t p = threadIdx.x; p <= F; p += kScatterThreads)
            cnt[p] = 0;
        __syncthreads();
        uint4 r[kScatterItems];
        uint32_t part[kScatterItems], rank[kScatterItems];
#pragma unroll
        for (int k = 0; k < kScatterItems; ++k) {
            const uint32_t i = k * kScatterThreads + threadIdx.x;
            if (i < tn) {
                r[k] = load_rec<RM>(S.in, t0 + i, L);
                part[k] = radix_part(key_mix(rec_key(r[k])), S.bits);
                rank[k] = atomicAdd(&cnt[part[k]], 1u);
            }
        }
        __syncthreads();
        block_scan_array<kScatterThreads>(cnt, F, wsum);
#pragma unroll
        for (int k = 0; k < kScatterItems; ++k) {
            const uint32_t i = k * kScatterThreads + threadIdx.x;
            if (i < tn)
                tile[cnt[part[k]] + rank[k]] = r[k];
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < tn; i += kScatterThreads) {
            const uint4 v = tile[i];
            const uint32_t p = radix_part(key_mix(rec_key(v)), S.bits);
            recs[goff[p] + (i - cnt[p])] = v;
        }
        __syncthreads();
        for (uint32_t p = threadIdx.x; p < F; p += kScatterThreads)
            goff[p] += cnt[p + 1] - cnt[p];
        __syncthreads();
    }
}

// Partition starts of every side (one block a side): starts[S.pbase + p] =
// the first record of partition p in recs, starts[S.pbase + F] = its end.
__global__ void k_mbounds(const MSide* __restrict__ sides, const uint32_t* __restrict__ offs,
                          uint32_t* __restrict__ starts)
{
    const MSide S = sides[blockIdx.x];
    const uint32_t F = 1u << S.bits;
    for (uint32_t p = threadIdx.x; p <= F; p += blockDim.x)
        starts[S.pbase + p] = p == F ? S.rbase + S.n : S.nblk ? offs[S.hbase + p * S.nblk] : S.rbase;
}

// The work items of join blockIdx.x (one block a join, as k_plan): offs[p]
// = first item of build partition p, offs[FR] = the join's items; its state
// is cleared and armed with that count.
__global__ __launch_bounds__(kPlanThreads) void k_mplan(const MJoin* __restrict__ joins, MState* __restrict__ state)
{
    __shared__ uint32_t wsum[kPlanThreads / 64 + 1];
    const PlanView v = joins[blockIdx.x].v;
    uint32_t* offs = const_cast<uint32_t*>(v.offs);
    const uint32_t per = (v.FR + kPlanThreads - 1) / kPlanThreads;
    const uint32_t p0 = threadIdx.x * per;
    uint32_t c[kPlanPer];
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t q = 0; q < kPlanPer; ++q) {
        c[q] = 0;
        if (q < per && p0 + q < v.FR) {
            uint32_t splits;
            c[q] = plan_count(v, p0 + q, &splits);
            sum += c[q];
        }
    }
    uint32_t total;
    uint32_t pre = block_excl_scan<kPlanThreads>(sum, wsum, &total);
#pragma unroll
    for (uint32_t q = 0; q < kPlanPer; ++q)
        if (q < per && p0 + q < v.FR) {
            offs[p0 + q] = pre;
            pre += c[q];
        }
    if (threadIdx.x == 0) {
        offs[v.FR] = total;
        MState& m = state[blockIdx.x];
        m.rows = 0;
        m.remaining = total;
        m.items = total;
        m.t_done = 0;
    }
}

// One block: ibase[j] = the first work item of join j, ibase[nj] = all
// items; clears the item counter and the error flag and stamps the batch's
// start (misc words 2-3).
__global__ __launch_bounds__(1024) void k_mitems(const MState* __restrict__ state, uint32_t nj,
                                                  uint32_t* __restrict__ ibase, uint32_t* __restrict__ misc,
                                                  unsigned long long* __restrict__ t0)
{
    __shared__ uint32_t wsum[1024 / 64 + 1];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0)
        carry = 0;
    __syncthreads();
    for (uint32_t b = 0; b < nj; b += 1024) {
        const uint32_t j = b + threadIdx.x;
        const uint32_t v = j < nj ? state[j].items : 0u;
        uint32_t total;
        const uint32_t pre = block_excl_scan<1024>(v, wsum, &total);
        if (j < nj)
            ibase[j] = carry + pre;
        __syncthreads();
        if (threadIdx.x == 0)
            carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ibase[nj] = carry;
        misc[0] = 0;   // item counter
        misc[1] = 0;   // error flag
        *t0 = __builtin_amdgcn_s_memrealtime();
    }
}

// Every join's work items from one counter (one block a CU, as k_join): the
// item's join by binary search over ibase, then the build partition's LDS
// table and the probe records streamed past it into that join's output. The
// last item of a join stamps its completion. A count of items above the
// host's bound (max_items) is refused with misc[1] = 1 before any item runs.
__global__ __launch_bounds__(kJoinThreads) void k_mjoin(const uint4* __restrict__ recs,
                                                         const MJoin* __restrict__ joins, uint32_t nj,
                                                         const uint32_t* __restrict__ ibase,
                                                         MState* __restrict__ state, uint32_t* __restrict__ misc,
                                                         uint32_t max_items, OutFormat fmt)
{
    extern __shared__ uint64_t lds64[];
    const uint32_t tid = threadIdx.x;
    uint32_t* lmisc = reinterpret_cast<uint32_t*>(lds64);   // [0] item, [2..33] wave bases
    uint32_t* wsum = lmisc + 40;
    uint64_t* table = lds64 + kMiscWords / 2;
    uint64_t* tkey = table;                                            // kTableCap
    uint2* tpay = reinterpret_cast<uint2*>(tkey + kTableCap);          // kTableCap
    uint32_t* dirw = reinterpret_cast<uint32_t*>(tpay + kTableCap);    // kTableBuckets / 2 + 1
    const bool kbp = fmt.words == 6 && fmt.src[0] == SRC_BK0 && fmt.src[1] == SRC_BK1 && fmt.src[2] == SRC_BP0 &&
                     fmt.src[3] == SRC_BP1 && fmt.src[4] == SRC_PP0 && fmt.src[5] == SRC_PP1;
    const uint32_t nitems = ibase[nj];
    if (nitems > max_items) {
        if (blockIdx.x == 0 && tid == 0)
            misc[1] = 1;
        return;
    }
    for (;;) {
        if (tid == 0)
            lmisc[0] = atomicAdd(&misc[0], 1u);
        __syncthreads();
        const uint32_t it = lmisc[0];
        if (it >= nitems)
            break;
        uint32_t lo = 0, hi = nj;   // the last join whose first item is at or below it
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (ibase[mid] <= it)
                lo = mid;
            else
                hi = mid;
        }
        const MJoin& J = joins[lo];
        const PlanView v = J.v;
        const uint32_t bits = J.bits, bkt_shift = 64 - bits - 12;
        const Item I = plan_item(v, it - ibase[lo]);
        build_buckets(
            recs, I.r_off, I.r_cnt, dirw, wsum,
            [&](const uint4& r) -> uint32_t {
                const uint64_t h = key_mix(rec_key(r));
                return radix_part(h, bits) == I.part ? (uint32_t)(h >> bkt_shift) & (kTableBuckets - 1) : ~0u;
            },
            [&](uint32_t j, const uint4& r, uint32_t) {
                tkey[j] = rec_key(r);
                tpay[j] = make_uint2(r.z, r.w);
            });
        const RadixTab tab{tkey, tpay, reinterpret_cast<const uint16_t*>(dirw), bits, bkt_shift, I.part};
        probe_rounds<kJoinThreads>(tab, recs + I.s_off, I.s_cnt, J.out, J.out_cap, &state[lo].rows, fmt, kbp, wsum,
                                   lmisc + 2);
        __syncthreads();
        if (tid == 0 && atomicSub(&state[lo].remaining, 1u) == 1u)
            state[lo].t_done = __builtin_amdgcn_s_memrealtime();
    }
}

// LDS bytes of the join kernel.
constexpr size_t join_lds_bytes(bool range)
{
    return kMiscWords * 4 +
           (range ? std::max<size_t>((size_t)kDenseCap * 8 + (kDenseCap + 31) / 32 * 4,
                                     (size_t)kCodeCap * 12 + (kTableBuckets / 2 + 1) * 4)
                  : (size_t)kTableCap * 16 + (kTableBuckets / 2 + 1) * 4);
}
static_assert(join_lds_bytes(false) <= kLdsBytes && join_lds_bytes(true) <= kLdsBytes, "join LDS budget");

// ======================================================================
// host side
// ======================================================================

static thread_local std::string g_err;

void set_error(const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

int fail(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

const char* last_error() { return g_err.c_str(); }

static int col_width(const pj_schema& s, uint32_t c)
{
    switch (s.type[c]) {
    case PJ_CT_INTEGER: return 4;
    case PJ_CT_LONG:
    case PJ_CT_DECIMAL:
    case PJ_CT_DATE:
    case PJ_CT_POINTER: return 8;
    case PJ_CT_CHAR: return (int)s.width[c];
    default: return -1;
    }
}

int schema_offsets(const pj_schema& s, int* off, int* size)
{
    if (s.ncols == 0 || s.ncols > PJ_MAX_COLS)
        return fail(PJ_ERR_INVALID, "schema must have 1..%d columns", PJ_MAX_COLS);
    int o = 0;
    for (uint32_t c = 0; c < s.ncols; ++c) {
        int w = col_width(s, c);
        if (w < 0 || (s.type[c] == PJ_CT_CHAR && w < 2))
            return fail(PJ_ERR_INVALID, "column %u: bad type %u / width %u", c, s.type[c], s.width[c]);
        off[c] = o;
        o += w;
    }
    *size = o;
    return PJ_OK;
}

static bool int_key(uint32_t t) { return t == PJ_CT_INTEGER || t == PJ_CT_LONG || t == PJ_CT_DATE; }

// The 16-byte record form: one INTEGER/LONG/DATE key column a side or two
// INTEGER columns, at most 8 payload bytes a side, output tuples of at most
// 32 bytes without CHAR columns, everything on 4-byte boundaries. Returns
// false when the descriptor needs the row-id form.
static bool record_layout(const pj_join_desc* d, const int* boff, const int* poff, int bsize, int psize, Layout* L)
{
    if (bsize % 4 || psize % 4)
        return false;
    SideLayout* sl[2] = {&L->build, &L->probe};
    const pj_schema* sc[2] = {&d->build, &d->probe};
    const int* so[2] = {boff, poff};
    const uint16_t* ja[2] = {d->buildjattr, d->probejattr};
    for (int side = 0; side < 2; ++side) {
        SideLayout& s = *sl[side];
        s.tw = (uint32_t)((side ? psize : bsize) / 4);
        s.tbytes = (uint32_t)(side ? psize : bsize);
        s.p0 = s.p1 = -1;
        if (d->nkeys == 1) {
            const uint32_t c = ja[side][0];
            if (so[side][c] % 4)
                return false;
            s.k0 = so[side][c] / 4;
            s.k1 = sc[side]->type[c] == PJ_CT_INTEGER ? -1 : s.k0 + 1;
        } else if (d->nkeys == 2) {
            const uint32_t a = ja[side][0], b = ja[side][1];
            if (sc[side]->type[a] != PJ_CT_INTEGER || sc[side]->type[b] != PJ_CT_INTEGER)
                return false;
            if (so[side][a] % 4 || so[side][b] % 4)
                return false;
            s.k0 = so[side][a] / 4;
            s.k1 = so[side][b] / 4;
        } else {
            return false;
        }
    }
    OutFormat& f = L->out;
    memset(&f, 0, sizeof(f));
    uint32_t words = 0;
    for (uint32_t j = 0; j < d->nproj; ++j) {
        const int side = d->proj_side[j];
        const pj_schema& s = *sc[side];
        const uint32_t c = d->proj_attr[j];
        const int w = col_width(s, c);
        if (s.type[c] == PJ_CT_CHAR)
            return false;
        const int nw = w / 4;
        if (words + nw > 8)
            return false;
        int keyidx = -1;
        for (uint32_t k = 0; k < d->nkeys; ++k)
            if (ja[side][k] == c)
                keyidx = (int)k;
        const uint8_t kbase = side ? SRC_PK0 : SRC_BK0, pbase = side ? SRC_PP0 : SRC_BP0;
        if (keyidx >= 0) {
            if (d->nkeys == 2)
                f.src[words++] = (uint8_t)(kbase + keyidx);
            else {
                f.src[words++] = kbase;
                if (nw == 2)
                    f.src[words++] = kbase + 1;
            }
            continue;
        }
        // a non-key column travels in the record's payload words
        SideLayout& sdl = *sl[side];
        if (so[side][c] % 4)
            return false;
        for (int q = 0; q < nw; ++q) {
            const int32_t word = so[side][c] / 4 + q;
            int slot;
            if (sdl.p0 == word) slot = 0;
            else if (sdl.p1 == word) slot = 1;
            else if (sdl.p0 < 0) { sdl.p0 = word; slot = 0; }
            else if (sdl.p1 < 0) { sdl.p1 = word; slot = 1; }
            else
                return false;
            f.src[words++] = (uint8_t)(pbase + slot);
        }
    }
    f.words = words;
    L->out_width = words * 4;
    for (int side = 0; side < 2; ++side) {
        SideLayout& s = *sl[side];
        s.ident = (s.tw == 4 && s.k0 == 0 && s.k1 == 1 && s.p0 == 2 && s.p1 == 3) ? 1u : 0u;
    }
    return true;
}

// The row-id form: records {key hash, row}, the join's pairs checked and
// projected on the tuples by the materialising pass.
static int rowid_layout(const pj_join_desc* d, const int* boff, const int* poff, int bsize, int psize, Layout* L)
{
    if (bsize > 65535 || psize > 65535)
        return fail(PJ_ERR_UNSUPPORTED, "device path takes tuples of at most 65535 bytes");
    const pj_schema* sc[2] = {&d->build, &d->probe};
    const int* so[2] = {boff, poff};
    const uint16_t* ja[2] = {d->buildjattr, d->probejattr};
    SideLayout* sl[2] = {&L->build, &L->probe};
    for (int side = 0; side < 2; ++side) {
        SideLayout& s = *sl[side];
        s = SideLayout();
        s.tw = 0;
        s.k0 = s.k1 = s.p0 = s.p1 = -1;
        s.ident = 0;
        s.hashed = 1;
        s.tbytes = (uint32_t)(side ? psize : bsize);
        s.nk = d->nkeys;
        s.hmask = (d->flags & PJ_FLAG_NARROW_ROW_HASH) ? 0x3FFull : ~0ull;
        for (uint32_t k = 0; k < d->nkeys; ++k) {
            s.koff[k] = (uint16_t)so[side][ja[side][k]];
            s.kint[k] = sc[side]->type[ja[side][k]] == PJ_CT_INTEGER ? 1 : 0;
        }
    }
    // the join's rows are (build row, probe row) pairs
    memset(&L->out, 0, sizeof(L->out));
    L->out.words = 2;
    L->out.src[0] = SRC_BP0;
    L->out.src[1] = SRC_PP0;
    RowSpec& r = L->rows;
    r = RowSpec();
    r.nk = d->nkeys;
    r.bw = (uint32_t)bsize;
    r.pw = (uint32_t)psize;
    bool words = bsize % 4 == 0 && psize % 4 == 0;
    for (uint32_t k = 0; k < d->nkeys; ++k) {
        r.key[k].boff = L->build.koff[k];
        r.key[k].poff = L->probe.koff[k];
        r.key[k].bint = L->build.kint[k];
        r.key[k].pint = L->probe.kint[k];
    }
    uint32_t dst = 0;
    for (uint32_t j = 0; j < d->nproj; ++j) {
        const int side = d->proj_side[j];
        const uint32_t c = d->proj_attr[j];
        const int w = col_width(*sc[side], c);
        RowCopy& q = r.copy[j];
        q.side = (uint8_t)side;
        q.chr = sc[side]->type[c] == PJ_CT_CHAR ? 1 : 0;
        q.src = (uint16_t)so[side][c];
        q.dst = (uint16_t)dst;
        q.width = (uint16_t)w;
        words = words && q.src % 4 == 0 && w % 4 == 0;
        dst += (uint32_t)w;
    }
    if (dst > 65535)
        return fail(PJ_ERR_UNSUPPORTED, "device path emits output tuples of at most 65535 bytes");
    r.ncopy = d->nproj;
    r.ow = dst;
    r.words = words && dst % 4 == 0 ? 1u : 0u;
    L->out_width = dst;
    L->rowid = 1;
    return PJ_OK;
}

// Mirrors the checks and schema construction of JoinOp::init and
// HashJoinOp::init (operators/join.cpp:224-326, 367-461): the key columns
// compare as the reference's ConjunctionEqualsEvaluator does
// (conjunctionevaluator.cpp:83-88 over rawcompfns.cpp:48-180, INTEGER and
// LONG by value, DATE only with DATE), then picks the record form or the
// row-id form.
int derive_layout(const pj_join_desc* d, Layout* L)
{
    int boff[PJ_MAX_COLS], poff[PJ_MAX_COLS], bsize, psize;
    int rc;
    if ((rc = schema_offsets(d->build, boff, &bsize)) || (rc = schema_offsets(d->probe, poff, &psize)))
        return rc;
    if (d->nkeys == 0 || d->nkeys > PJ_MAX_KEYS)
        return fail(PJ_ERR_INVALID, "buildjattr/probejattr must name 1..%d columns", PJ_MAX_KEYS);
    for (uint32_t k = 0; k < d->nkeys; ++k)
        if (d->buildjattr[k] >= d->build.ncols || d->probejattr[k] >= d->probe.ncols)
            return fail(PJ_ERR_INVALID, "join attribute %u out of range", k);
    if (d->nproj == 0 || d->nproj > PJ_MAX_COLS)
        return fail(PJ_ERR_INVALID, "projection must have 1..%d entries", PJ_MAX_COLS);
    for (uint32_t j = 0; j < d->nproj; ++j) {
        const pj_schema& s = d->proj_side[j] == PJ_SIDE_BUILD ? d->build : d->probe;
        if (d->proj_side[j] > PJ_SIDE_PROBE || d->proj_attr[j] >= s.ncols)
            return fail(PJ_ERR_INVALID, "projection entry %u out of range", j);
    }
    for (uint32_t k = 0; k < d->nkeys; ++k) {
        const uint32_t bt = d->build.type[d->buildjattr[k]], pt = d->probe.type[d->probejattr[k]];
        if (!int_key(bt) || !int_key(pt))
            return fail(PJ_ERR_UNSUPPORTED, "device path joins INTEGER, LONG or DATE keys");
        if ((bt == PJ_CT_DATE) != (pt == PJ_CT_DATE))
            return fail(PJ_ERR_UNSUPPORTED, "DATE keys join only DATE keys");
    }
    *L = Layout();
    if (record_layout(d, boff, poff, bsize, psize, L))
        return PJ_OK;
    *L = Layout();
    return rowid_layout(d, boff, poff, bsize, psize, L);
}

// The device memory the library's buffers hold, and the process's limit on
// it (0: none). A buffer that would pass the limit fails as a failed
// hipMalloc does.
std::atomic<uint64_t> g_dev_used{0}, g_dev_limit{0};

int DevBuf::ensure(size_t bytes)
{
    if (bytes <= cap)
        return PJ_OK;
    size_t want = bytes < 256 ? 256 : bytes;
    if (arena) {
        const size_t at = (arena->used + 255) / 256 * 256;
        if (at > arena->size || want > arena->size - at)
            return fail(PJ_ERR_UNSUPPORTED, "the fixed scratch of %zu bytes cannot hold %zu more bytes", arena->size,
                        want);
        p = arena->base + at;
        cap = want;
        arena->used = at + want;
        return PJ_OK;
    }
    release();
    const uint64_t lim = g_dev_limit.load();
    if (lim && g_dev_used.load() + want > lim)
        return fail(PJ_ERR_DEVICE, "a device buffer of %zu bytes would pass the limit of %llu bytes (%llu held)",
                    want, (unsigned long long)lim, (unsigned long long)g_dev_used.load());
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        p = nullptr;
        return fail(PJ_ERR_DEVICE, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    cap = want;
    g_dev_used += want;
    return PJ_OK;
}

void DevBuf::release()
{
    if (p && !arena) {
        (void)hipFree(p);
        g_dev_used -= cap;
    }
    p = nullptr;
    cap = 0;
}

}  // namespace pj

// ======================================================================
// handle and orchestration
// ======================================================================

using namespace pj;

namespace {

// One relation partitioned on the device.
struct Partitioned {
    uint32_t F = 1;
    uint64_t n = 0;
    std::vector<uint32_t> start;   // F + 1 partition starts (host copy)
};

struct CursorImpl {
    std::vector<uint8_t> rows;
    uint64_t next = 0, total = 0;
};

}  // namespace

struct pj_join_s {
    pj_join_desc desc;
    Layout lay;
    std::mutex mu;            // serialises device work on this handle
    std::mutex stage_mu;      // guards staged build tuples
    std::vector<uint8_t> staged;
    bool built = false;
    DevBuf d_raw_build, d_raw_probe, d_out_batch;
    DevBuf rpart, spart, hist, bsum, rstarts, sstarts, planoffs, counters;
    Partitioned R, S;
    uint32_t bR = 0, bS = 0;  // radix bits of R and of S (bR >= bS)
    // range partitions chosen by the learned model
    bool range = false;
    RangePlan plan;
    LearnedCtx lctx;
    DevBuf map_dev, pminmax, pids, plan_parts, coarse_of;
    std::vector<uint64_t> pmin, pmax;   // per build partition, ordered keys
    const void* raw_build = nullptr;    // device build tuples of the current run (radix re-run)
    uint64_t raw_nbuild = 0;
    float model_ms = 0.f;
    pj_timings last;
    hipEvent_t ev[10];   // 0-5 phases, 6-7 the model, 8-9 after each side's histogram
    bool ev_ok = false;
    uint32_t* host_ctr = nullptr;   // pinned: the join's counters, read back after each probe
    DevBuf gj_count, gj_rows;       // group-join: matches a build record, merged rows
    DevBuf pairs;                   // row-id form: the join's (build row, probe row) pairs
    // the probe side's appended layout: chunk tags, counts, chunk list
    DevBuf a_cmap, a_bhist, a_bsum, a_nchunks, a_clist;
    bool s_app = false;             // the last probe side took the appended layout
    // the probe side's second radix pass: fine counts, their scan, the
    // records in fine partitions and their starts
    DevBuf t_fine, t_fsum, t_recs, t_starts;
    bool s_two = false;             // the last probe side took two passes
    uint64_t pair_cap = 0;          // pairs the last probe phase could hold
    uint64_t pair_want = 0;         // pairs a previous run needed
};

constexpr uint32_t kMaxBits = 14;

uint32_t pj::choose_build_bits(uint64_t nr, int32_t forced)
{
    if (forced >= 0)
        return (uint32_t)std::min<int32_t>(forced, (int32_t)kMaxBits);
    for (uint32_t b = 0; b < kMaxBits; ++b) {
        double avg = (double)nr / (double)(1ull << b);
        if (avg + 5.0 * std::sqrt(avg) + 16.0 <= (double)kTableCap)
            return b;
    }
    return kMaxBits;
}

namespace {

constexpr uint32_t kMaxCoarse = 2048;   // probe partitions of a range plan (write-combining fan-out)

uint32_t scatter_tile(uint32_t F, bool range)
{
    const int64_t avail = (int64_t)kLdsBytes - 256 - (int64_t)(2 * F + 1) * 4 - 16 * 4;
    int64_t ts = avail / (range ? 18 : 16);
    ts = (ts / kScatterThreads) * kScatterThreads;
    ts = std::min<int64_t>(ts, (int64_t)kScatterThreads * kScatterItems);
    return ts > 0 ? (uint32_t)ts : 0;
}

// Records per write-combining chunk for a fan-out of F buffered by one block:
// the largest power of two (at most 16) whose buffers fit in 128 KiB of LDS;
// 1 selects the tile-sorted scatter (no padding).
uint32_t wc_chunk(uint32_t F)
{
    if (F <= 1)
        return 1;
    uint32_t B = 16;
    while (B > 1 && ((uint64_t)F * B * 16 > 131072 || (uint64_t)F * B * 16 + 8ull * F > kLdsBytes))
        B >>= 1;
    return B < 2 ? 1 : B;
}

// Blocks an input chunk of the write-combining scatter: two when one block
// could only buffer chunks of fewer than 8 records (64 bytes), which HBM
// writes at half the rate of whole 128-byte lines, and two blocks of half
// the partitions each reach 8.
uint32_t wc_split(uint32_t F)
{
    return F > 1 && wc_chunk(F) < 8 && wc_chunk((F + 1) / 2) >= 8 ? 2u : 1u;
}

// Blocks of a tile-sorted partition pass: about 64K tuples each for large
// relations, and for small ones enough blocks of at least 2K tuples to
// spread over the CUs (one block would do a 64K-tuple relation alone).
uint32_t hist_blocks(uint64_t n)
{
    uint64_t b = (n + 65535) / 65536;
    b = std::max<uint64_t>(b, std::min<uint64_t>((n + 2047) / 2048, 256));
    return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(b, 1), 2048);
}

// A key whose top bit is set: its partition is never 0 (for bits >= 1), so
// it pads partition 0; key 0 mixes to 0 and pads every other partition.
uint64_t pad_key_nonzero_part()
{
    for (uint64_t k = 1;; ++k)
        if (pj_fin64(k) >> 63)
            return k;
}

// Lets kernel fn use `bytes` of dynamic LDS. hipFuncSetAttribute costs
// microseconds of host time and a small join launches about ten kernels, so
// the largest size set so far is remembered per device and kernel.
int allow_lds(const void* fn, size_t bytes)
{
    static std::mutex mu;
    static std::map<std::pair<int, const void*>, size_t> done;
    int dev = 0;
    PJ_HIP(hipGetDevice(&dev));
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = done.find({dev, fn});
        if (it != done.end() && it->second >= bytes)
            return PJ_OK;
    }
    PJ_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    std::lock_guard<std::mutex> g(mu);
    size_t& v = done[{dev, fn}];
    v = std::max(v, bytes);
    return PJ_OK;
}

template <int RM, int B, bool RANGE, int SPLIT, bool APP>
int launch_wc_rm(const uint32_t* in, uint64_t n, const SideLayout& L, const PartFn& pf, uint64_t chunk,
                 uint32_t nblk, const uint32_t* offs, uint4* out, uint4 pad_other, uint4 pad_zero, const uint16_t* pid,
                 const AppendOut& app, size_t lds, hipStream_t st)
{
    if (int rc = allow_lds((const void*)k_scatter_wc<RM, B, RANGE, SPLIT, APP>, lds))
        return rc;
    hipLaunchKernelGGL((k_scatter_wc<RM, B, RANGE, SPLIT, APP>), dim3(nblk * SPLIT), dim3(kWcThreads), lds, st, in, n,
                       L, pf, chunk, nblk, offs, out, pad_other, pad_zero, pid, app);
    PJ_HIP(hipGetLastError());
    return PJ_OK;
}

template <int B, bool RANGE, int SPLIT, bool APP>
int launch_wc_t(int rm, const uint32_t* in, uint64_t n, const SideLayout& L, const PartFn& pf, uint64_t chunk,
                uint32_t nblk, const uint32_t* offs, uint4* out, const uint16_t* pid, const AppendOut& app,
                hipStream_t st)
{
    const uint32_t FL = SPLIT == 1 ? pf.F : (pf.F + 1) / 2;
    if (SPLIT > 1 && nblk % 8)
        return fail(PJ_ERR_INVALID, "split scatter needs a multiple of 8 input chunks");
    const size_t lds = (size_t)FL * B * 16 + 2 * FL * 4 + (APP ? (size_t)FL * 4 + 16 : 0);
    uint4 pad_other, pad_zero;
    if (RANGE) {
        // ordered key 0 lies in partition 0, ordered key ~0 in the last one
        pad_other = make_uint4(0, 0x80000000u, 0, 0);
        pad_zero = make_uint4(0xFFFFFFFFu, 0x7FFFFFFFu, 0, 0);
    } else {
        const uint64_t pk = pad_key_nonzero_part();
        pad_other = make_uint4(0, 0, 0, 0);
        pad_zero = make_uint4((uint32_t)pk, (uint32_t)(pk >> 32), 0, 0);
    }
    if (rm == REC_IDENT)
        return launch_wc_rm<REC_IDENT, B, RANGE, SPLIT, APP>(in, n, L, pf, chunk, nblk, offs, out, pad_other, pad_zero,
                                                             pid, app, lds, st);
    if (RANGE || rm == REC_WORDS)
        return launch_wc_rm<REC_WORDS, B, RANGE, SPLIT, APP>(in, n, L, pf, chunk, nblk, offs, out, pad_other,
                                                             pad_zero, pid, app, lds, st);
    return launch_wc_rm<REC_HASH, B, false, SPLIT, APP>(in, n, L, pf, chunk, nblk, offs, out, pad_other, pad_zero,
                                                        pid, app, lds, st);
}

template <int B>
int launch_wc(int rm, bool range, uint32_t split, const uint32_t* in, uint64_t n, const SideLayout& L,
              const PartFn& pf, uint64_t chunk, uint32_t nblk, const uint32_t* offs, uint4* out, const uint16_t* pid,
              hipStream_t st)
{
    const AppendOut none = {};
    if (split > 1)
        return range ? launch_wc_t<B, true, 2, false>(rm, in, n, L, pf, chunk, nblk, offs, out, pid, none, st)
                     : launch_wc_t<B, false, 2, false>(rm, in, n, L, pf, chunk, nblk, offs, out, nullptr, none, st);
    return range ? launch_wc_t<B, true, 1, false>(rm, in, n, L, pf, chunk, nblk, offs, out, pid, none, st)
                 : launch_wc_t<B, false, 1, false>(rm, in, n, L, pf, chunk, nblk, offs, out, nullptr, none, st);
}

template <int RM, bool RANGE>
int launch_hist_rm(const uint32_t* in, uint64_t n, const SideLayout& L, const PartFn& pf, uint64_t chunk,
                   uint32_t nblk, uint32_t pad, uint32_t* hist, unsigned long long* pmin, unsigned long long* pmax,
                   uint16_t* pid, size_t lds, hipStream_t st)
{
    if (int rc = allow_lds((const void*)k_hist<RM, RANGE>, lds))
        return rc;
    hipLaunchKernelGGL((k_hist<RM, RANGE>), dim3(nblk), dim3(kHistThreads), lds, st, in, n, L, pf, chunk, nblk, pad,
                       hist, pmin, pmax, pid);
    PJ_HIP(hipGetLastError());
    return PJ_OK;
}

template <bool RANGE>
int launch_hist(int rm, const uint32_t* in, uint64_t n, const SideLayout& L, const PartFn& pf, uint64_t chunk,
                uint32_t nblk, uint32_t pad, uint32_t* hist, unsigned long long* pmin, unsigned long long* pmax,
                uint16_t* pid, hipStream_t st)
{
    const size_t lds = (size_t)pf.F * 4 + (RANGE ? kMapBytes : 0) + (pmin ? (size_t)pf.F * 16 : 0);
    if (lds > kLdsBytes)
        return fail(PJ_ERR_UNSUPPORTED, "%u partitions exceed the histogram's LDS budget", pf.F);
    if (rm == REC_IDENT)
        return launch_hist_rm<REC_IDENT, RANGE>(in, n, L, pf, chunk, nblk, pad, hist, pmin, pmax, pid, lds, st);
    if (RANGE || rm == REC_WORDS)
        return launch_hist_rm<REC_WORDS, RANGE>(in, n, L, pf, chunk, nblk, pad, hist, pmin, pmax, pid, lds, st);
    return launch_hist_rm<REC_HASH, false>(in, n, L, pf, chunk, nblk, pad, hist, pmin, pmax, pid, lds, st);
}

template <int RM, bool RANGE>
int launch_tile_rm(const uint32_t* in, uint64_t n, const SideLayout& L, const PartFn& pf, uint64_t chunk,
                   uint32_t nblk, const uint32_t* offs, uint4* out, uint32_t ts, const uint16_t* pid, size_t lds,
                   hipStream_t st)
{
    if (int rc = allow_lds((const void*)k_scatter<RM, RANGE>, lds))
        return rc;
    hipLaunchKernelGGL((k_scatter<RM, RANGE>), dim3(nblk), dim3(kScatterThreads), lds, st, in, n, L, pf, chunk, nblk,
                       offs, out, ts, pid);
    PJ_HIP(hipGetLastError());
    return PJ_OK;
}

template <bool RANGE>
int launch_tile_scatter(int rm, const uint32_t* in, uint64_t n, const SideLayout& L, const PartFn& pf,
                        uint64_t chunk, uint32_t nblk, const uint32_t* offs, uint4* out, const uint16_t* pid,
                        hipStream_t st)
{
    const uint32_t ts = scatter_tile(pf.F, RANGE);
    if (ts == 0)
        return fail(PJ_ERR_UNSUPPORTED, "%u partitions exceed the scatter's LDS budget", pf.F);
    const size_t lds = (size_t)ts * (RANGE ? 18 : 16) + (2 * pf.F + 1) * 4 + 16 * 4;
    if (rm == REC_IDENT)
        return launch_tile_rm<REC_IDENT, RANGE>(in, n, L, pf, chunk, nblk, offs, out, ts, pid, lds, st);
    if (RANGE || rm == REC_WORDS)
        return launch_tile_rm<REC_WORDS, RANGE>(in, n, L, pf, chunk, nblk, offs, out, ts, pid, lds, st);
    return launch_tile_rm<REC_HASH, false>(in, n, L, pf, chunk, nblk, offs, out, ts, pid, lds, st);
}

// Blocks of the partition pass: about 512K tuples each, at least one per CU.
uint32_t part_blocks(uint64_t n)
{
    uint64_t b = n >> 19;
    b = std::max<uint64_t>(b, 256);
    b = std::min<uint64_t>(b, 2048);
    return (uint32_t)std::max<uint64_t>(std::min<uint64_t>(b, (n + 255) / 256), 1);
}

// Partitions raw tuples of one side into dst (canonical records, each
// block's share of a partition padded to whole write-combining chunks).
// With range = true the partition function is pf's boundaries; pminmax (2F
// ordered keys, build side only) then receives each partition's key range,
// and a write-combining scatter takes each tuple's partition from the
// histogram pass through h->pids instead of searching again.
// The shape of one partition pass over n tuples: blocks an input chunk
// (split), records a write-combining chunk (B; 1 selects the tile-sorted
// scatter), input chunks (nblk), the histogram's counters (m) and scan
// segments (nsb), and the records the partitioned copy can hold (cap).
struct PartGeom {
    uint32_t split, B, nblk, nsb;
    uint64_t chunk, m, cap;
};

PartGeom part_geom(uint64_t n, const PartFn& pf, bool range, bool write_combine)
{
    PartGeom g;
    const uint32_t F = pf.F;
    g.split = write_combine ? wc_split(F) : 1;
    g.B = write_combine ? wc_chunk(g.split > 1 ? (F + 1) / 2 : F) : 1;
    g.nblk = g.B > 1 ? part_blocks(n) : hist_blocks(n);
    if (g.split > 1)   // input chunks pair up blocks b and b + 8 (one XCD)
        g.nblk = (g.nblk + 7) / 8 * 8;
    g.chunk = (n + g.nblk - 1) / g.nblk;
    g.m = (uint64_t)F * g.nblk;
    g.nsb = (uint32_t)((g.m + kScanSeg - 1) / kScanSeg);
    g.cap = n + (uint64_t)F * g.nblk * (g.B - 1);
    return g;
}

// mark (optional) is recorded on st once the histogram and its scan are
// enqueued, before the scatter.
int partition_bufs(DevBuf& histb, DevBuf& bsumb, DevBuf& pidsb, const void* d_in, uint64_t n, const SideLayout& L,
                   const PartFn& pf, bool range, bool write_combine, DevBuf& dst, Partitioned& P, DevBuf& starts,
                   bool host_copy, unsigned long long* pminmax, hipStream_t st, hipEvent_t mark = nullptr)
{
    int rc;
    const uint32_t F = pf.F;
    P.F = F;
    P.n = n;
    P.start.assign(host_copy ? F + 1 : 0, 0);
    if ((rc = starts.ensure((F + 1) * 4)))
        return rc;
    if (n == 0) {   // every partition empty
        if (!host_copy) {
            hipLaunchKernelGGL(k_fill_u32, dim3((F + 1 + 255) / 256), dim3(256), 0, st, (uint32_t*)starts.p, F + 1, 0u);
            PJ_HIP(hipGetLastError());
        }
        if (mark)
            PJ_HIP(hipEventRecord(mark, st));
        return PJ_OK;
    }
    if (n >= (1ull << 31))
        return fail(PJ_ERR_UNSUPPORTED, "relations of 2^31 tuples or more");
    const PartGeom g = part_geom(n, pf, range, write_combine);
    const uint32_t split = g.split, B = g.B, nblk = g.nblk, nsb = g.nsb;
    const uint64_t chunk = g.chunk, m = g.m, cap = g.cap;
    if (cap >= (1ull << 32))
        return fail(PJ_ERR_UNSUPPORTED, "partitioned relation exceeds 2^32 records");
    if ((rc = dst.ensure(cap * 16)) || (rc = histb.ensure(m * 4)) || (rc = bsumb.ensure((nsb + 1) * 4)))
        return rc;
    const uint32_t* in = (const uint32_t*)d_in;
    if (L.hashed && range)
        return fail(PJ_ERR_UNSUPPORTED, "row-id records take radix partitions only");
    const int rm = L.hashed ? REC_HASH : (L.ident && ((uintptr_t)d_in % 16 == 0)) ? REC_IDENT : REC_WORDS;
    uint32_t* hist = (uint32_t*)histb.p;
    uint32_t* bsum = (uint32_t*)bsumb.p;
    unsigned long long* pmin = pminmax;
    unsigned long long* pmax = pminmax ? pminmax + F : nullptr;
    if (pminmax) {
        // pmin starts at ~0 (all bytes 0xFF), pmax at 0
        PJ_HIP(hipMemsetAsync(pmin, 0xFF, (size_t)F * 8, st));
        PJ_HIP(hipMemsetAsync(pmax, 0, (size_t)F * 8, st));
    }
    uint16_t* pid = nullptr;   // range partitions: each tuple's, from the histogram to the scatter
    if (range) {
        if (F > 65536 || !pf.map)
            return fail(PJ_ERR_INVALID, "range partitions need the learned function and at most 65536 partitions");
        if ((rc = pidsb.ensure(n * 2)))
            return rc;
        pid = (uint16_t*)pidsb.p;
    }
    rc = range ? launch_hist<true>(rm, in, n, L, pf, chunk, nblk, B, hist, pmin, pmax, pid, st)
               : launch_hist<false>(rm, in, n, L, pf, chunk, nblk, B, hist, nullptr, nullptr, nullptr, st);
    if (rc)
        return rc;
    if (m <= kScanSmall) {   // one launch for small histograms
        hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(kScanSmallThreads), 0, st, hist, (uint32_t)m, F, nblk,
                           (uint32_t*)starts.p);
    } else {
        hipLaunchKernelGGL(k_scan_reduce, dim3(nsb), dim3(kScanThreads), 0, st, hist, m, bsum);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, bsum, nsb);
        hipLaunchKernelGGL(k_scan_down, dim3(nsb), dim3(kScanThreads), 0, st, hist, m, (const uint32_t*)bsum);
        hipLaunchKernelGGL(k_part_bounds, dim3((F + 1 + 255) / 256), dim3(256), 0, st, hist, F, nblk,
                           (const uint32_t*)(bsum + nsb), (uint32_t*)starts.p);
    }
    PJ_HIP(hipGetLastError());
    if (mark)
        PJ_HIP(hipEventRecord(mark, st));
    uint4* out = (uint4*)dst.p;
    switch (B) {
    case 16: rc = launch_wc<16>(rm, range, split, in, n, L, pf, chunk, nblk, hist, out, pid, st); break;
    case 8: rc = launch_wc<8>(rm, range, split, in, n, L, pf, chunk, nblk, hist, out, pid, st); break;
    case 4: rc = launch_wc<4>(rm, range, split, in, n, L, pf, chunk, nblk, hist, out, pid, st); break;
    case 2: rc = launch_wc<2>(rm, range, split, in, n, L, pf, chunk, nblk, hist, out, pid, st); break;
    default:
        rc = range ? launch_tile_scatter<true>(rm, in, n, L, pf, chunk, nblk, hist, out, pid, st)
                   : launch_tile_scatter<false>(rm, in, n, L, pf, chunk, nblk, hist, out, nullptr, st);
    }
    if (rc)
        return rc;
    if (host_copy)
        PJ_HIP(hipMemcpyAsync(P.start.data(), starts.p, (F + 1) * 4, hipMemcpyDeviceToHost, st));
    return PJ_OK;
}

// A scatter block's lanes own its partitions, one each: at most 1024 a
// block, 2048 over a split pair.
constexpr uint32_t kAppendMaxF = 2 * kWcThreads;
static_assert(kAppendMaxF <= kPlaceMaxF, "the chunk list's tile sort holds every partition");

// Whether the probe side takes the appended layout: radix partitions whose
// build partitions are the probe partitions (the join reads each probe
// partition once: listed chunks cost the join a record in flight a thread,
// paid on every read), at least two partitions (one has no key outside it to
// pad chunks with) and at most kAppendMaxF, fewer than 2^30 tuples (slot
// numbers in 32 bits), and not in a fixed scratch (it needs about twice the
// room: each block of a split pair reserves slots for its whole input
// chunk).
bool want_append(const pj_join_s* h, const void* d_probe, uint64_t np)
{
    (void)d_probe;
    if (np == 0 || np >= (1ull << 30) || h->range || h->bS == 0 || (1u << h->bS) > kAppendMaxF ||
        h->bR != h->bS || (h->desc.flags & PJ_FLAG_LAYOUT_EXACT) || h->spart.arena)
        return false;
    return (h->desc.flags & PJ_FLAG_LAYOUT_APPEND) || np >= kAppendMin;
}

// Partitions the probe side into the appended layout (radix partitions):
// the write-combining scatter (8-record chunks) appends each block's chunks
// without a histogram pass before it, then the chunk tags are counted,
// scanned and placed into h->a_clist, partition by partition; sstarts[p] is
// partition p's first record in the list's order (8 a chunk). Marks event 9
// before the scatter.
int partition_append(pj_join_s* h, const void* d_in, uint64_t n, const SideLayout& L, const PartFn& pf,
                     hipStream_t st)
{
    int rc;
    const uint32_t F = pf.F;
    constexpr uint32_t B = kChunkRecs;
    Partitioned& P = h->S;
    P.F = F;
    P.n = n;
    P.start.clear();
    if (n >= (1ull << 31))
        return fail(PJ_ERR_UNSUPPORTED, "relations of 2^31 tuples or more");
    if (F > kAppendMaxF || F < 2)
        return fail(PJ_ERR_UNSUPPORTED, "the appended layout takes 2 to %u partitions", kAppendMaxF);
    const uint32_t split = F > 1024 ? 2u : 1u;   // 8-record buffers of at most 1024 partitions a block
    const uint32_t FL = split > 1 ? (F + 1) / 2 : F;
    uint32_t nblk = part_blocks(n);
    if (split > 1)
        nblk = (nblk + 7) / 8 * 8;
    const uint64_t chunk = (n + nblk - 1) / nblk;
    const uint32_t NB = nblk * split;
    const uint64_t region = (chunk + B - 1) / B + FL;   // the most chunks a block can write
    const uint64_t slots = (uint64_t)NB * region;
    const uint64_t m = (uint64_t)F * NB;
    const uint32_t nsb = (uint32_t)((m + kScanSeg - 1) / kScanSeg);
    if (slots * B >= (1ull << 32) || m >= (1ull << 32))
        return fail(PJ_ERR_UNSUPPORTED, "appended layout exceeds 2^32 records");
    if ((rc = h->sstarts.ensure((F + 1) * 4)) || (rc = h->spart.ensure(slots * B * 16)) ||
        (rc = h->a_cmap.ensure(slots * 2)) || (rc = h->a_bhist.ensure(m * 4)) ||
        (rc = h->a_bsum.ensure((nsb + 1) * 4)) || (rc = h->a_nchunks.ensure((size_t)NB * 4)) ||
        (rc = h->a_clist.ensure(slots * 4)))
        return rc;
    const uint32_t* in = (const uint32_t*)d_in;
    const int rm = L.hashed ? REC_HASH : (L.ident && ((uintptr_t)d_in % 16 == 0)) ? REC_IDENT : REC_WORDS;
    if (h->ev_ok && (h->desc.flags & PJ_FLAG_TIMING))
        PJ_HIP(hipEventRecord(h->ev[9], st));
    AppendOut app;
    app.cmap = (uint16_t*)h->a_cmap.p;
    app.bhist = (uint32_t*)h->a_bhist.p;
    app.nchunks = (uint32_t*)h->a_nchunks.p;
    app.region = region;
    uint4* out = (uint4*)h->spart.p;
    rc = split > 1 ? launch_wc_t<B, false, 2, true>(rm, in, n, L, pf, chunk, nblk, nullptr, out, nullptr, app, st)
                   : launch_wc_t<B, false, 1, true>(rm, in, n, L, pf, chunk, nblk, nullptr, out, nullptr, app, st);
    if (rc)
        return rc;
    uint32_t* bh = (uint32_t*)h->a_bhist.p;
    uint32_t* bsum = (uint32_t*)h->a_bsum.p;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nsb), dim3(kScanThreads), 0, st, bh, m, bsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, bsum, nsb);
    hipLaunchKernelGGL(k_scan_down, dim3(nsb), dim3(kScanThreads), 0, st, bh, m, (const uint32_t*)bsum);
    hipLaunchKernelGGL(k_chunk_place, dim3(NB), dim3(kPlaceThreads), 0, st, (const uint16_t*)h->a_cmap.p,
                       (const uint32_t*)h->a_nchunks.p, (const uint32_t*)bh, F, region, (uint32_t*)h->a_clist.p);
    hipLaunchKernelGGL(k_chunk_bounds, dim3((F + 1 + 255) / 256), dim3(256), 0, st, (const uint32_t*)bh, F, NB,
                       (const uint32_t*)(bsum + nsb), (uint32_t*)h->sstarts.p);
    PJ_HIP(hipGetLastError());
    return PJ_OK;
}

// Probe sides of at least this many tuples whose build partitions are at
// least 4 times finer than their probe partitions take two passes by
// default (PJ_FLAG_LAYOUT_EXACT / _TWO_PASS force one or two). The second
// pass reads and writes the probe side once; it saves the join D - 1 reads
// of it. Measured on 16M x 256M uniform: 2 build partitions a probe
// partition 5.65 ms a step in one pass, 7.00 in two; 4 of them 7.68 in one,
// 7.34 in two.
constexpr uint64_t kTwoPassMin = 1ull << 22;
constexpr uint32_t kTwoPassMinBits = 2;

// Whether the probe side takes two radix passes: radix partitions, 2 to 8
// fine partitions a probe partition (bR - bS of 1 to 3 bits; the planner
// takes at most 2^14 build partitions), not in a fixed scratch (the second
// pass needs room for another copy).
bool want_two_pass(const pj_join_s* h, uint64_t np)
{
    if (np == 0 || h->range || h->bR <= h->bS || h->bR - h->bS > 3 || (h->desc.flags & PJ_FLAG_LAYOUT_EXACT) ||
        h->spart.arena)
        return false;
    return (h->desc.flags & PJ_FLAG_LAYOUT_TWO_PASS) ||
           (np >= kTwoPassMin && h->bR - h->bS >= kTwoPassMinBits);
}

// Partitions the probe side in two radix passes (k_subsplit): one histogram
// over the 2^bR fine partitions gives the first pass's per-chunk counts
// (k_coarsen) and the second pass's places; h->t_recs then holds fine
// partition f's records at [t_starts[f], t_starts[f + 1]), no padding.
// Marks event 9 after the histogram and its scans.
int partition_two_pass(pj_join_s* h, const void* d_in, uint64_t n, const SideLayout& L, hipStream_t st)
{
    int rc;
    const uint32_t bS = h->bS, dbits = h->bR - h->bS;
    const PartFn pf = radix_fn(bS), pff = radix_fn(h->bR);
    const uint32_t F = pf.F, FF = pff.F;
    Partitioned& P = h->S;
    P.F = FF;
    P.n = n;
    P.start.clear();
    if (n >= (1ull << 31))
        return fail(PJ_ERR_UNSUPPORTED, "relations of 2^31 tuples or more");
    const PartGeom g = part_geom(n, pf, false, true);
    if (g.B != 8)
        return fail(PJ_ERR_UNSUPPORTED, "two passes take a first pass of 8-record chunks (%u partitions)", F);
    const uint64_t mf = (uint64_t)FF * g.nblk;
    const uint32_t nsf = (uint32_t)((mf + kScanSeg - 1) / kScanSeg);
    if (g.cap >= (1ull << 32) || mf >= (1ull << 32))
        return fail(PJ_ERR_UNSUPPORTED, "partitioned relation exceeds 2^32 records");
    if ((rc = h->spart.ensure(g.cap * 16)) || (rc = h->hist.ensure(g.m * 4)) ||
        (rc = h->bsum.ensure(((size_t)g.nsb + 1) * 4)) || (rc = h->t_fine.ensure(mf * 4)) ||
        (rc = h->t_fsum.ensure(((size_t)nsf + 1) * 4)) || (rc = h->t_recs.ensure(n * 16)) ||
        (rc = h->t_starts.ensure(((size_t)FF + 1) * 4)))
        return rc;
    const uint32_t* in = (const uint32_t*)d_in;
    const int rm = L.hashed ? REC_HASH : (L.ident && ((uintptr_t)d_in % 16 == 0)) ? REC_IDENT : REC_WORDS;
    uint32_t* fine = (uint32_t*)h->t_fine.p;
    uint32_t* fsum = (uint32_t*)h->t_fsum.p;
    uint32_t* coarse = (uint32_t*)h->hist.p;
    uint32_t* csum = (uint32_t*)h->bsum.p;
    if ((rc = launch_hist<false>(rm, in, n, L, pff, g.chunk, g.nblk, 1, fine, nullptr, nullptr, nullptr, st)))
        return rc;
    hipLaunchKernelGGL(k_coarsen, dim3((uint32_t)((g.m + 255) / 256)), dim3(256), 0, st, (const uint32_t*)fine, F,
                       g.nblk, dbits, g.B, coarse);
    hipLaunchKernelGGL(k_scan_reduce, dim3(g.nsb), dim3(kScanThreads), 0, st, coarse, g.m, csum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, csum, g.nsb);
    hipLaunchKernelGGL(k_scan_down, dim3(g.nsb), dim3(kScanThreads), 0, st, coarse, g.m, (const uint32_t*)csum);
    hipLaunchKernelGGL(k_scan_reduce, dim3(nsf), dim3(kScanThreads), 0, st, fine, mf, fsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, fsum, nsf);
    hipLaunchKernelGGL(k_scan_down, dim3(nsf), dim3(kScanThreads), 0, st, fine, mf, (const uint32_t*)fsum);
    hipLaunchKernelGGL(k_part_bounds, dim3((FF + 1 + 255) / 256), dim3(256), 0, st, (const uint32_t*)fine, FF, g.nblk,
                       (const uint32_t*)(fsum + nsf), (uint32_t*)h->t_starts.p);
    PJ_HIP(hipGetLastError());
    if (h->ev_ok && (h->desc.flags & PJ_FLAG_TIMING))
        PJ_HIP(hipEventRecord(h->ev[9], st));
    if ((rc = launch_wc<8>(rm, false, g.split, in, n, L, pf, g.chunk, g.nblk, coarse, (uint4*)h->spart.p, nullptr,
                           st)))
        return rc;
    hipLaunchKernelGGL(k_subsplit, dim3(g.nblk), dim3(kSubThreads), 0, st, (const uint4*)h->spart.p,
                       (const uint32_t*)coarse, (const uint32_t*)(csum + g.nsb), F, g.nblk, bS, dbits,
                       (const uint32_t*)fine, (uint4*)h->t_recs.p);
    PJ_HIP(hipGetLastError());
    return PJ_OK;
}

// The build side (&h->R) marks event 8 after its histogram, the probe side 9.
int partition_side(pj_join_s* h, const void* d_in, uint64_t n, const SideLayout& L, const PartFn& pf, bool range,
                   bool write_combine, DevBuf& dst, Partitioned& P, DevBuf& starts, bool host_copy,
                   unsigned long long* pminmax, hipStream_t st)
{
    const hipEvent_t mark =
        h->ev_ok && (h->desc.flags & PJ_FLAG_TIMING) ? h->ev[&P == &h->R ? 8 : 9] : nullptr;
    return partition_bufs(h->hist, h->bsum, h->pids, d_in, n, L, pf, range, write_combine, dst, P, starts, host_copy,
                          pminmax, st, mark);
}

int device_count_cus()
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

constexpr uint32_t kProbeSlicesPerCU = 8;

uint64_t probe_chunk(uint64_t ns)
{
    const uint64_t slices = (uint64_t)device_count_cus() * kProbeSlicesPerCU;
    return std::max<uint64_t>((ns + slices - 1) / slices, 32768);
}

// Table format of range build partition f: dense when it holds exactly
// the keys base..base+cnt-1 (checked for repeats in the kernel), code
// tables (split into chunks when larger than one) when its keys span less
// than 2^32; otherwise the plan cannot be used.
bool range_format(const pj_join_s* h, uint32_t f, uint32_t cnt, uint32_t* fmt)
{
    if (cnt == 0) {
        *fmt = kFmtCodes;
        return true;
    }
    const uint64_t lo = h->pmin[f], hi = h->pmax[f];
    if (hi - lo + 1 == (uint64_t)cnt && cnt <= kDenseCap) {
        *fmt = kFmtDense;
        return true;
    }
    *fmt = kFmtCodes;
    return hi - lo < (1ull << 32);
}

hipStream_t as_stream(void* s) { return (hipStream_t)s; }

void record(pj_join_s* h, int i, hipStream_t st)
{
    if (h->ev_ok && (h->desc.flags & PJ_FLAG_TIMING))
        (void)hipEventRecord(h->ev[i], st);
}

float elapsed(pj_join_s* h, int a, int b)
{
    float ms = 0.f;
    if (h->ev_ok && (h->desc.flags & PJ_FLAG_TIMING))
        (void)hipEventElapsedTime(&ms, h->ev[a], h->ev[b]);
    return ms;
}

int prepare_radix(pj_join_s* h, const void* d_build, uint64_t nb, hipStream_t st)
{
    h->range = false;
    h->bR = choose_build_bits(nb, h->desc.radix_bits);
    h->bS = std::min<uint32_t>(h->bR, 11);
    const PartFn pf = radix_fn(h->bR);
    return partition_side(h, d_build, nb, h->lay.build, pf, false, false, h->rpart, h->R, h->rstarts, false, nullptr,
                          st);
}

// Range partitions of the build side from the learned model. Leaves
// h->range false (and the radix plan to the caller) when the model declines
// or a partition cannot be held by either range table format.
int prepare_range(pj_join_s* h, const void* d_build, uint64_t nb, hipStream_t st)
{
    h->range = false;
    int rc;
    hipEvent_t m0 = h->ev[6], m1 = h->ev[7];
    if (h->ev_ok)
        (void)hipEventRecord(m0, st);
    if ((rc = learned_plan(h->lctx, d_build, nb, h->lay.build, kMaxCoarse, st, &h->plan)))
        return rc;
    if (h->ev_ok) {
        (void)hipEventRecord(m1, st);
        (void)hipEventSynchronize(m1);
        (void)hipEventElapsedTime(&h->model_ms, m0, m1);
    }
    if (!h->plan.ok)
        return PJ_OK;
    const RangePlan& pl = h->plan;
    if ((rc = h->map_dev.ensure(kMapBytes)) || (rc = h->pminmax.ensure((size_t)pl.nfine * 16)))
        return rc;
    PJ_HIP(hipMemcpyAsync(h->map_dev.p, pl.map.data(), kMapBytes, hipMemcpyHostToDevice, st));
    const PartFn pf = map_fn(pl.nfine, (const uint8_t*)h->map_dev.p, 1);
    if ((rc = partition_side(h, d_build, nb, h->lay.build, pf, true, false, h->rpart, h->R, h->rstarts, true,
                             (unsigned long long*)h->pminmax.p, st)))
        return rc;
    h->pmin.resize(pl.nfine);
    h->pmax.resize(pl.nfine);
    PJ_HIP(hipMemcpyAsync(h->pmin.data(), h->pminmax.p, (size_t)pl.nfine * 8, hipMemcpyDeviceToHost, st));
    PJ_HIP(hipMemcpyAsync(h->pmax.data(), (char*)h->pminmax.p + (size_t)pl.nfine * 8, (size_t)pl.nfine * 8,
                          hipMemcpyDeviceToHost, st));
    PJ_HIP(hipStreamSynchronize(st));
    // each build partition's table, for the work items derived on the device
    std::vector<PlanPart> parts(pl.nfine);
    for (uint32_t f = 0; f < pl.nfine; ++f) {
        uint32_t fmt;
        if (!range_format(h, f, h->R.start[f + 1] - h->R.start[f], &fmt)) {
            h->plan.ok = false;
            h->plan.why = "a build partition spans 2^32 keys or more";
            return PJ_OK;
        }
        PlanPart& q = parts[f];
        q.fmt = fmt;
        q.cap = fmt == kFmtDense ? kDenseCap : kCodeCap;
        // the partition's build keys span [pmin, pmax]: a probe key outside
        // it cannot be one of them (a key's partition is a function of the key)
        q.lo = h->pmin[f];
        q.hi = h->pmax[f];
        q.base = h->pmin[f];
    }
    if ((rc = h->plan_parts.ensure(parts.size() * sizeof(PlanPart))) ||
        (rc = h->coarse_of.ensure((size_t)pl.nfine * 4)))
        return rc;
    PJ_HIP(hipMemcpyAsync(h->plan_parts.p, parts.data(), parts.size() * sizeof(PlanPart), hipMemcpyHostToDevice, st));
    PJ_HIP(hipMemcpyAsync(h->coarse_of.p, pl.coarse_of.data(), (size_t)pl.nfine * 4, hipMemcpyHostToDevice, st));
    PJ_HIP(hipStreamSynchronize(st));   // the host arrays are locals
    h->range = true;
    return PJ_OK;
}

// Whether the learned partitioner should be tried: always when asked for,
// and under "auto" when radix would have to re-read probe partitions (more
// build than probe partitions).
bool want_learned(const pj_join_s* h, uint64_t nb)
{
    if (h->lay.rowid)   // hashed keys have no order to cut into ranges
        return false;
    if (h->desc.partitioner == PJ_PART_LEARNED)
        return true;
    if (h->desc.partitioner != PJ_PART_AUTO || h->desc.radix_bits >= 0)
        return false;
    return choose_build_bits(nb, -1) > 11;
}

int prepare_build(pj_join_s* h, const void* d_build, uint64_t nb, hipStream_t st)
{
    h->model_ms = 0.f;
    h->raw_build = d_build;
    h->raw_nbuild = nb;
    h->plan = RangePlan();
    int rc;
    record(h, 0, st);
    if (want_learned(h, nb) && nb > 0) {
        if ((rc = prepare_range(h, d_build, nb, st)))
            return rc;
    }
    if (!h->range)
        rc = prepare_radix(h, d_build, nb, st);
    record(h, 1, st);
    return rc;
}

// Enqueues one probe phase on st: partition S, count the work items, join,
// and copy the join's counters into slot (8 words of pinned host memory:
// [2..3] rows produced, [4] range tables that could not be held, [6] work
// items, [7] split tables; the row-id form copies 12 words, the join's
// pairs then in [2..3] and the materialising pass's key mismatches in
// [10]) when the stream reaches that copy. *empty: the join has no work
// (the slot is zeroed at once).
int probe_enqueue(pj_join_s* h, const void* d_probe, uint64_t np, void* d_out, uint64_t out_cap, uint32_t* slot,
                  bool* empty, hipStream_t st)
{
    int rc;
    *empty = false;
    record(h, 2, st);
    h->s_app = want_append(h, d_probe, np);
    h->s_two = !h->s_app && want_two_pass(h, np);
    const PartFn spf =
        h->range ? map_fn(h->plan.ncoarse, (const uint8_t*)h->map_dev.p, h->plan.div) : radix_fn(h->bS);
    if (h->s_app)
        rc = partition_append(h, d_probe, np, h->lay.probe, spf, st);
    else if (h->s_two)
        rc = partition_two_pass(h, d_probe, np, h->lay.probe, st);
    else
        rc = partition_side(h, d_probe, np, h->lay.probe, spf, h->range, true, h->spart, h->S, h->sstarts, false,
                            nullptr, st);
    if (rc)
        return rc;
    record(h, 3, st);
    if (h->R.n == 0 || np == 0) {
        *empty = true;
        for (int i = 0; i < (h->lay.rowid ? 12 : 8); ++i)
            slot[i] = 0;
        return PJ_OK;
    }
    // the work items, counted on the device from the partition starts
    PlanView v = {};
    v.FR = h->range ? h->plan.nfine : 1u << h->bR;
    v.shift = h->range || h->s_two ? 0u : h->bR - h->bS;   // two passes: probe partitions = build partitions
    v.rst = (const uint32_t*)h->rstarts.p;
    v.sst = (const uint32_t*)(h->s_two ? h->t_starts.p : h->sstarts.p);
    v.coarse_of = h->range ? (const uint32_t*)h->coarse_of.p : nullptr;
    v.pp = h->range ? (const PlanPart*)h->plan_parts.p : nullptr;
    v.cl = h->s_app ? (const uint32_t*)h->a_clist.p : nullptr;
    v.chunk = probe_chunk(np);
    if (v.FR > kPlanThreads * kPlanPer)
        return fail(PJ_ERR_UNSUPPORTED, "%u build partitions exceed the planner", v.FR);
    if ((rc = h->planoffs.ensure((size_t)(v.FR + 1) * 4)) || (rc = h->counters.ensure(64)))
        return rc;
    v.offs = (const uint32_t*)h->planoffs.p;
    uint32_t* counters = (uint32_t*)h->counters.p;
    hipLaunchKernelGGL(k_plan, dim3(1), dim3(kPlanThreads), 0, st, v, (uint32_t*)h->planoffs.p, counters);
    uint32_t* item_ctr = counters;
    unsigned long long* cursor = (unsigned long long*)(counters + 2);
    uint32_t* fail_ctr = counters + 4;
    const size_t lds = join_lds_bytes(h->range);
    const void* kfn = h->range ? (const void*)k_join<true, false>
                               : (h->s_app ? (const void*)k_join<false, true> : (const void*)k_join<false, false>);
    if ((rc = allow_lds(kfn, lds)))
        return rc;
    const uint32_t bkt_shift = 64 - h->bR - 12;
    // the row-id form's join writes (build row, probe row) pairs, checked and
    // projected afterwards
    void* jout = d_out;
    uint64_t jcap = out_cap;
    if (h->lay.rowid) {
        uint64_t want = std::max<uint64_t>(out_cap + out_cap / 64 + 1024, h->pair_want);
        if (h->pairs.arena && h->pairs.cap)   // fixed scratch: the pairs it was given
            want = std::min<uint64_t>(want, h->pairs.cap / 8);
        if ((rc = h->pairs.ensure(want * 8)))
            return rc;
        jout = h->pairs.p;
        jcap = h->pair_cap = h->pairs.cap / 8;
    }
    record(h, 4, st);
    const uint32_t grid = (uint32_t)device_count_cus();   // blocks without an item exit at once
    const uint32_t jbits = h->range ? 0u : h->bR, jshift = h->range ? 0u : bkt_shift;
    auto launch = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(kJoinThreads), lds, st, (const uint4*)h->rpart.p,
                           (const uint4*)(h->s_two ? h->t_recs.p : h->spart.p), v, item_ctr, (uint32_t*)jout, jcap,
                           cursor, h->lay.out, jbits, jshift, fail_ctr);
    };
    if (h->range)   // range plans keep the contiguous layout
        launch(k_join<true, false>);
    else
        h->s_app ? launch(k_join<false, true>) : launch(k_join<false, false>);
    PJ_HIP(hipGetLastError());
    if (h->lay.rowid &&
        (rc = rows_materialize((const uint2*)h->pairs.p, cursor, jcap, h->raw_build, d_probe, h->lay.rows, d_out,
                               out_cap, (unsigned long long*)(counters + 8), counters + 10, false, st)))
        return rc;
    record(h, 5, st);
    PJ_HIP(hipMemcpyAsync(slot, counters, h->lay.rowid ? 48 : 32, hipMemcpyDeviceToHost, st));
    return PJ_OK;
}

int probe_once(pj_join_s* h, const void* d_probe, uint64_t np, void* d_out, uint64_t out_cap, uint64_t* out_n,
               uint32_t* fails, hipStream_t st)
{
    *out_n = 0;
    *fails = 0;
    if (!h->host_ctr)
        PJ_HIP(hipHostMalloc((void**)&h->host_ctr, 64, hipHostMallocDefault));
    uint32_t* produced = h->host_ctr;
    bool empty = false;
    int rc = probe_enqueue(h, d_probe, np, d_out, out_cap, produced, &empty, st);
    if (rc)
        return rc;
    PJ_HIP(hipStreamSynchronize(st));
    if (empty) {
        memset(&h->last, 0, sizeof(h->last));
        return PJ_OK;
    }
    *out_n = (uint64_t)produced[2] | ((uint64_t)produced[3] << 32);
    *fails = produced[4];
    if (h->lay.rowid) {
        if (*out_n > h->pair_cap) {   // pairs were dropped: the caller runs the probe again with room for them
            h->pair_want = *out_n;
            return PJ_ERR_CAPACITY;
        }
        if (produced[10]) {
            // some pairs had equal hashes and unequal keys: write the rows
            // densely instead (the rows counter, words 8-9, is still zero)
            uint32_t* counters = (uint32_t*)h->counters.p;
            int rc2 = rows_materialize((const uint2*)h->pairs.p, (const unsigned long long*)(counters + 2),
                                       h->pair_cap, h->raw_build, d_probe, h->lay.rows, d_out, out_cap,
                                       (unsigned long long*)(counters + 8), counters + 10, true, st);
            if (rc2)
                return rc2;
            PJ_HIP(hipMemcpyAsync(produced + 8, counters + 8, 8, hipMemcpyDeviceToHost, st));
            PJ_HIP(hipStreamSynchronize(st));
            *out_n = (uint64_t)produced[8] | ((uint64_t)produced[9] << 32);
        }
    }
    pj_timings& t = h->last;
    t.build_partition_ms = elapsed(h, 0, 1) - h->model_ms;
    t.probe_partition_ms = elapsed(h, 2, 3);
    t.join_ms = elapsed(h, 4, 5);
    t.build_hist_ms = elapsed(h, 0, 8) - h->model_ms;
    t.probe_hist_ms = elapsed(h, 2, 9);
    t.model_ms = h->model_ms;
    t.total_ms = elapsed(h, 0, 5);
    t.partitions = h->range ? h->plan.ncoarse : 1u << h->bS;
    t.subpartitions = h->range ? (h->plan.nfine + h->plan.ncoarse - 1) / h->plan.ncoarse : 1u << (h->bR - h->bS);
    t.work_items = produced[6];
    t.used_learned = h->range ? 1u : 0u;
    t.spilled_items = produced[7];
    t.probe_appended = h->s_app ? 1u : 0u;
    t.probe_passes = h->s_two ? 2u : 1u;
    return PJ_OK;
}

// The probe phase. When a range plan's kernel met a partition it could not
// hold (repeated keys in a partition too large for a code table), the build
// side is partitioned again with radix bits and the probe phase repeats.
int probe_device(pj_join_s* h, const void* d_probe, uint64_t np, void* d_out, uint64_t out_cap, uint64_t* out_n,
                 hipStream_t st)
{
    uint32_t fails = 0;
    int rc = probe_once(h, d_probe, np, d_out, out_cap, out_n, &fails, st);
    if (rc == PJ_ERR_CAPACITY && h->lay.rowid) {
        if (h->pairs.arena)   // a fixed scratch cannot grow
            return fail(PJ_ERR_UNSUPPORTED, "the join's %llu row pairs exceed the %llu the fixed scratch holds",
                        (unsigned long long)h->pair_want, (unsigned long long)h->pair_cap);
        // the pair buffer was too small: once more with room for all
        rc = probe_once(h, d_probe, np, d_out, out_cap, out_n, &fails, st);
    }
    if (rc == PJ_OK && fails) {
        if (!h->raw_build)
            return fail(PJ_ERR_STATE, "range plan failed and the build tuples are gone");
        h->plan.ok = false;
        h->plan.why = "repeated keys in a dense-looking partition";
        if ((rc = prepare_radix(h, h->raw_build, h->raw_nbuild, st)))
            return rc;
        rc = probe_once(h, d_probe, np, d_out, out_cap, out_n, &fails, st);
    }
    if (rc)
        return rc;
    if (*out_n > out_cap)
        return fail(PJ_ERR_CAPACITY, "join produced %llu tuples, output holds %llu", (unsigned long long)*out_n,
                    (unsigned long long)out_cap);
    return PJ_OK;
}

}  // namespace

int pj::join_enqueue(pj_join_s* h, const void* d_build, uint64_t nb, const void* d_probe, uint64_t np, void* d_out,
                     uint64_t out_cap, uint32_t* slot, hipStream_t st)
{
    if (!h || !slot || (nb && !d_build) || (np && !d_probe) || (out_cap && !d_out))
        return fail(PJ_ERR_INVALID, "null argument");
    if (out_cap && ((uintptr_t)d_out % 16))
        return fail(PJ_ERR_INVALID, "the output buffer must be 16-byte aligned");
    if (h->desc.partitioner != PJ_PART_RADIX)
        return fail(PJ_ERR_INVALID, "enqueued joins use radix partitions");
    std::lock_guard<std::mutex> g(h->mu);
    int rc = prepare_build(h, d_build, nb, st);
    if (rc)
        return rc;
    bool empty;
    return probe_enqueue(h, d_probe, np, d_out, out_cap, slot, &empty, st);
}

// ---------------------------------------------------------------- batched joins

uint32_t pj::mbatch_scatter_tile(uint32_t maxF) { return scatter_tile(maxF, false); }

namespace {

size_t mbatch_scatter_lds(uint32_t maxF)
{
    return (size_t)scatter_tile(maxF, false) * 16 + (2 * (size_t)maxF + 1) * 4 + 16 * 4;
}

}  // namespace

int pj::mbatch_setup(const MBatch& b)
{
    int rc;
    const size_t hl = (size_t)b.maxF * 4, sl = mbatch_scatter_lds(b.maxF), jl = join_lds_bytes(false);
    if ((rc = allow_lds((const void*)k_mhist<REC_WORDS>, hl)) || (rc = allow_lds((const void*)k_mhist<REC_IDENT>, hl)) ||
        (rc = allow_lds((const void*)k_mscatter<REC_WORDS>, sl)) ||
        (rc = allow_lds((const void*)k_mscatter<REC_IDENT>, sl)) || (rc = allow_lds((const void*)k_mjoin, jl)))
        return rc;
    return PJ_OK;
}

int pj::mbatch_enqueue(MBatch& b, hipStream_t st)
{
    if (b.njoins == 0 || b.maxF > (1u << kMaxBatchBits) || b.max_items >= (1ull << 32))
        return fail(PJ_ERR_STATE, "no prepared batch");
    const MSide* sides = (const MSide*)b.sides.p;
    uint32_t* hist = (uint32_t*)b.hist.p;
    const SideLayout& L0 = b.lay.build;
    const SideLayout& L1 = b.lay.probe;
    if (b.nblocks) {
        const size_t hl = (size_t)b.maxF * 4, sl = mbatch_scatter_lds(b.maxF);
        const uint32_t ts = scatter_tile(b.maxF, false);
        if (ts == 0)
            return fail(PJ_ERR_UNSUPPORTED, "%u partitions exceed the batch scatter's LDS budget", b.maxF);
        if (b.rm == REC_IDENT)
            hipLaunchKernelGGL(k_mhist<REC_IDENT>, dim3(b.nblocks), dim3(kHistThreads), hl, st, sides, b.nsides, L0,
                               L1, hist);
        else
            hipLaunchKernelGGL(k_mhist<REC_WORDS>, dim3(b.nblocks), dim3(kHistThreads), hl, st, sides, b.nsides, L0,
                               L1, hist);
        const uint32_t nsb = (uint32_t)((b.nhist + kScanSeg - 1) / kScanSeg);
        uint32_t* bsum = (uint32_t*)b.bsum.p;
        hipLaunchKernelGGL(k_scan_reduce, dim3(nsb), dim3(kScanThreads), 0, st, hist, b.nhist, bsum);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, bsum, nsb);
        hipLaunchKernelGGL(k_scan_down, dim3(nsb), dim3(kScanThreads), 0, st, hist, b.nhist, (const uint32_t*)bsum);
        if (b.rm == REC_IDENT)
            hipLaunchKernelGGL(k_mscatter<REC_IDENT>, dim3(b.nblocks), dim3(kScatterThreads), sl, st, sides, b.nsides,
                               L0, L1, (const uint32_t*)hist, ts, b.maxF, (uint4*)b.recs.p);
        else
            hipLaunchKernelGGL(k_mscatter<REC_WORDS>, dim3(b.nblocks), dim3(kScatterThreads), sl, st, sides, b.nsides,
                               L0, L1, (const uint32_t*)hist, ts, b.maxF, (uint4*)b.recs.p);
    }
    hipLaunchKernelGGL(k_mbounds, dim3(b.nsides), dim3(256), 0, st, sides, (const uint32_t*)hist,
                       (uint32_t*)b.starts.p);
    MState* state = (MState*)b.state.p;
    uint32_t* misc = (uint32_t*)b.misc.p;
    hipLaunchKernelGGL(k_mplan, dim3(b.njoins), dim3(kPlanThreads), 0, st, (const MJoin*)b.joins.p, state);
    hipLaunchKernelGGL(k_mitems, dim3(1), dim3(1024), 0, st, (const MState*)state, b.njoins, (uint32_t*)b.ibase.p,
                       misc, (unsigned long long*)(misc + 2));
    hipLaunchKernelGGL(k_mjoin, dim3(device_count_cus()), dim3(kJoinThreads), join_lds_bytes(false), st,
                       (const uint4*)b.recs.p, (const MJoin*)b.joins.p, b.njoins, (const uint32_t*)b.ibase.p, state,
                       misc, (uint32_t)b.max_items, b.lay.out);
    PJ_HIP(hipGetLastError());
    return PJ_OK;
}

// ---------------------------------------------------------------- scratch
// The no-allocation entry point's device memory: one fixed array, the same
// symbol on every device (each device holds its own copy). Host code never
// passes the symbol itself to a kernel; it takes the array's address on each
// device once, with hipGetSymbolAddress, and uses that.
static __device__ __align__(256) uint8_t scratch[PJ_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;

struct ScratchSlot {
    std::mutex mu;        // calls on one device share its scratch: one at a time
    void* ptr = nullptr;  // the device's address of `scratch`
};
ScratchSlot g_scratch[kMaxDevices];
std::mutex g_scratch_mu;

// The current device and its address of `scratch` (looked up once a device).
int scratch_slot(ScratchSlot** out)
{
    int dev = 0;
    PJ_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return fail(PJ_ERR_DEVICE, "device %d beyond the %d the scratch table covers", dev, kMaxDevices);
    std::lock_guard<std::mutex> g(g_scratch_mu);
    ScratchSlot& s = g_scratch[dev];
    if (!s.ptr) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(PJ_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev,
                        hipGetErrorString(e));
        s.ptr = p;
    }
    *out = &s;
    return PJ_OK;
}

size_t align256(size_t x) { return (x + 255) / 256 * 256; }
size_t piece(size_t bytes) { return align256(bytes < 256 ? 256 : bytes); }

// Scratch bytes one radix join of nb x np tuples carves (every buffer the
// partition passes and the probe phase ensure, each a piece of its own),
// the row-id form's pairs aside.
size_t noalloc_need(const pj_join_s* h, uint64_t nb, uint64_t np)
{
    const uint32_t bR = choose_build_bits(nb, h->desc.radix_bits);
    const uint32_t bS = std::min<uint32_t>(bR, 11);
    size_t need = piece(((size_t)(1u << bR) + 1) * 4) + piece(((size_t)(1u << bS) + 1) * 4);
    if (nb) {
        const PartGeom g = part_geom(nb, radix_fn(bR), false, false);
        need += piece(g.cap * 16) + piece(g.m * 4) + piece(((size_t)g.nsb + 1) * 4);
    }
    if (np) {
        const PartGeom g = part_geom(np, radix_fn(bS), false, true);
        need += piece(g.cap * 16) + piece(g.m * 4) + piece(((size_t)g.nsb + 1) * 4);
    }
    need += piece(((size_t)(1u << bR) + 1) * 4) + piece(64);   // plan offsets, counters
    return need;
}

// For the duration of one call, the handle's device buffers are empty
// buffers carved from the scratch; the handle's own are put back after.
struct ScratchSwap {
    struct Saved {
        void* p;
        size_t cap;
        Arena* arena;
    };
    Arena ar;
    DevBuf* bufs[10];
    Saved saved[10];
    ScratchSwap(pj_join_s* h, void* base, size_t size)
        : ar{(uint8_t*)base, size, 0},
          bufs{&h->rpart, &h->spart, &h->hist, &h->bsum, &h->pids, &h->rstarts, &h->sstarts, &h->planoffs,
               &h->counters, &h->pairs}
    {
        for (int i = 0; i < 10; ++i) {
            saved[i] = Saved{bufs[i]->p, bufs[i]->cap, bufs[i]->arena};
            bufs[i]->p = nullptr;
            bufs[i]->cap = 0;
            bufs[i]->arena = &ar;
        }
    }
    ~ScratchSwap()
    {
        for (int i = 0; i < 10; ++i) {
            bufs[i]->p = saved[i].p;
            bufs[i]->cap = saved[i].cap;
            bufs[i]->arena = saved[i].arena;
        }
    }
};

}  // namespace

// ======================================================================
// C ABI
// ======================================================================

namespace pj {

int radix_records(RadixRecords& r, const void* d_in, uint64_t n, const SideLayout& L, uint32_t bits, hipStream_t st)
{
    Partitioned P;
    return partition_bufs(r.hist, r.bsum, r.pids, d_in, n, L, radix_fn(bits), false, true, r.recs, P, r.starts, false,
                          nullptr, st);
}

}  // namespace pj

extern "C" {

const char* pj_last_error(void) { return pj::last_error(); }
void pj_set_error(const char* msg) { pj::set_error("%s", msg ? msg : ""); }
const char* pj_version(void) { return "pjoin 0.1 gfx950"; }

int pj_set_device_limit(uint64_t bytes)
{
    pj::g_dev_limit.store(bytes);
    return PJ_OK;
}

int pj_device_bytes(uint64_t* held)
{
    if (!held)
        return fail(PJ_ERR_INVALID, "null argument");
    *held = pj::g_dev_used.load();
    return PJ_OK;
}

int pj_join_create(const pj_join_desc* desc, pj_join_t* out)
{
    if (!desc || !out)
        return fail(PJ_ERR_INVALID, "null argument");
    *out = nullptr;
    Layout L;
    int rc = derive_layout(desc, &L);
    if (rc)
        return rc;
    if (desc->partitioner > PJ_PART_AUTO)
        return fail(PJ_ERR_INVALID, "unknown partitioner %u", desc->partitioner);
    pj_join_s* h = new (std::nothrow) pj_join_s();
    if (!h)
        return fail(PJ_ERR_INVALID, "out of host memory");
    h->desc = *desc;
    h->lay = L;
    memset(&h->last, 0, sizeof(h->last));
    h->ev_ok = true;
    for (int i = 0; i < 10; ++i)
        if (hipEventCreate(&h->ev[i]) != hipSuccess)
            h->ev_ok = false;
    // pinned words the probe phase's counters are read back into (without a
    // device the handle still validates descriptors; the probe phase then
    // allocates them, and the no-allocation entry point refuses the handle)
    if (hipHostMalloc((void**)&h->host_ctr, 64, hipHostMallocDefault) != hipSuccess)
        h->host_ctr = nullptr;
    *out = h;
    return PJ_OK;
}

int pj_join_out_width(pj_join_t h, uint32_t* width)
{
    if (!h || !width)
        return fail(PJ_ERR_INVALID, "null argument");
    *width = h->lay.out_width;
    return PJ_OK;
}

int pj_join_destroy(pj_join_t h)
{
    if (!h)
        return PJ_OK;
    for (DevBuf* b : {&h->d_raw_build, &h->d_raw_probe, &h->d_out_batch, &h->rpart, &h->spart, &h->hist, &h->bsum,
                      &h->rstarts, &h->sstarts, &h->planoffs, &h->counters, &h->map_dev,
                      &h->pminmax, &h->pids, &h->plan_parts, &h->coarse_of, &h->pairs})
        b->release();
    if (h->ev_ok)
        for (int i = 0; i < 10; ++i)
            (void)hipEventDestroy(h->ev[i]);
    if (h->host_ctr)
        (void)hipHostFree(h->host_ctr);
    delete h;
    return PJ_OK;
}

int pj_join_run_device(pj_join_t h, const void* d_build, uint64_t nbuild, const void* d_probe, uint64_t nprobe,
                       void* d_out, uint64_t out_cap, uint64_t* out_n, void* stream)
{
    if (!h || !out_n || (nbuild && !d_build) || (nprobe && !d_probe) || (out_cap && !d_out))
        return fail(PJ_ERR_INVALID, "null argument");
    if (out_cap && ((uintptr_t)d_out % 16))
        return fail(PJ_ERR_INVALID, "the output buffer must be 16-byte aligned");
    std::lock_guard<std::mutex> g(h->mu);
    hipStream_t st = as_stream(stream);
    int rc = prepare_build(h, d_build, nbuild, st);
    if (rc)
        return rc;
    return probe_device(h, d_probe, nprobe, d_out, out_cap, out_n, st);
}

int pj_join_count_by_key(pj_join_t h, const void* d_build, uint64_t nbuild, const void* d_probe, uint64_t nprobe,
                         void* d_groups, uint64_t groups_cap, uint64_t* ngroups, void* stream)
{
    if (!h || !ngroups || (nbuild && !d_build) || (nprobe && !d_probe) || (groups_cap && !d_groups))
        return fail(PJ_ERR_INVALID, "null argument");
    if ((uintptr_t)d_groups % 4)
        return fail(PJ_ERR_INVALID, "the group buffer must be 4-byte aligned");
    *ngroups = 0;
    const pj_join_desc& d = h->desc;
    if (h->lay.rowid)
        return fail(PJ_ERR_UNSUPPORTED, "the count per key takes one INTEGER / LONG / DATE key or two INTEGER keys");
    // the group columns: the build key, one INTEGER / LONG / DATE column or two INTEGER columns
    pj_schema gs = {};
    gs.ncols = d.nkeys + 1;
    for (uint32_t k = 0; k < d.nkeys; ++k)
        gs.type[k] = d.build.type[d.buildjattr[k]];
    gs.type[d.nkeys] = PJ_CT_LONG;
    const uint32_t key_words = d.nkeys == 2 || gs.type[0] != PJ_CT_INTEGER ? 2u : 1u;
    std::lock_guard<std::mutex> g(h->mu);
    hipStream_t st = as_stream(stream);
    int rc;
    h->model_ms = 0.f;
    h->raw_build = d_build;
    h->raw_nbuild = nbuild;
    h->plan = RangePlan();
    if ((rc = prepare_radix(h, d_build, nbuild, st)))
        return rc;
    h->s_app = want_append(h, d_probe, nprobe);
    h->s_two = false;
    rc = h->s_app ? partition_append(h, d_probe, nprobe, h->lay.probe, radix_fn(h->bS), st)
                  : partition_side(h, d_probe, nprobe, h->lay.probe, radix_fn(h->bS), false, true, h->spart, h->S,
                                   h->sstarts, false, nullptr, st);
    if (rc)
        return rc;
    if (nbuild == 0 || nprobe == 0) {
        PJ_HIP(hipStreamSynchronize(st));
        return PJ_OK;
    }
    PlanView v = {};
    v.FR = 1u << h->bR;
    v.shift = h->bR - h->bS;
    v.rst = (const uint32_t*)h->rstarts.p;
    v.sst = (const uint32_t*)h->sstarts.p;
    v.cl = h->s_app ? (const uint32_t*)h->a_clist.p : nullptr;
    v.chunk = probe_chunk(nprobe);
    if (v.FR > kPlanThreads * kPlanPer)
        return fail(PJ_ERR_UNSUPPORTED, "%u build partitions exceed the planner", v.FR);
    const uint64_t nrec = h->rpart.cap / 16;   // records R_part can hold, padding included
    if ((rc = h->planoffs.ensure((size_t)(v.FR + 1) * 4)) || (rc = h->counters.ensure(64)) ||
        (rc = h->gj_count.ensure(nrec * 4)))
        return rc;
    if (!h->host_ctr)
        PJ_HIP(hipHostMalloc((void**)&h->host_ctr, 64, hipHostMallocDefault));
    v.offs = (const uint32_t*)h->planoffs.p;
    uint32_t* counters = (uint32_t*)h->counters.p;
    uint32_t* rcount = (uint32_t*)h->gj_count.p;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)((nrec + 255) / 256)), dim3(256), 0, st, rcount, (uint32_t)nrec,
                       0u);
    hipLaunchKernelGGL(k_plan, dim3(1), dim3(kPlanThreads), 0, st, v, (uint32_t*)h->planoffs.p, counters);
    const size_t lds = join_lds_bytes(false);
    const void* gfn = h->s_app ? (const void*)k_groupjoin<true> : (const void*)k_groupjoin<false>;
    if ((rc = allow_lds(gfn, lds)))
        return rc;
    auto launch = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3((uint32_t)device_count_cus()), dim3(kJoinThreads), lds, st,
                           (const uint4*)h->rpart.p, (const uint4*)h->spart.p, v, counters, rcount, h->bR,
                           64 - h->bR - 12, counters + 1);
    };
    h->s_app ? launch(k_groupjoin<true>) : launch(k_groupjoin<false>);
    PJ_HIP(hipGetLastError());
    PJ_HIP(hipMemcpyAsync(h->host_ctr, counters, 16, hipMemcpyDeviceToHost, st));
    PJ_HIP(hipStreamSynchronize(st));
    const bool dup = h->host_ctr[1] != 0;
    unsigned long long* cursor = (unsigned long long*)(counters + 2);
    const uint32_t emit_grid = (uint32_t)std::min<uint64_t>((nrec + 4095) / 4096, 4096);
    if (!dup) {   // every key on one record: the records' counts are the groups
        hipLaunchKernelGGL(k_groupjoin_emit, dim3(emit_grid), dim3(256), 0, st, (const uint4*)h->rpart.p,
                           (const uint32_t*)h->rstarts.p, v.FR, (const uint32_t*)rcount, key_words,
                           (uint32_t*)d_groups, groups_cap, cursor);
        PJ_HIP(hipGetLastError());
        PJ_HIP(hipMemcpyAsync(h->host_ctr, cursor, 8, hipMemcpyDeviceToHost, st));
        PJ_HIP(hipStreamSynchronize(st));
        *ngroups = *(const uint64_t*)h->host_ctr;
        if (*ngroups > groups_cap)
            return fail(PJ_ERR_CAPACITY, "%llu groups, the output holds %llu", (unsigned long long)*ngroups,
                        (unsigned long long)groups_cap);
        return PJ_OK;
    }
    // equal keys on several records: one row a record with matches, then the
    // sum of their counts per key
    const uint32_t row_bytes = (key_words + 2) * 4;
    if ((rc = h->gj_rows.ensure(nbuild * row_bytes)))
        return rc;
    hipLaunchKernelGGL(k_groupjoin_emit, dim3(emit_grid), dim3(256), 0, st, (const uint4*)h->rpart.p,
                       (const uint32_t*)h->rstarts.p, v.FR, (const uint32_t*)rcount, key_words,
                       (uint32_t*)h->gj_rows.p, nbuild, cursor);
    PJ_HIP(hipGetLastError());
    PJ_HIP(hipMemcpyAsync(h->host_ctr, cursor, 8, hipMemcpyDeviceToHost, st));
    PJ_HIP(hipStreamSynchronize(st));
    const uint64_t nrows = *(const uint64_t*)h->host_ctr;
    if (nrows > nbuild)
        return fail(PJ_ERR_STATE, "group-join emitted %llu rows for %llu build records", (unsigned long long)nrows,
                    (unsigned long long)nbuild);
    pj_agg_desc ad = {};
    ad.in = gs;
    ad.op = PJ_AGG_SUM;
    ad.nfields = d.nkeys;
    for (uint32_t k = 0; k < d.nkeys; ++k)
        ad.fields[k] = (uint16_t)k;
    ad.sumfield = (uint16_t)d.nkeys;
    ad.max_groups = std::max<uint64_t>(nrows, 1);
    pj_agg_t a = nullptr;
    if ((rc = pj_agg_create(&ad, &a)))
        return rc;
    rc = pj_agg_run_device(a, h->gj_rows.p, nrows, d_groups, groups_cap, ngroups, stream);
    pj_agg_destroy(a);
    return rc;
}

int pj_join_count_by_key_host(pj_join_t h, const void* build, uint64_t nbuild, const void* probe, uint64_t nprobe,
                              void* groups, uint64_t groups_cap, uint64_t* ngroups)
{
    if (!h || !ngroups || (nbuild && !build) || (nprobe && !probe) || (groups_cap && !groups))
        return fail(PJ_ERR_INVALID, "null argument");
    *ngroups = 0;
    int boff[PJ_MAX_COLS], poff[PJ_MAX_COLS], bsize = 0, psize = 0;
    int rc;
    if ((rc = schema_offsets(h->desc.build, boff, &bsize)) || (rc = schema_offsets(h->desc.probe, poff, &psize)))
        return rc;
    const uint32_t key_words = h->desc.nkeys == 2 || h->desc.build.type[h->desc.buildjattr[0]] != PJ_CT_INTEGER ? 2u
                                                                                                                : 1u;
    const uint64_t gw = (key_words + 2) * 4;
    DevBuf bdev, pdev, gdev;   // this call's own device copies
    if ((rc = bdev.ensure(std::max<uint64_t>(nbuild * (uint64_t)bsize, 16))) ||
        (rc = pdev.ensure(std::max<uint64_t>(nprobe * (uint64_t)psize, 16))) ||
        (rc = gdev.ensure(std::max<uint64_t>(groups_cap * gw, 16))))
        return rc;
    if (nbuild)
        PJ_HIP(hipMemcpy(bdev.p, build, nbuild * (uint64_t)bsize, hipMemcpyHostToDevice));
    if (nprobe)
        PJ_HIP(hipMemcpy(pdev.p, probe, nprobe * (uint64_t)psize, hipMemcpyHostToDevice));
    if ((rc = pj_join_count_by_key(h, bdev.p, nbuild, pdev.p, nprobe, gdev.p, groups_cap, ngroups, nullptr)))
        return rc;
    if (*ngroups)
        PJ_HIP(hipMemcpy(groups, gdev.p, *ngroups * gw, hipMemcpyDeviceToHost));
    return PJ_OK;
}

int pj_join_build_append(pj_join_t h, const void* tuples, uint64_t n)
{
    if (!h || (n && !tuples))
        return fail(PJ_ERR_INVALID, "null argument");
    if (h->built)
        return fail(PJ_ERR_STATE, "build side already finished");
    const size_t w = (size_t)h->lay.build.tbytes;
    std::lock_guard<std::mutex> g(h->stage_mu);
    const uint8_t* p = (const uint8_t*)tuples;
    h->staged.insert(h->staged.end(), p, p + n * w);
    return PJ_OK;
}

int pj_join_build_finish(pj_join_t h)
{
    if (!h)
        return fail(PJ_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(h->mu);
    if (h->built)
        return fail(PJ_ERR_STATE, "build side already finished");
    const size_t w = (size_t)h->lay.build.tbytes;
    const uint64_t n = h->staged.size() / w;
    int rc;
    if ((rc = h->d_raw_build.ensure(std::max<size_t>(h->staged.size(), 16))))
        return rc;
    if (n)
        PJ_HIP(hipMemcpy(h->d_raw_build.p, h->staged.data(), h->staged.size(), hipMemcpyHostToDevice));
    if ((rc = prepare_build(h, h->d_raw_build.p, n, nullptr)))
        return rc;
    PJ_HIP(hipStreamSynchronize(nullptr));
    h->built = true;
    return PJ_OK;
}

int pj_join_probe(pj_join_t h, const void* tuples, uint64_t n, pj_cursor* cur, void* out, uint64_t out_cap,
                  uint64_t* out_n)
{
    if (!h || !cur || !out_n || (out_cap && !out) || (n && !tuples))
        return fail(PJ_ERR_INVALID, "null argument");
    *out_n = 0;
    CursorImpl* ci = (CursorImpl*)cur->impl;
    const uint32_t ow = h->lay.out_width;
    if (tuples) {
        if (!h->built)
            return fail(PJ_ERR_STATE, "probe before pj_join_build_finish");
        if (ci && ci->next < ci->total)
            return fail(PJ_ERR_STATE, "cursor still holds %llu rows", (unsigned long long)(ci->total - ci->next));
        if (!ci) {
            ci = new CursorImpl();
            cur->impl = ci;
        }
        ci->rows.clear();
        ci->next = ci->total = 0;
        if (n) {
            std::lock_guard<std::mutex> g(h->mu);
            const size_t w = (size_t)h->lay.probe.tbytes;
            int rc;
            if ((rc = h->d_raw_probe.ensure(n * w)))
                return rc;
            PJ_HIP(hipMemcpy(h->d_raw_probe.p, tuples, n * w, hipMemcpyHostToDevice));
            uint64_t cap = std::max<uint64_t>(n, 1024), produced = 0;
            for (int attempt = 0; attempt < 2; ++attempt) {
                if ((rc = h->d_out_batch.ensure(cap * ow)))
                    return rc;
                rc = probe_device(h, h->d_raw_probe.p, n, h->d_out_batch.p, cap, &produced, nullptr);
                if (rc != PJ_ERR_CAPACITY)
                    break;
                cap = produced;
            }
            if (rc)
                return rc;
            ci->rows.resize(produced * ow);
            if (produced)
                PJ_HIP(hipMemcpy(ci->rows.data(), h->d_out_batch.p, produced * ow, hipMemcpyDeviceToHost));
            ci->total = produced;
        }
    }
    if (!ci) {
        cur->pending = 0;
        return PJ_OK;
    }
    const uint64_t take = std::min<uint64_t>(out_cap, ci->total - ci->next);
    if (take)
        memcpy(out, ci->rows.data() + ci->next * ow, take * ow);
    ci->next += take;
    *out_n = take;
    cur->pending = ci->total - ci->next;
    return PJ_OK;
}

int pj_cursor_release(pj_cursor* cur)
{
    if (cur && cur->impl) {
        delete (CursorImpl*)cur->impl;
        cur->impl = nullptr;
        cur->pending = 0;
    }
    return PJ_OK;
}

int pj_join_reset(pj_join_t h)
{
    if (!h)
        return fail(PJ_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(h->mu);
    std::lock_guard<std::mutex> g2(h->stage_mu);
    h->staged.clear();
    h->staged.shrink_to_fit();
    h->built = false;
    h->R = Partitioned();
    return PJ_OK;
}

int pj_join_last_timings(pj_join_t h, pj_timings* t)
{
    if (!h || !t)
        return fail(PJ_ERR_INVALID, "null argument");
    *t = h->last;
    return PJ_OK;
}

int pj_join_run_device_noalloc(pj_join_t h, const void* d_build, uint64_t nbuild, const void* d_probe,
                               uint64_t nprobe, void* d_out, uint64_t out_cap, uint64_t* out_n, void* stream)
{
    if (!h || !out_n || (nbuild && !d_build) || (nprobe && !d_probe) || (out_cap && !d_out))
        return fail(PJ_ERR_INVALID, "null argument");
    if (out_cap && ((uintptr_t)d_out % 16))
        return fail(PJ_ERR_INVALID, "the output buffer must be 16-byte aligned");
    *out_n = 0;
    if (nbuild >= (1ull << 31) || nprobe >= (1ull << 31))
        return fail(PJ_ERR_UNSUPPORTED, "relations of 2^31 tuples or more");
    if (!h->host_ctr)
        return fail(PJ_ERR_STATE, "the handle has no pinned counter words (created without a device?)");
    ScratchSlot* slot = nullptr;
    int rc;
    if ((rc = scratch_slot(&slot)))
        return rc;
    std::lock_guard<std::mutex> g(h->mu);
    if (h->built)
        return fail(PJ_ERR_STATE, "the handle holds a staged build side (pj_join_reset first)");
    const size_t need = noalloc_need(h, nbuild, nprobe);
    const size_t size = sizeof(scratch);
    if (need > size)
        return fail(PJ_ERR_UNSUPPORTED,
                    "a %llu x %llu join needs %zu bytes of scratch; the fixed scratch holds %zu (use "
                    "pj_join_run_device)",
                    (unsigned long long)nbuild, (unsigned long long)nprobe, need, size);
    std::lock_guard<std::mutex> gs(slot->mu);
    ScratchSwap sw(h, slot->ptr, size);
    if (h->lay.rowid) {   // the (build row, probe row) pairs take what the partition passes leave
        const size_t room = (size - need - 256) / 8 * 8;
        if (room < 8 * 1024)
            return fail(PJ_ERR_UNSUPPORTED, "the fixed scratch leaves no room for the join's row pairs");
        if ((rc = h->pairs.ensure(room)))
            return rc;
    }
    hipStream_t st = as_stream(stream);
    h->model_ms = 0.f;
    h->raw_build = d_build;
    h->raw_nbuild = nbuild;
    h->plan.ok = false;
    record(h, 0, st);
    if ((rc = prepare_radix(h, d_build, nbuild, st)))
        return rc;
    record(h, 1, st);
    return probe_device(h, d_probe, nprobe, d_out, out_cap, out_n, st);
}

}  // extern "C"
