// pj_probe.h -- the probe side shared by the join kernels: output row
// assembly and the round-based streaming probe with wave-aggregated output
// reservation.
//
// A table view T provides lookup(s, nm, first) (count a probe record's
// matches and return the slot of the first), mkey(s) / match(j, mk) / next(j)
// (walk later slots for many-to-many matches), bkey(j, s) (the build key of a
// matching slot) and bpay(j) (its payload words).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pj_device.h"

namespace pj {

constexpr int kProbeItems = 8;

__device__ __forceinline__ uint32_t pick_word(uint64_t bk, uint2 bp, const uint4& s, uint32_t src)
{
    switch (src) {
    case SRC_BK0: return (uint32_t)bk;
    case SRC_BK1: return (uint32_t)(bk >> 32);
    case SRC_BP0: return bp.x;
    case SRC_BP1: return bp.y;
    case SRC_PK0: return s.x;
    case SRC_PK1: return s.y;
    case SRC_PP0: return s.z;
    default: return s.w;
    }
}

__device__ __forceinline__ void write_row(uint32_t* __restrict__ out, uint64_t row, const OutFormat& f, bool kbp,
                                          uint64_t bk, uint2 bp, const uint4& s)
{
    uint32_t* o = out + row * f.words;
    if (kbp) {   // (build key, build payload, probe payload), 24 bytes
        st_stream2(o, (uint32_t)bk, (uint32_t)(bk >> 32));
        st_stream2(o + 2, bp.x, bp.y);
        st_stream2(o + 4, s.z, s.w);
        return;
    }
    uint32_t w[8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
        w[i] = (uint32_t)i < f.words ? pick_word(bk, bp, s, f.src[i]) : 0u;
    if (f.words == 4) {   // 16-byte rows leave as one vector store
        u32x4 v = {w[0], w[1], w[2], w[3]};
        __builtin_nontemporal_store(v, reinterpret_cast<u32x4*>(o));
        return;
    }
    if ((f.words & 1u) == 0) {
#pragma unroll
        for (int i = 0; i < 8; i += 2)
            if ((uint32_t)i < f.words)
                st_stream2(o + i, w[i], w[i + 1]);
        return;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i)
        if ((uint32_t)i < f.words)
            o[i] = w[i];
}

// Streams the probe records [0, s_cnt) of Sp past a built table (NT
// threads a block). The probe runs in rounds of NT * PI records (PI = 8 by
// default): every wave probes a 64 * PI-record slice, the
// block reserves the round's output rows with one atomic, and the lanes of
// each wave write adjacent output rows. Loads are software-pipelined: a
// record's slot is refilled with the next round's record as soon as its
// lookup is done.
//
// CL: the records are not contiguous but listed by chunk: record i is record
// i % kChunkRecs of chunk clp[i / kChunkRecs] of Sp (chunks of kChunkRecs
// records each). Lane l of a wave holds the chunk number of its slice's
// chunk l, loaded two rounds ahead of the records it locates, and hands it
// to the lanes that load that chunk's records.
constexpr uint32_t kChunkRecs = 8;

template <int NT, class T, int PI = kProbeItems, bool CL = false>
__device__ __forceinline__ void probe_rounds(const T& tab, const uint4* __restrict__ Sp, uint32_t s_cnt,
                                             uint32_t* __restrict__ out, uint64_t out_cap,
                                             unsigned long long* __restrict__ out_cursor, const OutFormat& fmt,
                                             bool kbp, uint32_t* wsum, uint32_t* wbase_lds,
                                             const uint32_t* __restrict__ clp = nullptr)
{
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    constexpr uint32_t kSlice = 64 * PI;
    constexpr uint32_t kRound = (NT / 64) * kSlice;
    constexpr uint32_t kSliceChunks = kSlice / kChunkRecs;
    static_assert(!CL || (kSlice % kChunkRecs == 0 && kSliceChunks <= 64), "a slice's chunks fit a wave");
    // chunk number of the lane's chunk in this wave's slice of round r0
    auto chunk_at = [&](uint32_t r0) -> uint32_t {
        const uint32_t c = (r0 + wave * kSlice) / kChunkRecs + lane;
        return lane < kSliceChunks && c * kChunkRecs < s_cnt ? clp[c] : 0u;
    };
    // address of item k of this lane in the slice whose chunk numbers are in ch
    auto rec_at = [&](uint32_t ch, int k) -> const uint4* {
        const uint32_t c = __shfl(ch, k * (64 / kChunkRecs) + (lane / kChunkRecs), 64);
        return Sp + ((uint64_t)c * kChunkRecs + (lane % kChunkRecs));
    };
    uint32_t ch_next = 0;   // CL: chunk numbers of the next round's slice
    uint4 nxt[PI];
    // The first round's loads are unconditional, so that they leave back to
    // back: a lane past s_cnt loads the last record again (the list's first
    // chunk when CL), and its result is never looked up.
    if (CL) {
        const uint32_t ch0 = chunk_at(0);
        ch_next = chunk_at(kRound);
#pragma unroll
        for (int k = 0; k < PI; ++k)
            nxt[k] = ld_stream(rec_at(ch0, k));
    } else if (s_cnt) {
#pragma unroll
        for (int k = 0; k < PI; ++k) {
            const uint32_t i = wave * kSlice + k * 64 + lane;
            nxt[k] = ld_stream(Sp + (i < s_cnt ? i : s_cnt - 1));
        }
    }
    for (uint32_t r0 = 0; r0 < s_cnt; r0 += kRound) {
        const uint32_t t0 = r0 + wave * kSlice;
        const uint32_t ch_after = CL ? chunk_at(r0 + 2 * kRound) : 0u;
        uint4 s[PI];
        uint32_t nm[PI], first[PI];
        bool multi = false;
#pragma unroll
        for (int k = 0; k < PI; ++k) {
            const uint32_t i = t0 + k * 64 + lane;
            s[k] = nxt[k];
            nm[k] = 0;
            first[k] = 0;
            if (i < s_cnt)
                tab.lookup(s[k], nm[k], first[k]);
            multi |= nm[k] > 1;
            if (CL) {
                const uint4* a = rec_at(ch_next, k);
                if (i + kRound < s_cnt)
                    nxt[k] = ld_stream(a);
            } else if (i + kRound < s_cnt) {
                nxt[k] = ld_stream(Sp + i + kRound);
            }
        }
        ch_next = ch_after;
        const bool wmulti = __any(multi);
        uint64_t m[PI];
        uint32_t mine = 0, pre = 0, wtotal = 0;
        if (!wmulti) {
#pragma unroll
            for (int k = 0; k < PI; ++k) {
                m[k] = __ballot(nm[k] != 0);
                wtotal += (uint32_t)__popcll(m[k]);
            }
        } else {
#pragma unroll
            for (int k = 0; k < PI; ++k)
                mine += nm[k];
            uint32_t x = mine;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t y = __shfl_up(x, d, 64);
                if (lane >= (uint32_t)d)
                    x += y;
            }
            pre = x - mine;
            wtotal = __shfl(x, 63);
        }
        if (lane == 0)
            wsum[wave] = wtotal;
        __syncthreads();
        if (wave == 0) {
            const uint32_t v = lane < NT / 64 ? wsum[lane] : 0u;
            uint32_t x = v;
#pragma unroll
            for (int d = 1; d < (int)(NT / 64); d <<= 1) {
                const uint32_t y = __shfl_up(x, d, 64);
                if (lane >= (uint32_t)d)
                    x += y;
            }
            const uint32_t btotal = __shfl(x, NT / 64 - 1);
            unsigned long long base = 0;
            if (lane == 0 && btotal)
                base = atomicAdd(out_cursor, (unsigned long long)btotal);
            base = __shfl(base, 0);
            if (lane < NT / 64) {
                const uint64_t wb = base + (x - v);
                wbase_lds[2 * lane] = (uint32_t)wb;
                wbase_lds[2 * lane + 1] = (uint32_t)(wb >> 32);
            }
        }
        __syncthreads();
        if (!wtotal)
            continue;
        const uint64_t wbase = (uint64_t)wbase_lds[2 * wave] | ((uint64_t)wbase_lds[2 * wave + 1] << 32);
        if (wmulti) {
            // many-to-many: a lane writes all matches of its probes
            uint64_t row = wbase + pre;
#pragma unroll
            for (int k = 0; k < PI; ++k) {
                const uint64_t mk = tab.mkey(s[k]);
                uint32_t left = nm[k];
                for (uint32_t j = first[k]; left; j = tab.next(j))
                    if (tab.match(j, mk)) {
                        if (row < out_cap)
                            write_row(out, row, fmt, false, tab.bkey(j, s[k]), tab.bpay(j), s[k]);
                        row++;
                        left--;
                    }
            }
            continue;
        }
        // at most one match a probe: lanes of a wave write adjacent rows
        uint64_t row0 = wbase;
#pragma unroll
        for (int k = 0; k < PI; ++k) {
            if (nm[k]) {
                const uint64_t row = row0 + (uint64_t)__popcll(m[k] & below);
                if (row < out_cap)
                    write_row(out, row, fmt, kbp, tab.bkey(first[k], s[k]), tab.bpay(first[k]), s[k]);
            }
            row0 += (uint64_t)__popcll(m[k]);
        }
    }
}

}  // namespace pj
