// pj_device.h -- device helpers shared by the partition, join and model
// kernels: the canonical record, streaming loads and stores, and the two
// partition functions (radix bits of a key mix, or the learned function).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pj_internal.h"

namespace pj {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// Streaming (non-temporal) 16-byte load and 8-byte store.
__device__ __forceinline__ uint4 ld_stream(const uint4* p)
{
    u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st_stream2(uint32_t* p, uint32_t a, uint32_t b)
{
    u32x2 v = {a, b};
    __builtin_nontemporal_store(v, reinterpret_cast<u32x2*>(p));
}

__device__ __forceinline__ uint64_t rec_key(uint4 r) { return (uint64_t)r.x | ((uint64_t)r.y << 32); }

// The key as an unsigned value whose order is the signed order of the
// normalised 64-bit key: range partitions are ranges of this value.
__host__ __device__ __forceinline__ uint64_t ord_key(uint64_t k) { return k ^ 0x8000000000000000ull; }

// Key mix used for radix partition, sub-partition and LDS bucket bits.
__device__ __forceinline__ uint64_t key_mix(uint64_t k) { return pj_fin64(k); }

__device__ __forceinline__ uint32_t radix_part(uint64_t h, uint32_t bits)
{
    return bits ? (uint32_t)(h >> (64 - bits)) : 0u;
}

// Loads at any byte address (packed tuples put columns at any offset): one
// 32-bit load when the address is 4-byte aligned, else four byte loads.
__device__ __forceinline__ uint32_t ld32u(const uint8_t* p)
{
    if (((uintptr_t)p & 3u) == 0)
        return *reinterpret_cast<const uint32_t*>(p);
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
__device__ __forceinline__ uint64_t ld64u(const uint8_t* p)
{
    return (uint64_t)ld32u(p) | ((uint64_t)ld32u(p + 4) << 32);
}

// Hash of a tuple's key columns, each normalised to int64 (an INTEGER
// column sign-extended), so that equal keys of either width hash alike.
__device__ __forceinline__ uint64_t row_key_hash(const uint8_t* t, const SideLayout& L)
{
    uint64_t h = 0x9e3779b97f4a7c15ull * (L.nk + 1);
    for (uint32_t k = 0; k < L.nk; ++k) {
        const uint8_t* p = t + L.koff[k];
        const uint64_t v = L.kint[k] ? (uint64_t)(int64_t)(int32_t)ld32u(p) : ld64u(p);
        h = pj_fin64(h ^ v) + 0xd6e8feb86659fd93ull;
    }
    return pj_fin64(h);
}

// Record modes of the partition passes: the tuple's own words (REC_WORDS),
// the tuple is the record (REC_IDENT), or the row-id form (REC_HASH).
enum { REC_WORDS = 0, REC_IDENT = 1, REC_HASH = 2 };

template <int RM>
__device__ __forceinline__ uint4 load_rec(const uint32_t* __restrict__ in, uint64_t i, const SideLayout& L)
{
    if (RM == REC_IDENT) {
        return ld_stream(reinterpret_cast<const uint4*>(in) + i);
    } else if (RM == REC_HASH) {
        const uint64_t h = row_key_hash(reinterpret_cast<const uint8_t*>(in) + i * L.tbytes, L) & L.hmask;
        return make_uint4((uint32_t)h, (uint32_t)(h >> 32), (uint32_t)i, 0u);
    } else {
        const uint32_t* t = in + i * L.tw;
        uint4 r;
        r.x = t[L.k0];
        r.y = L.k1 >= 0 ? t[L.k1] : (uint32_t)((int32_t)r.x >> 31);
        r.z = L.p0 >= 0 ? t[L.p0] : 0u;
        r.w = L.p1 >= 0 ? t[L.p1] : 0u;
        return r;
    }
}

// Radix partition of a key: the top pf.bits bits of the key mix.
__device__ __forceinline__ uint32_t part_of(uint64_t key, const PartFn& pf)
{
    return radix_part(key_mix(key), pf.bits);
}

// The learned partition function over a copy of its blob (pj_internal.h).
// Every step is one single-precision operation rounded to nearest, so that
// the build side's and the probe side's kernels, compiled apart, put every
// key in the same partition. __fmul_rn and __fsub_rn are plain * and - in
// this toolchain's headers, and hipcc's default contraction (fast, which
// ignores #pragma clang fp contract) fuses such a multiply and subtract into
// an FMA where it sees fit, differently in each kernel they are inlined into:
// two evaluations of this map in one kernel disagreed by one partition on
// 214 of 64M clustered keys that way. The library is therefore built with
// -ffp-contract=off; so built, the device's partitions equal an IEEE
// single-precision evaluation on the host for every one of those keys.
struct MapView {
    const uint8_t* base;
    __device__ __forceinline__ const MapHdr& hdr() const { return *reinterpret_cast<const MapHdr*>(base); }
    __device__ __forceinline__ uint64_t seg_lo(uint32_t s) const
    {
        return reinterpret_cast<const uint64_t*>(base + kMapOffSegLo)[s];
    }
    __device__ __forceinline__ uint64_t seg_cut(uint32_t s) const
    {
        return reinterpret_cast<const uint64_t*>(base + kMapOffSegCut)[s];
    }
    // The build partition of ordered key u: the segment that owns u (the
    // last whose cut is at or below it, found from the directory slice of u
    // and a binary search inside it), the leaf of u's position in it, and
    // that leaf's line, clamped to the segment's partitions.
    __device__ __forceinline__ uint32_t fine(uint64_t u) const
    {
        const MapHdr& h = hdr();
        uint32_t s = 0;
        if (u >= h.lo0) {
            const uint16_t* dir = reinterpret_cast<const uint16_t*>(base + kMapOffDir);
            uint64_t d = (u - h.lo0) >> h.dshift;
            if (d >= kMapDir)
                d = kMapDir - 1;
            uint32_t lo = dir[d], hi = dir[d + 1];
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) >> 1;
                if (seg_cut(mid) <= u)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            s = lo;
        }
        const uint2 meta = reinterpret_cast<const uint2*>(base + kMapOffSegMeta)[s];
        const uint32_t leaf0 = meta.x & 0xFFFFu, nl = meta.x >> 16, first = meta.y & 0xFFFFu, np = meta.y >> 16;
        const float inv = reinterpret_cast<const float*>(base + kMapOffSegInv)[s];
        const uint64_t slo = seg_lo(s);
        // |u - slo| rounded to single precision, from 32 bits where it fits
        // (the same value as from 64 bits, in one conversion)
        const uint64_t dist = u >= slo ? u - slo : slo - u;
        const float fd = dist < (1ull << 32) ? __uint2float_rn((uint32_t)dist) : __ull2float_rn(dist);
        const float fl = u >= slo ? __fmul_rn(fd, inv) : -__fmul_rn(fd, inv);
        const uint32_t j = fl < 0.f ? 0u : fl < (float)nl ? (uint32_t)fl : nl - 1;
        const float x = __fsub_rn(fl, (float)j);
        const float2 cd = reinterpret_cast<const float2*>(base + kMapOffLeaf)[leaf0 + j];
        const float pf = __fmaf_rn(cd.y, x, cd.x);
        const uint32_t p = pf > 0.f ? (pf < (float)np ? (uint32_t)pf : np - 1) : 0u;
        return first + (p < np ? p : np - 1);
    }
};

// The probe partition of build partition fine (below 2^16) under pf.
__device__ __forceinline__ uint32_t coarse_of_fine(uint32_t fine, const PartFn& pf)
{
    return pf.div > 1 ? __umulhi(fine, pf.divm) : fine;
}

// Copies the blob into LDS (the caller synchronises before the first use).
__device__ __forceinline__ MapView load_map(uint8_t* lds, const uint8_t* map)
{
    const uint4* src = reinterpret_cast<const uint4*>(map);
    uint4* dst = reinterpret_cast<uint4*>(lds);
    for (uint32_t i = threadIdx.x; i < kMapBytes / 16; i += blockDim.x)
        dst[i] = src[i];
    return MapView{lds};
}

}  // namespace pj
