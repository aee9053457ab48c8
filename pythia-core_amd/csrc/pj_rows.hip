// pj_rows.hip -- the row-id form's second pass: key check and projection on
// the tuples themselves.
//
// What it replaces in the reference (paths relative to its tree):
//   ConjunctionEqualsEvaluator::eval over the key columns
//       conjunctionevaluator.h:65-73, comparator.h:76-80,
//       rawcompfns.cpp:48 (IntIntEqual), :78 (IntLongEqual),
//       :141 (LongIntEqual), :171 (LongLongEqual; DATE compares as LONG)
//   HashJoinOp::constructOutputTuple / Schema::writeData
//       operators/join.cpp:648-679, schema.inl:47-93
// The join kernel pairs build and probe rows whose 64-bit key hashes are
// equal. Here every pair's key columns are compared on the packed tuples, by
// value after widening INTEGER columns to 64 bits, and the projected columns
// are copied into the packed output row as writeData copies them (bit copies;
// strncpy for CHAR), so keys of any number of columns and projections of any
// width and type come out exactly as the reference writes them.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pj_device.h"
#include "pj_internal.h"

namespace pj {

namespace {

constexpr uint32_t kRowThreads = 256;
constexpr uint32_t kCompactThreads = 1024;

__device__ __forceinline__ int64_t row_key_value(const uint8_t* p, uint32_t is_int)
{
    return is_int ? (int64_t)(int32_t)ld32u(p) : (int64_t)ld64u(p);
}

// Every key column equal (the loop is unrolled over the largest key, so the
// specification stays in kernel-argument registers).
__device__ __forceinline__ bool row_keys_equal(const uint8_t* bt, const uint8_t* pt, const RowSpec& s)
{
    bool eq = true;
#pragma unroll
    for (uint32_t k = 0; k < PJ_MAX_KEYS; ++k)
        if (k < s.nk)
            eq = eq && row_key_value(bt + s.key[k].boff, s.key[k].bint) ==
                           row_key_value(pt + s.key[k].poff, s.key[k].pint);
    return eq;
}

// The projected columns of (bt, pt) into the output row at o.
template <bool WORDS>
__device__ __forceinline__ void row_copy(uint8_t* __restrict__ o, const uint8_t* bt, const uint8_t* pt,
                                         const RowSpec& s)
{
#pragma unroll
    for (uint32_t j = 0; j < PJ_MAX_COLS; ++j) {
        if (j >= s.ncopy)
            break;
        const RowCopy c = s.copy[j];
        const uint8_t* src = (c.side ? pt : bt) + c.src;
        uint8_t* dst = o + c.dst;
        if (c.chr) {   // writeData's strncpy: the bytes up to the first NUL, then zeros
            bool end = false;
            for (uint32_t b = 0; b < c.width; ++b) {
                const uint8_t v = end ? 0 : src[b];
                end = end || v == 0;
                dst[b] = v;
            }
        } else if (WORDS) {
            const uint32_t* s4 = reinterpret_cast<const uint32_t*>(src);
            uint32_t* d4 = reinterpret_cast<uint32_t*>(dst);
            for (uint32_t w = 0; w < c.width / 4u; ++w)
                d4[w] = s4[w];
        } else {
            for (uint32_t b = 0; b < c.width; ++b)
                dst[b] = src[b];
        }
    }
}

// Pair i -> output row i. A pair whose keys differ (equal hashes of unequal
// keys) is counted in *nfail; its row is left unwritten.
template <bool WORDS>
__global__ __launch_bounds__(kRowThreads) void k_rows(const uint2* __restrict__ pairs,
                                                       const unsigned long long* __restrict__ ncand, uint64_t pair_cap,
                                                       const uint8_t* __restrict__ B, const uint8_t* __restrict__ P,
                                                       RowSpec spec, uint8_t* __restrict__ out, uint64_t out_cap,
                                                       uint32_t* __restrict__ nfail)
{
    const uint64_t n = std::min<uint64_t>(*ncand, pair_cap);
    uint32_t bad = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * kRowThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kRowThreads) {
        const uint2 pr = pairs[i];
        const uint8_t* bt = B + (uint64_t)pr.x * spec.bw;
        const uint8_t* pt = P + (uint64_t)pr.y * spec.pw;
        if (!row_keys_equal(bt, pt, spec)) {
            ++bad;
            continue;
        }
        if (i < out_cap)
            row_copy<WORDS>(out + i * spec.ow, bt, pt, spec);
    }
    for (int d = 32; d > 0; d >>= 1)
        bad += __shfl_xor(bad, d, 64);
    if ((threadIdx.x & 63) == 0 && bad)
        atomicAdd(nfail, bad);
}

// The pairs whose keys are equal, written densely: each block reserves the
// rows of its 1024 pairs with one atomic on *rows.
template <bool WORDS>
__global__ __launch_bounds__(kCompactThreads) void k_rows_compact(const uint2* __restrict__ pairs,
                                                                   const unsigned long long* __restrict__ ncand,
                                                                   uint64_t pair_cap, const uint8_t* __restrict__ B,
                                                                   const uint8_t* __restrict__ P, RowSpec spec,
                                                                   uint8_t* __restrict__ out, uint64_t out_cap,
                                                                   unsigned long long* __restrict__ rows)
{
    __shared__ uint32_t wpre[kCompactThreads / 64];
    __shared__ unsigned long long base_s;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    const uint64_t n = std::min<uint64_t>(*ncand, pair_cap);
    for (uint64_t t0 = (uint64_t)blockIdx.x * kCompactThreads; t0 < n; t0 += (uint64_t)gridDim.x * kCompactThreads) {
        const uint64_t i = t0 + threadIdx.x;
        bool ok = false;
        const uint8_t *bt = B, *pt = P;
        if (i < n) {
            const uint2 pr = pairs[i];
            bt = B + (uint64_t)pr.x * spec.bw;
            pt = P + (uint64_t)pr.y * spec.pw;
            ok = row_keys_equal(bt, pt, spec);
        }
        const uint64_t m = __ballot(ok);
        if (lane == 0)
            wpre[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t t = 0;
            for (uint32_t w = 0; w < kCompactThreads / 64; ++w) {
                const uint32_t c = wpre[w];
                wpre[w] = t;
                t += c;
            }
            base_s = t ? atomicAdd(rows, (unsigned long long)t) : 0ull;
        }
        __syncthreads();
        if (ok) {
            const uint64_t row = base_s + wpre[wave] + (uint64_t)__popcll(m & below);
            if (row < out_cap)
                row_copy<WORDS>(out + row * spec.ow, bt, pt, spec);
        }
        __syncthreads();
    }
}

}  // namespace

int rows_materialize(const uint2* pairs, const unsigned long long* ncand, uint64_t pair_cap, const void* d_build,
                     const void* d_probe, const RowSpec& spec, void* d_out, uint64_t out_cap,
                     unsigned long long* rows, uint32_t* nfail, bool compact, hipStream_t st)
{
    if (pair_cap == 0)
        return PJ_OK;
    const uint8_t* B = (const uint8_t*)d_build;
    const uint8_t* P = (const uint8_t*)d_probe;
    uint8_t* out = (uint8_t*)d_out;
    if (compact) {
        const uint32_t grid = (uint32_t)std::min<uint64_t>((pair_cap + kCompactThreads - 1) / kCompactThreads, 2048);
        if (spec.words)
            hipLaunchKernelGGL(k_rows_compact<true>, dim3(grid), dim3(kCompactThreads), 0, st, pairs, ncand, pair_cap,
                               B, P, spec, out, out_cap, rows);
        else
            hipLaunchKernelGGL(k_rows_compact<false>, dim3(grid), dim3(kCompactThreads), 0, st, pairs, ncand,
                               pair_cap, B, P, spec, out, out_cap, rows);
    } else {
        const uint32_t grid = (uint32_t)std::min<uint64_t>((pair_cap + kRowThreads - 1) / kRowThreads, 4096);
        if (spec.words)
            hipLaunchKernelGGL(k_rows<true>, dim3(grid), dim3(kRowThreads), 0, st, pairs, ncand, pair_cap, B, P, spec,
                               out, out_cap, nfail);
        else
            hipLaunchKernelGGL(k_rows<false>, dim3(grid), dim3(kRowThreads), 0, st, pairs, ncand, pair_cap, B, P,
                               spec, out, out_cap, nfail);
    }
    PJ_HIP(hipGetLastError());
    return PJ_OK;
}

}  // namespace pj
