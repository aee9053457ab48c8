// pj_smj.hip -- sort-merge join on the device, the counterpart of the
// reference's SortMergeJoinOp / MPSMJoinOp (operators/join.cpp:905-1791):
// both inputs sorted by the join key, then merged; the output multiset is
// the hash join's. It takes the same descriptor and writes the same rows as
// pj_join_run_device, so the two can be timed against each other.
//
//   k_smj_extract  tuples -> (key, payload) arrays, and the keys' min / max
//   radix sort     rocPRIM's pair sort of each side, over only the low bits in
//                  which that side's keys differ (the bits above the highest
//                  bit of min ^ max are equal in every key of the side)
//   k_smj_count    a block takes a run of sorted probe keys and the window of
//                  build keys they can meet; each key's run of equal build
//                  keys (start and length) by binary search in the window
//   exclusive scan the run lengths -> output row of each probe record
//   k_smj_emit     one output row per (probe record, build record of its run)
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <mutex>

#include "pj_device.h"
#include "pj_internal.h"
#include "pj_probe.h"

namespace pj {
namespace {

constexpr uint32_t kSmjThreads = 256;
constexpr uint32_t kSmjPer = 16;   // sorted probe records a thread of k_smj_count (on average)

template <bool IDENT>
__global__ __launch_bounds__(kSmjThreads) void k_smj_extract(const uint32_t* __restrict__ in, uint64_t n,
                                                              SideLayout L, uint64_t* __restrict__ keys,
                                                              uint64_t* __restrict__ vals,
                                                              unsigned long long* __restrict__ minmax)
{
    uint64_t lo = ~0ull, hi = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * kSmjThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kSmjThreads) {
        const uint4 r = load_rec<IDENT>(in, i, L);
        const uint64_t k = rec_key(r);
        keys[i] = k;
        vals[i] = (uint64_t)r.z | ((uint64_t)r.w << 32);
        lo = k < lo ? k : lo;
        hi = k > hi ? k : hi;
    }
    for (int d = 32; d > 0; d >>= 1) {
        const uint64_t a = __shfl_xor(lo, d, 64), b = __shfl_xor(hi, d, 64);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&minmax[0], (unsigned long long)lo);
        atomicMax(&minmax[1], (unsigned long long)hi);
    }
}

__global__ void k_smj_minmax_init(unsigned long long* minmax)
{
    if (threadIdx.x < 4)
        minmax[threadIdx.x] = (threadIdx.x & 1) ? 0ull : ~0ull;
}

// Smallest m in [j, n] with a[m] >= k (GT: a[m] > k), for a sorted
// ascending: steps of 1, 2, 4, ... from j, then a binary search.
template <bool GT>
__device__ __forceinline__ uint64_t gallop(const uint64_t* __restrict__ a, uint64_t n, uint64_t j, uint64_t k)
{
    auto before = [&](uint64_t m) { return GT ? a[m] <= k : a[m] < k; };
    if (j >= n || !before(j))
        return j;
    uint64_t lo = j, step = 1;   // invariant: before(lo)
    uint64_t hi = n;
    while (lo + step < n) {
        if (!before(lo + step)) {
            hi = lo + step;
            break;
        }
        lo += step;
        step <<= 1;
    }
    // before(lo), and hi == n or !before(hi): the answer is in (lo, hi]
    while (hi - lo > 1) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (before(mid))
            lo = mid;
        else
            hi = mid;
    }
    return hi;
}

// Smallest m in [a0, a1) with a[m] >= k (GT: a[m] > k), or a1.
template <bool GT>
__device__ __forceinline__ uint64_t bound_in(const uint64_t* __restrict__ a, uint64_t a0, uint64_t a1, uint64_t k)
{
    while (a0 < a1) {
        const uint64_t mid = a0 + (a1 - a0) / 2;
        if (GT ? a[mid] <= k : a[mid] < k)
            a0 = mid + 1;
        else
            a1 = mid;
    }
    return a0;
}

// A block takes kSmjPer * kSmjThreads consecutive sorted probe keys: thread 0
// finds the window of build keys they can meet (galloping from 0 to the
// first key's run, then to the end of the last key's run), and every thread
// binary-searches its keys (adjacent keys on adjacent lanes) in that window.
__global__ __launch_bounds__(kSmjThreads) void k_smj_count(const uint64_t* __restrict__ rk, uint64_t nr,
                                                            const uint64_t* __restrict__ sk, uint64_t ns,
                                                            uint32_t* __restrict__ lo, uint32_t* __restrict__ cnt)
{
    __shared__ uint64_t win[2];
    const uint64_t i0 = (uint64_t)blockIdx.x * kSmjPer * kSmjThreads;
    if (i0 >= ns)
        return;
    const uint64_t i1 = i0 + (uint64_t)kSmjPer * kSmjThreads < ns ? i0 + (uint64_t)kSmjPer * kSmjThreads : ns;
    if (threadIdx.x == 0) {
        const uint64_t w0 = gallop<false>(rk, nr, 0, sk[i0]);
        win[0] = w0;
        win[1] = gallop<true>(rk, nr, w0, sk[i1 - 1]);
    }
    __syncthreads();
    const uint64_t w0 = win[0], w1 = win[1];
    for (uint64_t i = i0 + threadIdx.x; i < i1; i += kSmjThreads) {
        const uint64_t k = sk[i];
        const uint64_t j = bound_in<false>(rk, w0, w1, k);
        uint64_t e = j;   // end of the run: usually at most one build key (a primary key)
        if (e < w1 && rk[e] == k && ++e < w1 && rk[e] == k)
            e = bound_in<true>(rk, e, w1, k);
        lo[i] = (uint32_t)j;
        cnt[i] = (uint32_t)(e - j);
    }
}

__global__ __launch_bounds__(kSmjThreads) void k_smj_emit(const uint64_t* __restrict__ rk,
                                                           const uint64_t* __restrict__ rv,
                                                           const uint64_t* __restrict__ sk,
                                                           const uint64_t* __restrict__ sv, uint64_t ns,
                                                           const uint32_t* __restrict__ lo,
                                                           const uint32_t* __restrict__ cnt,
                                                           const uint64_t* __restrict__ off, uint32_t* __restrict__ out,
                                                           OutFormat fmt, uint32_t kbp)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kSmjThreads + threadIdx.x; i < ns;
         i += (uint64_t)gridDim.x * kSmjThreads) {
        const uint32_t c = cnt[i];
        if (!c)
            continue;
        const uint64_t k = sk[i], v = sv[i];
        const uint4 s = make_uint4((uint32_t)k, (uint32_t)(k >> 32), (uint32_t)v, (uint32_t)(v >> 32));
        const uint64_t row = off[i], b = lo[i];
        for (uint32_t m = 0; m < c; ++m) {
            const uint64_t bvv = rv[b + m];
            write_row(out, row + m, fmt, kbp != 0, rk[b + m], make_uint2((uint32_t)bvv, (uint32_t)(bvv >> 32)), s);
        }
    }
}

// Low bits in which keys between mn and mx (unsigned) can differ.
uint32_t key_bits(uint64_t mn, uint64_t mx)
{
    const uint64_t d = mn ^ mx;
    return d ? 64u - (uint32_t)__builtin_clzll(d) : 1u;
}

}  // namespace
}  // namespace pj

using namespace pj;

struct pj_smj_s {
    Layout lay;
    pj_schema desc_build, desc_probe;
    std::mutex mu;
    DevBuf keys[2][2], vals[2][2];   // [side][unsorted, sorted]
    DevBuf lo, cnt, off, temp, minmax;
    unsigned long long* host = nullptr;   // pinned: min, max of each side, then the row total
    pj_smj_timings last = {};
    hipEvent_t ev[6];
    bool ev_ok = false;
    DevBuf host_b, host_p, host_o;   // the host path's device copies
    std::mutex host_mu;              // guards them
};

namespace {

int side_records(pj_smj_s* h, int side, const void* d_in, uint64_t n, const SideLayout& L, hipStream_t st)
{
    const bool ident = L.ident && ((uintptr_t)d_in % 16 == 0);
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + kSmjThreads - 1) / kSmjThreads, 4096);
    unsigned long long* mm = (unsigned long long*)h->minmax.p + 2 * side;
    if (ident)
        hipLaunchKernelGGL(k_smj_extract<true>, dim3(grid), dim3(kSmjThreads), 0, st, (const uint32_t*)d_in, n, L,
                           (uint64_t*)h->keys[side][0].p, (uint64_t*)h->vals[side][0].p, mm);
    else
        hipLaunchKernelGGL(k_smj_extract<false>, dim3(grid), dim3(kSmjThreads), 0, st, (const uint32_t*)d_in, n, L,
                           (uint64_t*)h->keys[side][0].p, (uint64_t*)h->vals[side][0].p, mm);
    PJ_HIP(hipGetLastError());
    return PJ_OK;
}

// Temporary bytes of a pair sort of n records over `bits` low key bits.
int sort_bytes(uint64_t n, uint32_t bits, size_t* bytes)
{
    *bytes = 0;
    PJ_HIP(rocprim::radix_sort_pairs(nullptr, *bytes, (uint64_t*)nullptr, (uint64_t*)nullptr, (uint64_t*)nullptr,
                                     (uint64_t*)nullptr, (size_t)n, 0u, bits, (hipStream_t)0));
    return PJ_OK;
}

// h->temp must already hold the sort's temporary bytes: growing it here
// would free memory that work queued earlier on the stream may still use.
int sort_side(pj_smj_s* h, int side, uint64_t n, uint32_t bits, hipStream_t st)
{
    uint64_t* k0 = (uint64_t*)h->keys[side][0].p;
    uint64_t* k1 = (uint64_t*)h->keys[side][1].p;
    uint64_t* v0 = (uint64_t*)h->vals[side][0].p;
    uint64_t* v1 = (uint64_t*)h->vals[side][1].p;
    size_t bytes = 0;
    int rc;
    if ((rc = sort_bytes(n, bits, &bytes)))
        return rc;
    if (bytes > h->temp.cap)
        return fail(PJ_ERR_STATE, "sort needs %zu temporary bytes, %zu reserved", bytes, h->temp.cap);
    PJ_HIP(rocprim::radix_sort_pairs(h->temp.p, bytes, k0, k1, v0, v1, (size_t)n, 0u, bits, st));
    return PJ_OK;
}

}  // namespace

extern "C" {

int pj_smj_create(const pj_join_desc* desc, pj_smj_t* out)
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
                    "the sort-merge join takes one INTEGER / LONG / DATE key or two INTEGER keys, at most 8 payload "
                    "bytes a side and output tuples of at most 32 bytes without CHAR columns");
    pj_smj_s* h = new (std::nothrow) pj_smj_s();
    if (!h)
        return fail(PJ_ERR_INVALID, "out of host memory");
    h->lay = L;
    h->desc_build = desc->build;
    h->desc_probe = desc->probe;
    *out = h;
    return PJ_OK;
}

int pj_smj_out_width(pj_smj_t h, uint32_t* width)
{
    if (!h || !width)
        return fail(PJ_ERR_INVALID, "null argument");
    *width = h->lay.out_width;
    return PJ_OK;
}

int pj_smj_run_device(pj_smj_t h, const void* d_build, uint64_t nbuild, const void* d_probe, uint64_t nprobe,
                      void* d_out, uint64_t out_cap, uint64_t* out_n, void* stream)
{
    if (!h || !out_n || (nbuild && !d_build) || (nprobe && !d_probe) || (out_cap && !d_out))
        return fail(PJ_ERR_INVALID, "null argument");
    if (out_cap && ((uintptr_t)d_out % 16))
        return fail(PJ_ERR_INVALID, "the output buffer must be 16-byte aligned");
    if (nbuild >= (1ull << 31) || nprobe >= (1ull << 31))
        return fail(PJ_ERR_UNSUPPORTED, "relations of 2^31 tuples or more");
    std::lock_guard<std::mutex> g(h->mu);
    hipStream_t st = (hipStream_t)stream;
    *out_n = 0;
    h->last = pj_smj_timings{};
    if (nbuild == 0 || nprobe == 0)
        return PJ_OK;
    int rc;
    if (!h->ev_ok) {
        for (int i = 0; i < 6; ++i)
            PJ_HIP(hipEventCreate(&h->ev[i]));
        h->ev_ok = true;
    }
    if (!h->host)
        PJ_HIP(hipHostMalloc((void**)&h->host, 64, hipHostMallocDefault));
    if ((rc = h->minmax.ensure(64)))
        return rc;
    // every buffer sized before the first launch (nothing is freed under queued work)
    size_t tb[3] = {0, 0, 0};
    if ((rc = sort_bytes(nbuild, 64, &tb[0])) || (rc = sort_bytes(nprobe, 64, &tb[1])))
        return rc;
    PJ_HIP(rocprim::exclusive_scan(nullptr, tb[2], (const uint32_t*)nullptr, (uint64_t*)nullptr, (uint64_t)0,
                                   (size_t)nprobe, rocprim::plus<uint64_t>(), st));
    if ((rc = h->temp.ensure(std::max<size_t>(std::max(tb[0], std::max(tb[1], tb[2])), 16))) ||
        (rc = h->lo.ensure(nprobe * 4)) || (rc = h->cnt.ensure(nprobe * 4)) || (rc = h->off.ensure(nprobe * 8)))
        return rc;
    for (int side = 0; side < 2; ++side) {
        const uint64_t n = side ? nprobe : nbuild;
        if ((rc = h->keys[side][0].ensure(n * 8)) || (rc = h->vals[side][0].ensure(n * 8)) ||
            (rc = h->keys[side][1].ensure(n * 8)) || (rc = h->vals[side][1].ensure(n * 8)))
            return rc;
    }
    PJ_HIP(hipEventRecord(h->ev[0], st));
    hipLaunchKernelGGL(k_smj_minmax_init, dim3(1), dim3(64), 0, st, (unsigned long long*)h->minmax.p);
    if ((rc = side_records(h, 0, d_build, nbuild, h->lay.build, st)) ||
        (rc = side_records(h, 1, d_probe, nprobe, h->lay.probe, st)))
        return rc;
    PJ_HIP(hipMemcpyAsync(h->host, h->minmax.p, 32, hipMemcpyDeviceToHost, st));
    PJ_HIP(hipEventRecord(h->ev[1], st));
    PJ_HIP(hipStreamSynchronize(st));
    const uint32_t rbits = key_bits(h->host[0], h->host[1]);
    const uint32_t sbits = key_bits(h->host[2], h->host[3]);
    if ((rc = sort_side(h, 0, nbuild, rbits, st)) || (rc = sort_side(h, 1, nprobe, sbits, st)))
        return rc;
    PJ_HIP(hipEventRecord(h->ev[2], st));
    const uint64_t* rk = (const uint64_t*)h->keys[0][1].p;
    const uint64_t* rv = (const uint64_t*)h->vals[0][1].p;
    const uint64_t* sk = (const uint64_t*)h->keys[1][1].p;
    const uint64_t* sv = (const uint64_t*)h->vals[1][1].p;
    uint32_t* lo = (uint32_t*)h->lo.p;
    uint32_t* cnt = (uint32_t*)h->cnt.p;
    uint64_t* off = (uint64_t*)h->off.p;
    const uint64_t per_block = (uint64_t)kSmjPer * kSmjThreads;
    hipLaunchKernelGGL(k_smj_count, dim3((uint32_t)((nprobe + per_block - 1) / per_block)), dim3(kSmjThreads), 0, st,
                       rk, nbuild, sk, nprobe, lo, cnt);
    PJ_HIP(hipGetLastError());
    size_t bytes = tb[2];
    PJ_HIP(rocprim::exclusive_scan(h->temp.p, bytes, cnt, off, (uint64_t)0, (size_t)nprobe,
                                   rocprim::plus<uint64_t>(), st));
    // the row total: the last probe record's row plus its run
    PJ_HIP(hipMemcpyAsync(h->host + 4, off + nprobe - 1, 8, hipMemcpyDeviceToHost, st));
    PJ_HIP(hipMemcpyAsync((char*)(h->host + 5), cnt + nprobe - 1, 4, hipMemcpyDeviceToHost, st));
    PJ_HIP(hipEventRecord(h->ev[3], st));
    PJ_HIP(hipStreamSynchronize(st));
    const uint64_t total = h->host[4] + (uint32_t)h->host[5];
    *out_n = total;
    if (total > out_cap)
        return fail(PJ_ERR_CAPACITY, "%llu output tuples, the buffer holds %llu", (unsigned long long)total,
                    (unsigned long long)out_cap);
    const OutFormat& f = h->lay.out;
    const bool kbp = f.words == 6 && f.src[0] == SRC_BK0 && f.src[1] == SRC_BK1 && f.src[2] == SRC_BP0 &&
                     f.src[3] == SRC_BP1 && f.src[4] == SRC_PP0 && f.src[5] == SRC_PP1;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((nprobe + kSmjThreads - 1) / kSmjThreads, 8192);
    hipLaunchKernelGGL(k_smj_emit, dim3(grid), dim3(kSmjThreads), 0, st, rk, rv, sk, sv, nprobe, lo, cnt, off,
                       (uint32_t*)d_out, f, kbp ? 1u : 0u);
    PJ_HIP(hipGetLastError());
    PJ_HIP(hipEventRecord(h->ev[4], st));
    PJ_HIP(hipEventSynchronize(h->ev[4]));
    float ms = 0.f;
    pj_smj_timings& t = h->last;
    if (hipEventElapsedTime(&ms, h->ev[0], h->ev[1]) == hipSuccess)
        t.extract_ms = ms;
    if (hipEventElapsedTime(&ms, h->ev[1], h->ev[2]) == hipSuccess)
        t.sort_ms = ms;
    if (hipEventElapsedTime(&ms, h->ev[2], h->ev[4]) == hipSuccess)
        t.merge_ms = ms;
    if (hipEventElapsedTime(&ms, h->ev[0], h->ev[4]) == hipSuccess)
        t.total_ms = ms;
    t.build_key_bits = rbits;
    t.probe_key_bits = sbits;
    return PJ_OK;
}

int pj_smj_run_host(pj_smj_t h, const void* build, uint64_t nbuild, const void* probe, uint64_t nprobe, void* out,
                    uint64_t out_cap, uint64_t* out_n)
{
    if (!h || !out_n || (nbuild && !build) || (nprobe && !probe) || (out_cap && !out))
        return fail(PJ_ERR_INVALID, "null argument");
    *out_n = 0;
    std::lock_guard<std::mutex> g(h->host_mu);
    int boff[PJ_MAX_COLS], poff[PJ_MAX_COLS], bsize = 0, psize = 0;
    int rc;
    if ((rc = schema_offsets(h->desc_build, boff, &bsize)) || (rc = schema_offsets(h->desc_probe, poff, &psize)))
        return rc;
    const uint64_t ow = h->lay.out_width;
    if ((rc = h->host_b.ensure(std::max<uint64_t>(nbuild * (uint64_t)bsize, 16))) ||
        (rc = h->host_p.ensure(std::max<uint64_t>(nprobe * (uint64_t)psize, 16))) ||
        (rc = h->host_o.ensure(std::max<uint64_t>(out_cap * ow, 16))))
        return rc;
    if (nbuild)
        PJ_HIP(hipMemcpy(h->host_b.p, build, nbuild * (uint64_t)bsize, hipMemcpyHostToDevice));
    if (nprobe)
        PJ_HIP(hipMemcpy(h->host_p.p, probe, nprobe * (uint64_t)psize, hipMemcpyHostToDevice));
    if ((rc = pj_smj_run_device(h, h->host_b.p, nbuild, h->host_p.p, nprobe, h->host_o.p, out_cap, out_n, nullptr)))
        return rc;
    if (*out_n)
        PJ_HIP(hipMemcpy(out, h->host_o.p, *out_n * ow, hipMemcpyDeviceToHost));
    return PJ_OK;
}

int pj_smj_last_timings(pj_smj_t h, pj_smj_timings* t)
{
    if (!h || !t)
        return fail(PJ_ERR_INVALID, "null argument");
    *t = h->last;
    return PJ_OK;
}

int pj_smj_destroy(pj_smj_t h)
{
    if (!h)
        return PJ_OK;
    if (h->ev_ok)
        for (int i = 0; i < 6; ++i)
            (void)hipEventDestroy(h->ev[i]);
    if (h->host)
        (void)hipHostFree(h->host);
    delete h;
    return PJ_OK;
}

}  // extern "C"
