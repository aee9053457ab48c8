// pj_measure.hip -- measurement kernels: the rates HBM and the memory
// system give the access patterns the join is built from (streaming read,
// write and copy; random 16-byte gathers and scatters; an LDS histogram over
// mixed 64-bit keys; device atomics), so that a kernel's time can be set
// against what the machine does for the same pattern. No counterpart in the
// reference, which has no device.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pj_device.h"
#include "pj_internal.h"

namespace pj {
namespace {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
constexpr uint32_t kThreads = 256;

__global__ __launch_bounds__(kThreads) void k_read(const u32x4* __restrict__ in, uint64_t n, uint32_t* __restrict__ sink)
{
    uint32_t acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kThreads) {
        const u32x4 v = __builtin_nontemporal_load(in + i);
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    sink[(uint64_t)blockIdx.x * kThreads + threadIdx.x] = acc;
}

__global__ __launch_bounds__(kThreads) void k_write(u32x4* __restrict__ out, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kThreads) {
        const u32x4 v = {(uint32_t)i, (uint32_t)(i >> 32), 1u, 2u};
        __builtin_nontemporal_store(v, out + i);
    }
}

__global__ __launch_bounds__(kThreads) void k_copy(const u32x4* __restrict__ in, u32x4* __restrict__ out, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kThreads)
        __builtin_nontemporal_store(__builtin_nontemporal_load(in + i), out + i);
}

// out[i] = in[idx[i]]: 16-byte records read at random.
__global__ __launch_bounds__(kThreads) void k_gather(const u32x4* __restrict__ in, const uint32_t* __restrict__ idx,
                                                     u32x4* __restrict__ out, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kThreads)
        __builtin_nontemporal_store(in[idx[i]], out + i);
}

// out[idx[i]] = in[i]: 16-byte records written at random.
__global__ __launch_bounds__(kThreads) void k_scatter(const u32x4* __restrict__ in, const uint32_t* __restrict__ idx,
                                                      u32x4* __restrict__ out, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kThreads)
        out[idx[i]] = __builtin_nontemporal_load(in + i);
}

// A histogram of 2^bits bins over the top bits of the mixed key of every
// 16-byte record (its first 8 bytes), in LDS; each block writes its bins.
// The counting half of a partition pass, with nothing else.
__global__ __launch_bounds__(kThreads) void k_histogram(const u32x4* __restrict__ in, uint64_t n, uint32_t bits,
                                                        uint32_t* __restrict__ hist)
{
    extern __shared__ uint32_t cnt[];
    const uint32_t F = 1u << bits;
    for (uint32_t b = threadIdx.x; b < F; b += kThreads)
        cnt[b] = 0;
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kThreads) {
        const u32x4 v = __builtin_nontemporal_load(in + i);
        atomicAdd(&cnt[radix_part(key_mix(((uint64_t)v.y << 32) | v.x), bits)], 1u);
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < F; b += kThreads)
        hist[(uint64_t)blockIdx.x * F + b] = cnt[b];
}

// n random 16-byte reads from a table of 2^bits records at the front of
// in, at hashed positions: the lookup rate of a table of that size, which
// the caches hold in part or whole.
__global__ __launch_bounds__(kThreads) void k_lookup(const u32x4* __restrict__ in, uint64_t n, uint32_t bits,
                                                     uint32_t* __restrict__ sink)
{
    uint32_t acc = 0;
    const uint64_t mask = (1ull << bits) - 1;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kThreads) {
        const u32x4 v = in[key_mix(i) & mask];
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    sink[(uint64_t)blockIdx.x * kThreads + threadIdx.x] = acc;
}

// The shape of a probe against one table with no partitioning: stream n
// records in, look each up at random in a table of 2^bits records at the
// front of in, stream a 16-byte result out.
__global__ __launch_bounds__(kThreads) void k_probe_stream(const u32x4* __restrict__ in, uint64_t n, uint32_t bits,
                                                           u32x4* __restrict__ out)
{
    const uint64_t mask = (1ull << bits) - 1;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kThreads) {
        const u32x4 s = __builtin_nontemporal_load(in + i);
        const u32x4 t = in[key_mix(((uint64_t)s.y << 32) | s.x) & mask];
        const u32x4 o = {s.x, s.y, t.z, s.z};
        __builtin_nontemporal_store(o, out + i);
    }
}

__global__ __launch_bounds__(kThreads) void k_zero(uint32_t* __restrict__ a, uint32_t n)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n)
        a[i] = 0;
}

// k atomics a lane on `addresses` counters (a power of two), each result
// used, spread by a hash of (lane, step).
__global__ __launch_bounds__(kThreads) void k_atomic(uint32_t* __restrict__ ctr, uint32_t addresses, uint32_t k,
                                                     uint32_t* __restrict__ sink)
{
    const uint32_t t = blockIdx.x * kThreads + threadIdx.x;
    uint32_t acc = 0;
    for (uint32_t s = 0; s < k; ++s) {
        uint32_t a = t * 0x9e3779b9u + s * 0x85ebca6bu;
        a ^= a >> 15;
        a *= 0x2c1b3c6du;
        a ^= a >> 12;
        acc += atomicAdd(&ctr[a & (addresses - 1)], 1u);
    }
    sink[t] = acc;
}

}  // namespace
}  // namespace pj

using namespace pj;

extern "C" {

int pj_measure(uint32_t kind, void* d_a, void* d_b, const uint32_t* d_idx, uint64_t n, uint32_t param, uint32_t reps,
               pj_measure_result* res, void* stream)
{
    if (!res || reps == 0)
        return fail(PJ_ERR_INVALID, "null result or no repetitions");
    *res = pj_measure_result{};
    hipStream_t st = (hipStream_t)stream;
    const u32x4* a = (const u32x4*)d_a;
    u32x4* b = (u32x4*)d_b;
    uint32_t grid = 0;
    size_t lds = 0;
    uint64_t bytes = 0, ops = 0;
    switch (kind) {
    case PJ_MEASURE_READ:
    case PJ_MEASURE_WRITE:
    case PJ_MEASURE_COPY:
        grid = param ? param : 8192;
        if (!d_b || (kind != PJ_MEASURE_WRITE && !d_a) || n == 0)
            return fail(PJ_ERR_INVALID, "streaming measures need their buffers and n > 0");
        if (kind == PJ_MEASURE_READ && (uint64_t)grid * kThreads * 4 > n * 16)
            return fail(PJ_ERR_INVALID, "the read measure's sink (grid x 256 words in d_b) must fit in n records");
        bytes = n * 16 * (kind == PJ_MEASURE_COPY ? 2 : 1);
        ops = n;
        break;
    case PJ_MEASURE_GATHER:
    case PJ_MEASURE_SCATTER:
        grid = 8192;
        if (!d_a || !d_b || !d_idx || n == 0)
            return fail(PJ_ERR_INVALID, "gather and scatter need in, out, an index per record and n > 0");
        bytes = n * (16 + 4 + 16);
        ops = n;
        break;
    case PJ_MEASURE_HISTOGRAM:
        grid = 2048;
        if (!d_a || !d_b || n == 0 || param == 0 || param > 15)
            return fail(PJ_ERR_INVALID, "the histogram measure needs 1 to 15 bits and its buffers");
        if ((uint64_t)grid << param > n * 4)
            return fail(PJ_ERR_INVALID, "the histogram's bins (2048 x 2^bits words in d_b) must fit in n records");
        lds = (size_t)4 << param;
        bytes = n * 16;
        ops = n;
        break;
    case PJ_MEASURE_LOOKUP:
        grid = 8192;
        if (!d_a || !d_b || n == 0 || param == 0 || param > 40 || (1ull << param) > n ||
            (uint64_t)grid * kThreads * 4 > n * 16)
            return fail(PJ_ERR_INVALID, "the lookup measure needs a table of 2^param records within n records");
        bytes = n * 16;
        ops = n;
        break;
    case PJ_MEASURE_PROBE:
        grid = 8192;
        if (!d_a || !d_b || n == 0 || param == 0 || param > 40 || (1ull << param) > n)
            return fail(PJ_ERR_INVALID, "the probe measure needs a table of 2^param records within n records");
        bytes = n * 32;
        ops = n;
        break;
    case PJ_MEASURE_ATOMIC:
        grid = 2048;
        if (!d_a || !d_b || param == 0 || (param & (param - 1)) || param > n * 4 ||
            (uint64_t)grid * kThreads * 4 > n * 16)
            return fail(PJ_ERR_INVALID, "the atomic measure needs a power-of-two count of counters that fits in n records");
        ops = (uint64_t)grid * kThreads * 64;
        break;
    default:
        return fail(PJ_ERR_INVALID, "unknown measure %u", kind);
    }
    if (kind == PJ_MEASURE_HISTOGRAM && lds > 65536)
        PJ_HIP(hipFuncSetAttribute((const void*)k_histogram, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipEvent_t e0, e1;
    PJ_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(PJ_ERR_DEVICE, "hipEventCreate failed");
    }
    float best = 1e30f;
    int rc = PJ_OK;
    for (uint32_t r = 0; r < reps + 1 && rc == PJ_OK; ++r) {   // the first launch is a warm-up
        if (kind == PJ_MEASURE_ATOMIC) {   // the counters start at zero: a fill kernel, not a memset node
            hipLaunchKernelGGL(k_zero, dim3((param + kThreads - 1) / kThreads), dim3(kThreads), 0, st, (uint32_t*)d_a,
                               param);
        }
        (void)hipEventRecord(e0, st);
        switch (kind) {
        case PJ_MEASURE_READ:
            hipLaunchKernelGGL(k_read, dim3(grid), dim3(kThreads), 0, st, a, n, (uint32_t*)d_b);
            break;
        case PJ_MEASURE_WRITE:
            hipLaunchKernelGGL(k_write, dim3(grid), dim3(kThreads), 0, st, b, n);
            break;
        case PJ_MEASURE_COPY:
            hipLaunchKernelGGL(k_copy, dim3(grid), dim3(kThreads), 0, st, a, b, n);
            break;
        case PJ_MEASURE_GATHER:
            hipLaunchKernelGGL(k_gather, dim3(grid), dim3(kThreads), 0, st, a, d_idx, b, n);
            break;
        case PJ_MEASURE_SCATTER:
            hipLaunchKernelGGL(k_scatter, dim3(grid), dim3(kThreads), 0, st, a, d_idx, b, n);
            break;
        case PJ_MEASURE_HISTOGRAM:
            hipLaunchKernelGGL(k_histogram, dim3(grid), dim3(kThreads), lds, st, a, n, param, (uint32_t*)d_b);
            break;
        case PJ_MEASURE_LOOKUP:
            hipLaunchKernelGGL(k_lookup, dim3(grid), dim3(kThreads), 0, st, a, n, param, (uint32_t*)d_b);
            break;
        case PJ_MEASURE_PROBE:
            hipLaunchKernelGGL(k_probe_stream, dim3(grid), dim3(kThreads), 0, st, a, n, param, b);
            break;
        case PJ_MEASURE_ATOMIC:
            hipLaunchKernelGGL(k_atomic, dim3(grid), dim3(kThreads), 0, st, (uint32_t*)d_a, param, 64u,
                               (uint32_t*)d_b);
            break;
        }
        (void)hipEventRecord(e1, st);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess)
            e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess)
            e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess)
            rc = fail(PJ_ERR_DEVICE, "measure %u: %s", kind, hipGetErrorString(e));
        else if (r > 0)
            best = std::min(best, ms);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc)
        return rc;
    res->ms = best;
    res->bytes = bytes;
    res->ops = ops;
    res->gbps = bytes ? (double)bytes / (best * 1e-3) / 1e9 : 0.0;
    res->gops = (double)ops / (best * 1e-3) / 1e9;
    return PJ_OK;
}

}  // extern "C"
