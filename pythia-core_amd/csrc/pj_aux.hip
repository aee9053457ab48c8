// pj_aux.hip -- PartitionOp's exactrange partitioner, the order-independent
// output checksum, and the workload generators, on gfx950.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "pj_internal.h"
#include "pj_scan.h"

namespace pj {
namespace {

// -------------------------------------------------------------- checksum
// Same fold as the checksum the tests compute on the host: tuple bytes, 8 at
// a time little-endian (last chunk zero-padded), through the splitmix64
// finaliser; the relation's checksum is the sum mod 2^64.
__device__ __forceinline__ uint64_t tuple_mix(const uint32_t* t, uint32_t words, uint32_t width)
{
    uint64_t h = 0x9e3779b97f4a7c15ull ^ width;
    for (uint32_t w = 0; w < words; w += 2) {
        uint64_t c = t[w];
        if (w + 1 < words)
            c |= (uint64_t)t[w + 1] << 32;
        h = pj_fin64(h ^ c) + 0x9e3779b97f4a7c15ull;
    }
    return pj_fin64(h);
}

constexpr int kSumThreads = 256;

__global__ __launch_bounds__(kSumThreads) void k_checksum(const uint32_t* __restrict__ t, uint64_t n, uint32_t words,
                                                          uint32_t width, unsigned long long* __restrict__ sum,
                                                          uint32_t* __restrict__ x32)
{
    __shared__ unsigned long long ssum[kSumThreads / 64];
    __shared__ uint32_t sx[kSumThreads / 64];
    uint64_t s = 0;
    uint32_t x = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * kSumThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kSumThreads) {
        const uint32_t* r = t + i * words;
        s += tuple_mix(r, words, width);
        for (uint32_t w = 0; w < words; ++w)
            x ^= r[w];
    }
    for (int d = 32; d >= 1; d >>= 1) {
        s += __shfl_xor(s, d, 64);
        x ^= __shfl_xor(x, d, 64);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        ssum[wid] = s;
        sx[wid] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long a = 0;
        uint32_t b = 0;
        for (int w = 0; w < kSumThreads / 64; ++w) {
            a += ssum[w];
            b ^= sx[w];
        }
        atomicAdd(sum, a);
        atomicXor(x32, b);
    }
}

// ------------------------------------------------------------ generators
__global__ void k_gen_pkfk(pj_pkfk_params p, uint64_t* __restrict__ build, uint64_t* __restrict__ probe)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < p.n_build + p.n_probe; i += stride) {
        uint64_t k, v;
        if (i < p.n_build) {
            pj_pkfk_build(&p, i, &k, &v);
            build[2 * i] = k;
            build[2 * i + 1] = v;
        } else {
            const uint64_t j = i - p.n_build;
            pj_pkfk_probe(&p, j, &k, &v);
            probe[2 * j] = k;
            probe[2 * j + 1] = v;
        }
    }
}

// Batch join tuples: (int64 key, int32 payload), 12 bytes packed.
__global__ void k_gen_batch(pj_pkfk_params p, uint32_t* __restrict__ build, uint32_t* __restrict__ probe)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < p.n_build + p.n_probe; i += stride) {
        uint64_t k, v;
        uint32_t* t;
        if (i < p.n_build) {
            pj_pkfk_build(&p, i, &k, &v);
            t = build + 3 * i;
        } else {
            pj_pkfk_probe(&p, i - p.n_build, &k, &v);
            t = probe + 3 * (i - p.n_build);
        }
        t[0] = (uint32_t)k;
        t[1] = (uint32_t)(k >> 32);
        t[2] = (uint32_t)v;
    }
}

struct PairTuple {   // (int32 a, int32 b, int64 payload), 16 bytes packed
    uint32_t a, b;
    uint64_t pay;
};

__global__ void k_gen_pairs_zipf(uint64_t seed, uint64_t n_build, uint64_t n_probe, const double* __restrict__ cdf,
                                 PairTuple* __restrict__ build, PairTuple* __restrict__ probe)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_build + n_probe; i += stride) {
        PairTuple t;
        if (i < n_build) {
            pj_pair_build(seed, i, &t.a, &t.b, &t.pay);
            build[i] = t;
        } else {
            pj_pair_zipf_probe(cdf, n_build, seed, i - n_build, &t.a, &t.b, &t.pay);
            probe[i - n_build] = t;
        }
    }
}

__global__ void k_gen_clustered(const pj_cluster_table* __restrict__ t, uint64_t n_probe,
                                const double* __restrict__ cdf, uint64_t* __restrict__ build,
                                uint64_t* __restrict__ probe)
{
    __shared__ pj_cluster_table st;   // 24.6 KiB: the table, once per block
    for (uint32_t w = threadIdx.x; w < sizeof(pj_cluster_table) / 8; w += blockDim.x)
        reinterpret_cast<uint64_t*>(&st)[w] = reinterpret_cast<const uint64_t*>(t)[w];
    __syncthreads();
    const uint64_t nb = st.n_keys;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nb + n_probe; i += stride) {
        uint64_t k, v;
        if (i < nb) {
            pj_clustered_build(&st, i, &k, &v);
            build[2 * i] = k;
            build[2 * i + 1] = v;
        } else {
            const uint64_t j = i - nb;
            pj_clustered_probe_with_rank(&st, pj_zipf_rank(cdf, nb, st.seed, j), j, &k, &v);
            probe[2 * j] = k;
            probe[2 * j + 1] = v;
        }
    }
}

template <class F>
void host_parallel(uint64_t n, F f)
{
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([=] {
            const uint64_t b = n * t / nt, e = n * (t + 1) / nt;
            for (uint64_t i = b; i < e; ++i)
                f(i);
        });
    for (auto& x : th)
        x.join();
}

// ------------------------------------------------------ exactrange scatter
// hash.h:462-466 (ExactRangeValueHasher::hash): (v - min) / bucketrange.
struct RangeFn {
    int64_t min, bucketrange;
    uint32_t off_words, is_int;
    __device__ __forceinline__ uint32_t operator()(const uint32_t* t) const
    {
        int64_t v;
        if (is_int)
            v = (int32_t)t[off_words];
        else
            v = (int64_t)((uint64_t)t[off_words] | ((uint64_t)t[off_words + 1] << 32));
        return (uint32_t)((int64_t)((uint64_t)v - (uint64_t)min) / bucketrange);   // wrapping difference
    }
};

constexpr int kPartThreads = 256;

__global__ __launch_bounds__(kPartThreads) void k_range_hist(const uint32_t* __restrict__ in, uint64_t n,
                                                             uint32_t words, RangeFn fn, uint32_t B, uint64_t chunk,
                                                             uint32_t nblk, uint32_t* __restrict__ hist,
                                                             uint32_t* __restrict__ bad)
{
    __shared__ uint32_t cnt[128];
    for (uint32_t p = threadIdx.x; p < B; p += kPartThreads)
        cnt[p] = 0;
    __syncthreads();
    const uint64_t beg = (uint64_t)blockIdx.x * chunk, end = beg + chunk < n ? beg + chunk : n;
    for (uint64_t i = beg + threadIdx.x; i < end; i += kPartThreads) {
        uint32_t p = fn(in + i * words);
        if (p < B)
            atomicAdd(&cnt[p], 1u);
        else
            atomicOr(bad, 1u);
    }
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < B; p += kPartThreads)
        hist[(uint64_t)p * nblk + blockIdx.x] = cnt[p];
}

__global__ __launch_bounds__(kPartThreads) void k_range_scatter(const uint32_t* __restrict__ in, uint64_t n,
                                                                uint32_t words, RangeFn fn, uint32_t B,
                                                                uint64_t chunk, uint32_t nblk,
                                                                const uint32_t* __restrict__ offs,
                                                                uint32_t* __restrict__ out)
{
    __shared__ uint32_t cur[128];
    for (uint32_t p = threadIdx.x; p < B; p += kPartThreads)
        cur[p] = offs[(uint64_t)p * nblk + blockIdx.x];
    __syncthreads();
    const uint64_t beg = (uint64_t)blockIdx.x * chunk, end = beg + chunk < n ? beg + chunk : n;
    for (uint64_t i = beg + threadIdx.x; i < end; i += kPartThreads) {
        const uint32_t* t = in + i * words;
        uint32_t p = fn(t);
        if (p >= B)
            continue;
        uint32_t dst = atomicAdd(&cur[p], 1u);
        uint32_t* o = out + (uint64_t)dst * words;
        for (uint32_t w = 0; w < words; ++w)
            o[w] = t[w];
    }
}

}  // namespace
}  // namespace pj

using namespace pj;

extern "C" {

int pj_checksum_device(const void* d_tuples, uint64_t n, uint32_t width, uint64_t* sum, uint32_t* xor32, void* stream)
{
    if (!sum || !xor32 || (n && !d_tuples))
        return fail(PJ_ERR_INVALID, "null argument");
    if (width == 0 || width % 4)
        return fail(PJ_ERR_UNSUPPORTED, "checksum needs a tuple width that is a multiple of 4");
    hipStream_t st = (hipStream_t)stream;
    void* acc = nullptr;
    PJ_HIP(hipMalloc(&acc, 16));
    int rc = PJ_OK;
    unsigned long long s = 0;
    uint32_t x = 0;
    do {
        if (hipMemsetAsync(acc, 0, 16, st) != hipSuccess) {
            rc = fail(PJ_ERR_DEVICE, "hipMemsetAsync failed");
            break;
        }
        if (n) {
            uint64_t blocks = std::min<uint64_t>((n + kSumThreads - 1) / kSumThreads, 4096);
            hipLaunchKernelGGL(k_checksum, dim3((uint32_t)blocks), dim3(kSumThreads), 0, st,
                               (const uint32_t*)d_tuples, n, width / 4, width, (unsigned long long*)acc,
                               (uint32_t*)((char*)acc + 8));
            if (hipGetLastError() != hipSuccess) {
                rc = fail(PJ_ERR_DEVICE, "checksum launch failed");
                break;
            }
        }
        if (hipMemcpyAsync(&s, acc, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(&x, (char*)acc + 8, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            rc = fail(PJ_ERR_DEVICE, "checksum readback failed");
            break;
        }
    } while (0);
    (void)hipFree(acc);
    *sum = s;
    *xor32 = x;
    return rc;
}

int pj_gen_pkfk(uint64_t seed, uint64_t domain, uint64_t n_build, uint64_t n_probe, uint64_t n_miss, void* build_dst,
                void* probe_dst, int on_device, void* stream)
{
    if ((n_build && !build_dst) || (n_probe && !probe_dst))
        return fail(PJ_ERR_INVALID, "null argument");
    if (n_probe && n_build == 0 && n_miss < n_probe)
        return fail(PJ_ERR_INVALID, "probe tuples need build keys unless every probe misses");
    if (n_miss > n_probe || domain < n_build + n_miss)
        return fail(PJ_ERR_INVALID, "domain must hold n_build + n_miss distinct keys");
    pj_pkfk_params p = {seed, domain, n_build, n_probe, n_miss};
    if (on_device) {
        hipStream_t st = (hipStream_t)stream;
        if (n_build + n_probe) {
            hipLaunchKernelGGL(k_gen_pkfk, dim3(4096), dim3(256), 0, st, p, (uint64_t*)build_dst, (uint64_t*)probe_dst);
            PJ_HIP(hipGetLastError());
        }
        PJ_HIP(hipStreamSynchronize(st));
        return PJ_OK;
    }
    uint64_t* b = (uint64_t*)build_dst;
    uint64_t* q = (uint64_t*)probe_dst;
    host_parallel(n_build + n_probe, [&](uint64_t i) {
        uint64_t k, v;
        if (i < n_build) {
            pj_pkfk_build(&p, i, &k, &v);
            b[2 * i] = k;
            b[2 * i + 1] = v;
        } else {
            pj_pkfk_probe(&p, i - n_build, &k, &v);
            q[2 * (i - n_build)] = k;
            q[2 * (i - n_build) + 1] = v;
        }
    });
    return PJ_OK;
}

int pj_batch_sizes(uint64_t seed, uint32_t njoins, uint64_t nmin, uint64_t nmax, uint64_t* nbuild)
{
    if (!nbuild || nmin == 0 || nmax < nmin)
        return fail(PJ_ERR_INVALID, "need 0 < nmin <= nmax and an output array");
    const double a = std::log((double)nmin), b = std::log((double)nmax);
    for (uint32_t j = 0; j < njoins; ++j) {
        const double u = (double)(pj_rand64(seed, 7, j) >> 11) * (1.0 / 9007199254740992.0);   // [0, 1)
        uint64_t n = (uint64_t)std::floor(std::exp(a + u * (b - a)));
        nbuild[j] = std::min(nmax, std::max(nmin, n));
    }
    return PJ_OK;
}

int pj_gen_batch_join(uint64_t seed, uint32_t j, uint64_t nbuild, uint64_t nprobe, void* build_dst, void* probe_dst,
                      int on_device, void* stream)
{
    if ((nbuild && !build_dst) || (nprobe && !probe_dst) || (nprobe && !nbuild))
        return fail(PJ_ERR_INVALID, "null argument or probe tuples without build tuples");
    if (nbuild > (1ull << 40))
        return fail(PJ_ERR_INVALID, "at most 2^40 build tuples");
    pj_pkfk_params p = {pj_batch_seed(seed, j), 1ull << 40, nbuild, nprobe, 0};
    if (on_device) {
        hipStream_t st = (hipStream_t)stream;
        if (nbuild + nprobe) {
            hipLaunchKernelGGL(k_gen_batch, dim3(1024), dim3(256), 0, st, p, (uint32_t*)build_dst, (uint32_t*)probe_dst);
            PJ_HIP(hipGetLastError());
        }
        return PJ_OK;
    }
    uint32_t* b = (uint32_t*)build_dst;
    uint32_t* q = (uint32_t*)probe_dst;
    host_parallel(nbuild + nprobe, [&](uint64_t i) {
        uint64_t k, v;
        uint32_t* t;
        if (i < nbuild) {
            pj_pkfk_build(&p, i, &k, &v);
            t = b + 3 * i;
        } else {
            pj_pkfk_probe(&p, i - nbuild, &k, &v);
            t = q + 3 * (i - nbuild);
        }
        t[0] = (uint32_t)k;
        t[1] = (uint32_t)(k >> 32);
        t[2] = (uint32_t)v;
    });
    return PJ_OK;
}

int pj_zipf_cdf(uint64_t n, double z, double* cdf)
{
    // generator_zipf.cpp:91-113: cdf[r] = sum_{i<=r} (i+1)^-z / H, summed
    // in order in double precision
    if (!cdf || n == 0)
        return fail(PJ_ERR_INVALID, "null argument");
    double h = 0;
    for (uint64_t i = 0; i < n; ++i) {
        h += std::pow((double)(i + 1), -z);
        cdf[i] = h;
    }
    for (uint64_t i = 0; i < n; ++i)
        cdf[i] /= h;
    return PJ_OK;
}

int pj_gen_pairs_zipf(uint64_t seed, uint64_t n_build, uint64_t n_probe, const double* cdf, void* build_dst,
                      void* probe_dst, int on_device, void* stream)
{
    if ((n_build && !build_dst) || (n_probe && (!probe_dst || !cdf || n_build == 0)))
        return fail(PJ_ERR_INVALID, "null argument");
    if (n_build >= (1ull << 62))
        return fail(PJ_ERR_INVALID, "too many build tuples");
    if (on_device) {
        hipStream_t st = (hipStream_t)stream;
        double* dcdf = nullptr;
        if (n_probe) {
            PJ_HIP(hipMalloc(&dcdf, n_build * sizeof(double)));
            if (hipMemcpy(dcdf, cdf, n_build * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
                (void)hipFree(dcdf);
                return fail(PJ_ERR_DEVICE, "cdf upload failed");
            }
        }
        hipLaunchKernelGGL(k_gen_pairs_zipf, dim3(4096), dim3(256), 0, st, seed, n_build, n_probe, dcdf,
                           (PairTuple*)build_dst, (PairTuple*)probe_dst);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess)
            e = hipStreamSynchronize(st);
        if (dcdf)
            (void)hipFree(dcdf);
        if (e != hipSuccess)
            return fail(PJ_ERR_DEVICE, "pair generator failed: %s", hipGetErrorString(e));
        return PJ_OK;
    }
    PairTuple* b = (PairTuple*)build_dst;
    PairTuple* q = (PairTuple*)probe_dst;
    host_parallel(n_build + n_probe, [&](uint64_t i) {
        PairTuple t;
        if (i < n_build) {
            pj_pair_build(seed, i, &t.a, &t.b, &t.pay);
            b[i] = t;
        } else {
            pj_pair_zipf_probe(cdf, n_build, seed, i - n_build, &t.a, &t.b, &t.pay);
            q[i - n_build] = t;
        }
    });
    return PJ_OK;
}

uint64_t pj_cluster_table_size(void) { return sizeof(pj_cluster_table); }

int pj_clusters_make(uint64_t seed, uint32_t nclusters, uint64_t n_keys, uint32_t log2_wmin, uint32_t log2_wmax,
                     double weight_z, void* table_out)
{
    pj_cluster_table* t = (pj_cluster_table*)table_out;
    if (!t || nclusters == 0 || nclusters > PJ_MAX_CLUSTERS || log2_wmin > log2_wmax || log2_wmax > 40)
        return fail(PJ_ERR_INVALID, "bad cluster parameters");
    memset(t, 0, sizeof(*t));
    t->nclusters = nclusters;
    t->seed = seed;
    t->n_keys = n_keys;
    std::vector<uint64_t> centre(nclusters), width(nclusters), count(nclusters);
    std::vector<double> weight(nclusters);
    double wsum = 0;
    for (uint32_t c = 0; c < nclusters; ++c) {
        centre[c] = pj_rand64(seed, 10, c) >> 2;   // [0, 2^62)
        const double u = (double)(pj_rand64(seed, 11, c) >> 11) / 9007199254740992.0;
        width[c] = (uint64_t)std::floor(std::pow(2.0, log2_wmin + u * (log2_wmax - log2_wmin)));
        weight[c] = std::pow((double)(c + 1), -weight_z);
        wsum += weight[c];
    }
    uint64_t placed = 0;
    for (uint32_t c = 0; c < nclusters; ++c) {
        count[c] = std::min<uint64_t>(width[c], (uint64_t)std::llround((double)n_keys * weight[c] / wsum));
        placed += count[c];
    }
    if (placed > n_keys) {   // rounding: trim from the lightest clusters
        for (uint32_t c = nclusters; c-- > 0 && placed > n_keys;) {
            const uint64_t d = std::min<uint64_t>(count[c], placed - n_keys);
            count[c] -= d;
            placed -= d;
        }
    }
    for (uint32_t c = 0; c < nclusters && placed < n_keys; ++c) {   // remainder to clusters with room
        const uint64_t d = std::min<uint64_t>(width[c] - count[c], n_keys - placed);
        count[c] += d;
        placed += d;
    }
    if (placed != n_keys)
        return fail(PJ_ERR_INVALID, "clusters hold only %llu of %llu keys", (unsigned long long)placed,
                    (unsigned long long)n_keys);
    std::vector<uint32_t> order(nclusters);
    for (uint32_t c = 0; c < nclusters; ++c)
        order[c] = c;
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return centre[a] < centre[b]; });
    uint64_t prev_end = 0, g = 0;
    for (uint32_t k = 0; k < nclusters; ++k) {
        const uint32_t c = order[k];
        uint64_t ctr = std::max(centre[c], prev_end);   // keep clusters disjoint
        t->centre[k] = ctr;
        t->width[k] = width[c];
        t->start[k] = g;
        g += count[c];
        prev_end = ctr + width[c];
    }
    t->start[nclusters] = g;
    return PJ_OK;
}

int pj_gen_clustered(const void* table, uint64_t n_probe, const double* cdf, void* build_dst, void* probe_dst,
                     int on_device, void* stream)
{
    const pj_cluster_table* t = (const pj_cluster_table*)table;
    if (!t || (t->n_keys && !build_dst) || (n_probe && (!probe_dst || !cdf || t->n_keys == 0)))
        return fail(PJ_ERR_INVALID, "null argument");
    if (on_device) {
        hipStream_t st = (hipStream_t)stream;
        void* dt = nullptr;
        double* dcdf = nullptr;
        PJ_HIP(hipMalloc(&dt, sizeof(*t)));
        hipError_t e = hipMemcpy(dt, t, sizeof(*t), hipMemcpyHostToDevice);
        if (e == hipSuccess && n_probe) {
            e = hipMalloc(&dcdf, t->n_keys * sizeof(double));
            if (e == hipSuccess)
                e = hipMemcpy(dcdf, cdf, t->n_keys * sizeof(double), hipMemcpyHostToDevice);
        }
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_gen_clustered, dim3(4096), dim3(256), 0, st, (const pj_cluster_table*)dt, n_probe,
                               dcdf, (uint64_t*)build_dst, (uint64_t*)probe_dst);
            e = hipGetLastError();
            if (e == hipSuccess)
                e = hipStreamSynchronize(st);
        }
        (void)hipFree(dt);
        if (dcdf)
            (void)hipFree(dcdf);
        if (e != hipSuccess)
            return fail(PJ_ERR_DEVICE, "clustered generator failed: %s", hipGetErrorString(e));
        return PJ_OK;
    }
    uint64_t* b = (uint64_t*)build_dst;
    uint64_t* q = (uint64_t*)probe_dst;
    const uint64_t nb = t->n_keys;
    host_parallel(nb + n_probe, [&](uint64_t i) {
        uint64_t k, v;
        if (i < nb) {
            pj_clustered_build(t, i, &k, &v);
            b[2 * i] = k;
            b[2 * i + 1] = v;
        } else {
            const uint64_t j = i - nb;
            pj_clustered_probe_with_rank(t, pj_zipf_rank(cdf, nb, t->seed, j), j, &k, &v);
            q[2 * j] = k;
            q[2 * j + 1] = v;
        }
    });
    return PJ_OK;
}

int pj_partition_exactrange_device(const pj_schema* schema, uint32_t attr, int64_t range_min, int64_t range_max,
                                   uint32_t buckets, const void* d_in, uint64_t n, void* d_out, uint64_t* part_sizes,
                                   void* stream)
{
    if (!schema || !part_sizes || (n && (!d_in || !d_out)))
        return fail(PJ_ERR_INVALID, "null argument");
    // PartitionOp::init (operators/partition.cpp:44-105): buckets == threads
    // and fewer than MAX_THREADS (128); the column must be numeric.
    if (buckets == 0 || buckets >= 128)
        return fail(PJ_ERR_INVALID, "buckets must be in 1..127");
    if (attr >= schema->ncols || schema->ncols > PJ_MAX_COLS)
        return fail(PJ_ERR_INVALID, "attr out of range");
    uint32_t off = 0, width = 0;
    for (uint32_t c = 0; c < schema->ncols; ++c) {
        uint32_t w;
        switch (schema->type[c]) {
        case PJ_CT_INTEGER: w = 4; break;
        case PJ_CT_CHAR: w = schema->width[c]; break;
        case PJ_CT_LONG: case PJ_CT_DECIMAL: case PJ_CT_DATE: case PJ_CT_POINTER: w = 8; break;
        default: return fail(PJ_ERR_INVALID, "column %u has a bad type", c);
        }
        if (c == attr)
            off = width;
        width += w;
    }
    const uint32_t t = schema->type[attr];
    if (!(t == PJ_CT_INTEGER || t == PJ_CT_LONG || t == PJ_CT_DATE))
        return fail(PJ_ERR_INVALID, "partition attribute must be INTEGER, LONG or DATE");
    if (width % 4 || off % 4)
        return fail(PJ_ERR_UNSUPPORTED, "device partitioner needs 4-byte multiples");
    if (range_max < range_min)
        return fail(PJ_ERR_INVALID, "empty range");
    if ((uint64_t)range_max - (uint64_t)range_min > (uint64_t)INT64_MAX - buckets)
        return fail(PJ_ERR_INVALID, "range [%lld, %lld] over %u buckets overflows a LONG", (long long)range_min,
                    (long long)range_max, buckets);
    if (n >= (1ull << 32))
        return fail(PJ_ERR_UNSUPPORTED, "relations of 2^32 tuples or more");
    RangeFn fn;
    fn.min = range_min;
    fn.bucketrange = (range_max - range_min + (int64_t)buckets) / (int64_t)buckets;  // hash.h:455-460
    fn.off_words = off / 4;
    fn.is_int = t == PJ_CT_INTEGER;
    std::fill(part_sizes, part_sizes + buckets, 0);
    if (n == 0)
        return PJ_OK;
    hipStream_t st = (hipStream_t)stream;
    const uint32_t nblk = (uint32_t)std::min<uint64_t>(std::max<uint64_t>((n + 16383) / 16384, 1), 2048);
    const uint64_t chunk = (n + nblk - 1) / nblk;
    const uint64_t m = (uint64_t)buckets * nblk;
    const uint32_t nsb = (uint32_t)((m + kScanSeg - 1) / kScanSeg);
    void *hist = nullptr, *bsum = nullptr, *bad = nullptr;
    int rc = PJ_OK;
    std::vector<uint32_t> starts(buckets);
    uint32_t badv = 0;
    do {
        if (hipMalloc(&hist, m * 4) != hipSuccess || hipMalloc(&bsum, (nsb + 1) * 4) != hipSuccess ||
            hipMalloc(&bad, 4) != hipSuccess) {
            rc = fail(PJ_ERR_DEVICE, "hipMalloc failed");
            break;
        }
        if (hipMemsetAsync(bad, 0, 4, st) != hipSuccess) {
            rc = fail(PJ_ERR_DEVICE, "hipMemsetAsync failed");
            break;
        }
        hipLaunchKernelGGL(k_range_hist, dim3(nblk), dim3(kPartThreads), 0, st, (const uint32_t*)d_in, n, width / 4,
                           fn, buckets, chunk, nblk, (uint32_t*)hist, (uint32_t*)bad);
        hipLaunchKernelGGL(k_scan_reduce, dim3(nsb), dim3(kScanThreads), 0, st, (uint32_t*)hist, m, (uint32_t*)bsum);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, (uint32_t*)bsum, nsb);
        hipLaunchKernelGGL(k_scan_down, dim3(nsb), dim3(kScanThreads), 0, st, (uint32_t*)hist, m,
                           (const uint32_t*)bsum);
        hipLaunchKernelGGL(k_range_scatter, dim3(nblk), dim3(kPartThreads), 0, st, (const uint32_t*)d_in, n,
                           width / 4, fn, buckets, chunk, nblk, (const uint32_t*)hist, (uint32_t*)d_out);
        if (hipGetLastError() != hipSuccess) {
            rc = fail(PJ_ERR_DEVICE, "partition launch failed");
            break;
        }
        for (uint32_t p = 0; p < buckets; ++p)
            if (hipMemcpyAsync(&starts[p], (uint32_t*)hist + (uint64_t)p * nblk, 4, hipMemcpyDeviceToHost, st) !=
                hipSuccess) {
                rc = fail(PJ_ERR_DEVICE, "readback failed");
                break;
            }
        if (rc)
            break;
        if (hipMemcpyAsync(&badv, bad, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            rc = fail(PJ_ERR_DEVICE, "readback failed");
            break;
        }
    } while (0);
    if (hist)
        (void)hipFree(hist);
    if (bsum)
        (void)hipFree(bsum);
    if (bad)
        (void)hipFree(bad);
    if (rc)
        return rc;
    if (badv)
        return fail(PJ_ERR_INVALID, "a key lies outside [range_min, range_max]");
    uint64_t total = 0;
    for (uint32_t p = 0; p < buckets; ++p) {
        uint64_t e = p + 1 < buckets ? starts[p + 1] : n;
        part_sizes[p] = e - starts[p];
        total += part_sizes[p];
    }
    if (total != n)
        return fail(PJ_ERR_STATE, "partition sizes add up to %llu of %llu tuples", (unsigned long long)total,
                    (unsigned long long)n);
    return PJ_OK;
}

int pj_partition_exactrange(const pj_schema* schema, uint32_t attr, int64_t range_min, int64_t range_max,
                            uint32_t buckets, const void* in, uint64_t n, void* out, uint64_t* part_sizes)
{
    if (!schema || !part_sizes || (n && (!in || !out)))
        return fail(PJ_ERR_INVALID, "null argument");
    uint32_t width = 0;
    for (uint32_t c = 0; c < schema->ncols && c < PJ_MAX_COLS; ++c)
        width += schema->type[c] == PJ_CT_INTEGER ? 4 : schema->type[c] == PJ_CT_CHAR ? schema->width[c] : 8;
    void *din = nullptr, *dout = nullptr;
    const size_t bytes = std::max<size_t>(n * width, 16);
    if (hipMalloc(&din, bytes) != hipSuccess || hipMalloc(&dout, bytes) != hipSuccess) {
        if (din)
            (void)hipFree(din);
        return fail(PJ_ERR_DEVICE, "hipMalloc of %zu bytes failed", bytes);
    }
    int rc = PJ_OK;
    if (n && hipMemcpy(din, in, n * width, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(PJ_ERR_DEVICE, "upload failed");
    if (rc == PJ_OK)
        rc = pj_partition_exactrange_device(schema, attr, range_min, range_max, buckets, din, n, dout, part_sizes,
                                            nullptr);
    if (rc == PJ_OK && n && hipMemcpy(out, dout, n * width, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(PJ_ERR_DEVICE, "download failed");
    (void)hipFree(din);
    (void)hipFree(dout);
    return rc;
}

}  // extern "C"
