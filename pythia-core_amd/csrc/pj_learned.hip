// pj_learned.hip -- the learned partitioner: a two-stage piecewise-linear
// CDF model of the build keys, trained on the GPU, that places range
// boundaries for the join's partitions.
//
// There is no counterpart in the reference: HashJoinOp partitions by hash
// bits (TupleHasher, hash.h:109-197) and PartitionOp by fixed key ranges
// (ExactRangeValueHasher, partition.cpp:166-200). Here:
//   1. every 100th build key is sampled (k_sample) and the sample is sorted
//      (LSD radix sort: k_sort_hist, scan, k_sort_scatter);
//   2. the root cuts the sorted sample into segments at gaps wider than
//      2^32 (k_seg_*) and gives each segment leaves of equal key width in
//      proportion to its sample count;
//   3. each leaf is a line rank ~ a + b x, fitted by least squares over the
//      samples of a window of neighbouring leaves (k_leaf_moments,
//      k_leaf_fit: centred FP32 moments merged with Chan's update);
//   4. training steps minimise that loss plus a continuity term at inner
//      leaf edges (k_gd_forward: residuals and their sums per leaf;
//      k_gd_backward: gradient, and a Newton step on each leaf's 2x2 block);
//   5. boundaries put equal LDS cost into every build partition, where a key
//      costs 1/kDenseCap in leaves the model finds dense (every integer
//      present) and 1/kCodeCap elsewhere, and one boundary sits in every gap
//      between segments, so a build partition never spans a gap.
// Partitioning with the boundaries is exact integer work (pj_join.hip); the
// model only decides where the ranges start.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pj_device.h"
#include "pj_internal.h"
#include "pj_scan.h"

namespace pj {
namespace {

constexpr uint64_t kSampleStride = 100;    // 1% sample
constexpr uint64_t kGap = 1ull << 32;      // segment cut
constexpr uint32_t kWindow = 16;           // least-squares window, leaves either side
constexpr uint64_t kLeafSamples = 128;     // samples per leaf the root aims for (up to 4096 leaves)
constexpr float kLambda = 64.f;            // continuity weight (in samples)
constexpr uint32_t kSteps = 3;             // training steps in a join
constexpr float kEta = 0.8f;               // step size of the Newton steps
constexpr double kFill = 0.85;             // planned fill of an LDS table
constexpr double kFillMax = 0.92;          // highest fill accepted to save a probe re-read
constexpr double kDenseDensity = 0.97;     // keys per key value above which a segment counts as dense

constexpr uint32_t kMaxFine = 4096;        // build partitions the range kernels' LDS can hold
static_assert(kMaxFine <= 65536, "build partitions travel as 16-bit ids and divide by multiply-high");

constexpr uint32_t kSortThreads = 1024;
constexpr uint32_t kSortChunk = 4096;
constexpr uint32_t kFitThreads = 256;

// ---------------------------------------------------------------- sample

template <bool IDENT>
__global__ void k_sample(const uint32_t* __restrict__ in, uint64_t stride, uint64_t ns, SideLayout L,
                         uint64_t* __restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ns)
        out[i] = ord_key(rec_key(load_rec<IDENT>(in, i * stride, L)));
}

// ---------------------------------------------------------------- sort
// One pass of an LSD radix sort on an 8-bit digit: per-block digit counts,
// an exclusive scan over (digit, block), and a stable scatter in which the
// lanes of a wave that share a digit are found with eight ballots.

__global__ __launch_bounds__(kSortThreads) void k_sort_hist(const uint64_t* __restrict__ keys, uint64_t n,
                                                             uint32_t shift, uint32_t nblk, uint32_t* __restrict__ hist)
{
    __shared__ uint32_t cnt[256];
    if (threadIdx.x < 256)
        cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t beg = (uint64_t)blockIdx.x * kSortChunk;
    const uint64_t end = beg + kSortChunk < n ? beg + kSortChunk : n;
    for (uint64_t i = beg + threadIdx.x; i < end; i += kSortThreads)
        atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
    __syncthreads();
    if (threadIdx.x < 256)
        hist[(uint64_t)threadIdx.x * nblk + blockIdx.x] = cnt[threadIdx.x];
}

__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(const uint64_t* __restrict__ in,
                                                                uint64_t* __restrict__ out, uint64_t n,
                                                                uint32_t shift, uint32_t nblk,
                                                                const uint32_t* __restrict__ offs)
{
    __shared__ uint32_t base[256];
    __shared__ uint32_t wcnt[kSortThreads / 64][256];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    if (tid < 256)
        base[tid] = offs[(uint64_t)tid * nblk + blockIdx.x];
    const uint64_t beg = (uint64_t)blockIdx.x * kSortChunk;
    const uint64_t end = beg + kSortChunk < n ? beg + kSortChunk : n;
    for (uint64_t t0 = beg; t0 < end; t0 += kSortThreads) {
        for (uint32_t j = tid; j < (kSortThreads / 64) * 256; j += kSortThreads)
            (&wcnt[0][0])[j] = 0;
        __syncthreads();
        const uint64_t i = t0 + tid;
        const bool valid = i < end;
        const uint64_t k = valid ? in[i] : 0ull;
        const uint32_t d = (uint32_t)(k >> shift) & 255u;
        uint64_t same = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t set = __ballot(valid && bit);
            same &= bit ? set : ~set;
        }
        const uint32_t rank = (uint32_t)__popcll(same & below);
        if (valid && rank == 0)
            wcnt[wave][d] = (uint32_t)__popcll(same);
        __syncthreads();
        if (tid < 256) {
            uint32_t run = base[tid];
            for (uint32_t w = 0; w < kSortThreads / 64; ++w) {
                const uint32_t c = wcnt[w][tid];
                wcnt[w][tid] = run;
                run += c;
            }
            base[tid] = run;
        }
        __syncthreads();
        if (valid)
            out[wcnt[wave][d] + rank] = k;
        __syncthreads();
    }
}

// ---------------------------------------------------------------- root

__global__ void k_seg_flags(const uint64_t* __restrict__ keys, uint64_t n, uint32_t* __restrict__ flags)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        flags[i] = (i == 0 || keys[i] - keys[i - 1] > kGap) ? 1u : 0u;
}

// scanned = exclusive prefix of the flags; segment s starts at sample i.
// Writes each segment's first and last key and its first sample.
__global__ void k_seg_starts(const uint64_t* __restrict__ keys, uint64_t n, const uint32_t* __restrict__ scanned,
                             uint32_t maxseg, uint64_t* __restrict__ seg_lo, uint64_t* __restrict__ seg_hi,
                             uint32_t* __restrict__ seg_start)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const bool first = i == 0 || keys[i] - keys[i - 1] > kGap;
    const uint32_t s = scanned[i] + (first ? 1u : 0u) - 1u;   // segment of sample i
    if (s >= maxseg)
        return;
    if (first) {
        seg_lo[s] = keys[i];
        seg_start[s] = (uint32_t)i;
    }
    if (i == n - 1 || keys[i + 1] - keys[i] > kGap)
        seg_hi[s] = keys[i];
}

__device__ __forceinline__ uint32_t seg_of(const pj_model* __restrict__ M, uint64_t u)
{
    uint32_t lo = 0, hi = M->nseg;   // last s with seg_lo[s] <= u, else 0
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (M->seg_lo[mid] <= u)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t leaf_of(const pj_model* __restrict__ M, uint64_t u)
{
    const uint32_t s = seg_of(M, u);
    const uint32_t l0 = M->seg_leaf0[s], nl = M->seg_leaf0[s + 1] - l0;
    if (u < M->seg_lo[s])
        return l0;
    const uint64_t q = (u - M->seg_lo[s]) / M->seg_lw[s];
    return l0 + (uint32_t)(q < nl - 1 ? q : nl - 1);
}

// x of ordered key u in its own leaf j, in [0, 1) (up to FP32 rounding).
__device__ __forceinline__ float leaf_x(const pj_model* __restrict__ M, uint32_t j, uint64_t u)
{
    const uint32_t s = M->leaf_seg[j];
    const uint64_t lo = M->seg_lo[s] + (uint64_t)(j - M->seg_leaf0[s]) * M->seg_lw[s];
    return (float)(u - lo) / (float)M->seg_lw[s];
}

// leaf_start[l] = first sample whose leaf is >= l; leaf_start[nleaves] = n.
__global__ void k_leaf_starts(pj_model* __restrict__ M, const uint64_t* __restrict__ keys, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int32_t j = (int32_t)leaf_of(M, keys[i]);
    const int32_t jp = i ? (int32_t)leaf_of(M, keys[i - 1]) : -1;
    for (int32_t l = jp + 1; l <= j; ++l)
        M->leaf_start[l] = (uint32_t)i;
    if (i == n - 1)
        for (uint32_t l = (uint32_t)j + 1; l <= M->nleaves; ++l)
            M->leaf_start[l] = (uint32_t)n;
}

// ---------------------------------------------------------------- fit

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        v += __shfl_xor(v, d, 64);
    return v;
}

// Sum over the block (kFitThreads threads); every thread gets the total.
__device__ __forceinline__ float block_sum(float v, float* red)
{
    v = wave_sum(v);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0)
        red[wave] = v;
    __syncthreads();
    float t = 0.f;
#pragma unroll
    for (uint32_t w = 0; w < kFitThreads / 64; ++w)
        t += red[w];
    return t;
}

struct Moments {   // centred moments of (x, y) over a set of samples
    float n, mx, my, cxx, cxy;
};

// Per leaf, in its own frame (x in [0,1), y = rank - leaf_start): N, means,
// centred sums of squares and products (two passes for accuracy). Also
// stores every sample's x (own frame) in xs for the training passes.
__global__ __launch_bounds__(kFitThreads) void k_leaf_moments(const pj_model* __restrict__ M,
                                                               const uint64_t* __restrict__ keys,
                                                               Moments* __restrict__ mom, float* __restrict__ xs)
{
    __shared__ float red[kFitThreads / 64];
    const uint32_t j = blockIdx.x;
    const uint32_t s0 = M->leaf_start[j], s1 = M->leaf_start[j + 1];
    float sx = 0.f, sy = 0.f;
    for (uint32_t i = s0 + threadIdx.x; i < s1; i += kFitThreads) {
        const float x = leaf_x(M, j, keys[i]);
        xs[i] = x;
        sx += x;
        sy += (float)(i - s0);
    }
    const float n = (float)(s1 - s0);
    sx = block_sum(sx, red);
    sy = block_sum(sy, red);
    const float mx = n > 0.f ? sx / n : 0.f, my = n > 0.f ? sy / n : 0.f;
    float cxx = 0.f, cxy = 0.f;
    for (uint32_t i = s0 + threadIdx.x; i < s1; i += kFitThreads) {
        const float dx = xs[i] - mx, dy = (float)(i - s0) - my;
        cxx += dx * dx;
        cxy += dx * dy;
    }
    cxx = block_sum(cxx, red);
    cxy = block_sum(cxy, red);
    if (threadIdx.x == 0)
        mom[j] = Moments{n, mx, my, cxx, cxy};
}

struct LeafStat {   // window statistics in the leaf's frame, for the Newton steps
    float n, sx, sxx;
};

__device__ __forceinline__ void window_of(const pj_model* __restrict__ M, uint32_t j, uint32_t* w0, uint32_t* w1)
{
    const uint32_t s = M->leaf_seg[j];
    const uint32_t l0 = M->seg_leaf0[s], l1 = M->seg_leaf0[s + 1];
    *w0 = j >= l0 + kWindow ? j - kWindow : l0;
    *w1 = j + kWindow + 1 <= l1 ? j + kWindow + 1 : l1;
}

// Least-squares line of leaf j over the samples of its window, from the
// window's moments shifted into leaf j's frame (x + (m - j), y + start_m -
// start_j) and merged pairwise (Chan, Golub, LeVeque 1979).
__global__ void k_leaf_fit(const pj_model* __restrict__ M, const Moments* __restrict__ mom, float2* __restrict__ ab,
                           LeafStat* __restrict__ stat)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M->nleaves)
        return;
    uint32_t w0, w1;
    window_of(M, j, &w0, &w1);
    float N = 0.f, MX = 0.f, MY = 0.f, CXX = 0.f, CXY = 0.f;
    for (uint32_t m = w0; m < w1; ++m) {
        const Moments q = mom[m];
        if (q.n == 0.f)
            continue;
        const float qx = q.mx + (float)((int32_t)m - (int32_t)j);
        const float qy = q.my + (float)((int64_t)M->leaf_start[m] - (int64_t)M->leaf_start[j]);
        const float nn = N + q.n;
        const float dx = qx - MX, dy = qy - MY;
        const float f = N * q.n / nn;
        CXX += q.cxx + dx * dx * f;
        CXY += q.cxy + dx * dy * f;
        MX += dx * q.n / nn;
        MY += dy * q.n / nn;
        N = nn;
    }
    float b = 0.f, a = MY;
    if (N >= 2.f && CXX > 1e-12f * N) {
        b = CXY / CXX;
        if (b < 0.f)
            b = 0.f;
        a = MY - b * MX;
    }
    if (N == 0.f)
        a = 0.f;
    ab[j] = make_float2(a, b);
    stat[j] = LeafStat{N, N * MX, CXX + N * MX * MX};
}

// ---------------------------------------------------------------- training
// Loss L = (1/n) [ sum_j sum_{i in window(j)} (a_j + b_j x_ij - y_ij)^2
//                + lambda * sum_{inner edges j|j+1} delta_j^2 ],
// delta_j = (start_j + a_j + b_j) - (start_{j+1} + a_{j+1}).

struct Resid {   // sums of r, r x and r^2 over leaf j's window
    float r, rx, rr;
};

// Forward: residuals of leaf j's line over its window, reduced per leaf
// (xs: every sample's x in its own leaf's frame).
__global__ __launch_bounds__(kFitThreads) void k_gd_forward(const pj_model* __restrict__ M,
                                                             const float* __restrict__ xs,
                                                             const float2* __restrict__ ab, Resid* __restrict__ res)
{
    __shared__ float red[kFitThreads / 64];
    const uint32_t j = blockIdx.x;
    uint32_t w0, w1;
    window_of(M, j, &w0, &w1);
    const float2 p = ab[j];
    const uint32_t sj = M->leaf_start[j];
    float r1 = 0.f, rx = 0.f, rr = 0.f;
    for (uint32_t m = w0; m < w1; ++m) {
        const float shift = (float)((int32_t)m - (int32_t)j);
        const uint32_t s0 = M->leaf_start[m], s1 = M->leaf_start[m + 1];
        for (uint32_t i = s0 + threadIdx.x; i < s1; i += kFitThreads) {
            const float x = xs[i] + shift;
            const float y = (float)((int64_t)i - (int64_t)sj);
            const float r = p.x + p.y * x - y;
            r1 += r;
            rx += r * x;
            rr += r * r;
        }
    }
    r1 = block_sum(r1, red);
    rx = block_sum(rx, red);
    rr = block_sum(rr, red);
    if (threadIdx.x == 0)
        res[j] = Resid{r1, rx, rr};
}

// Backward: dL/da_j, dL/db_j from the residual sums and the continuity
// terms; then (eta > 0) a Newton step on leaf j's 2x2 block of the Hessian,
// written to ab_out. Adds leaf j's share of the loss to *loss.
__global__ void k_gd_backward(const pj_model* __restrict__ M, const float2* __restrict__ ab_in,
                              const Resid* __restrict__ res, const LeafStat* __restrict__ stat, float lambda,
                              float eta, float2* __restrict__ ab_out, float* __restrict__ grad,
                              double* __restrict__ loss)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t L = M->nleaves;
    if (j >= L)
        return;
    const float inv_n = 1.f / (float)M->n;
    const uint32_t s = M->leaf_seg[j];
    const bool hr = j + 1 < M->seg_leaf0[s + 1], hl = j > M->seg_leaf0[s];
    const float2 p = ab_in[j];
    float dj = 0.f, dl = 0.f;
    if (hr) {
        const float2 q = ab_in[j + 1];
        dj = (float)((int64_t)M->leaf_start[j] - (int64_t)M->leaf_start[j + 1]) + p.x + p.y - q.x;
    }
    if (hl) {
        const float2 q = ab_in[j - 1];
        dl = (float)((int64_t)M->leaf_start[j - 1] - (int64_t)M->leaf_start[j]) + q.x + q.y - p.x;
    }
    const Resid r = res[j];
    const float ga = inv_n * (2.f * r.r + 2.f * lambda * (dj - dl));
    const float gb = inv_n * (2.f * r.rx + 2.f * lambda * dj);
    if (grad) {
        grad[j] = ga;
        grad[L + j] = gb;
    }
    atomicAdd(loss, (double)(inv_n * (r.rr + lambda * dj * dj)));
    if (eta > 0.f) {
        const LeafStat st = stat[j];
        const float haa = 2.f * inv_n * (st.n + lambda * ((hr ? 1.f : 0.f) + (hl ? 1.f : 0.f)));
        const float hab = 2.f * inv_n * (st.sx + (hr ? lambda : 0.f));
        const float hbb = 2.f * inv_n * (st.sxx + (hr ? lambda : 0.f));
        const float det = haa * hbb - hab * hab;
        float2 o = p;
        if (det > 1e-30f && haa > 0.f) {
            o.x = p.x - eta * (hbb * ga - hab * gb) / det;
            o.y = p.y - eta * (haa * gb - hab * ga) / det;
        } else if (haa > 0.f) {
            o.x = p.x - eta * ga / haa;
        }
        if (o.y < 0.f)
            o.y = 0.f;
        ab_out[j] = o;
    }
}

// ---------------------------------------------------------------- host

constexpr uint32_t kMaxSegs = PJ_MODEL_MAX_SEGS;
constexpr uint32_t kLeaves = PJ_MODEL_LEAVES;

// Device scratch layout of a fit (offsets into ctx buffers).
struct FitBufs {
    uint64_t* sorted;   // n
    uint64_t* alt;      // n
    uint32_t* scan;     // n (flags, then their prefix), and sort histograms
    uint32_t* bsum;
    pj_model* dm;
    float2* ab[2];
    Moments* mom;
    LeafStat* stat;
    Resid* res;
    double* loss;
    float* grad;
    uint64_t* seg_lo;
    uint64_t* seg_hi;
    uint32_t* seg_start;
    float* xs;          // n: each sample's x in its own leaf
};

int scan_u32(uint32_t* a, uint64_t m, uint32_t* bsum, hipStream_t st)
{
    const uint32_t nsb = (uint32_t)((m + kScanSeg - 1) / kScanSeg);
    hipLaunchKernelGGL(k_scan_reduce, dim3(nsb), dim3(kScanThreads), 0, st, a, m, bsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, bsum, nsb);
    hipLaunchKernelGGL(k_scan_down, dim3(nsb), dim3(kScanThreads), 0, st, a, m, (const uint32_t*)bsum);
    PJ_HIP(hipGetLastError());
    return PJ_OK;
}

uint32_t scan_blocks(uint64_t m) { return (uint32_t)((m + kScanSeg - 1) / kScanSeg); }

// Sorts keys (n of them) using alt as the ping-pong buffer; the result is in
// keys after the eight passes.
int sort_keys(uint64_t* keys, uint64_t* alt, uint64_t n, uint32_t* hist, uint32_t* bsum, hipStream_t st)
{
    if (n < 2)
        return PJ_OK;
    const uint32_t nblk = (uint32_t)((n + kSortChunk - 1) / kSortChunk);
    const uint64_t m = 256ull * nblk;
    uint64_t *src = keys, *dst = alt;
    for (uint32_t shift = 0; shift < 64; shift += 8) {
        hipLaunchKernelGGL(k_sort_hist, dim3(nblk), dim3(kSortThreads), 0, st, src, n, shift, nblk, hist);
        int rc = scan_u32(hist, m, bsum, st);
        if (rc)
            return rc;
        hipLaunchKernelGGL(k_sort_scatter, dim3(nblk), dim3(kSortThreads), 0, st, src, dst, n, shift, nblk,
                           (const uint32_t*)hist);
        PJ_HIP(hipGetLastError());
        std::swap(src, dst);
    }
    return PJ_OK;
}

uint64_t sort_scratch_words(uint64_t n)
{
    const uint64_t nblk = (n + kSortChunk - 1) / kSortChunk;
    return std::max<uint64_t>(256ull * nblk, n) + 16;
}

int alloc_fit(LearnedCtx& c, uint64_t n, FitBufs* f)
{
    int rc;
    const uint64_t sw = sort_scratch_words(n);
    if ((rc = c.sample.ensure(n * 8 + 64)) || (rc = c.tmp.ensure(n * 12 + 64)) || (rc = c.hist.ensure(sw * 4)) ||
        (rc = c.bsum.ensure((scan_blocks(sw) + 2) * 4)) || (rc = c.model.ensure(sizeof(pj_model))) ||
        (rc = c.params.ensure(2 * kLeaves * sizeof(float2))) ||
        (rc = c.moments.ensure(kLeaves * (sizeof(Moments) + sizeof(LeafStat)))) ||
        (rc = c.resid.ensure(kLeaves * sizeof(Resid) + 2 * kLeaves * sizeof(float))) ||
        (rc = c.loss.ensure(64)) || (rc = c.segs.ensure(kMaxSegs * 20 + 64)))
        return rc;
    f->sorted = (uint64_t*)c.sample.p;
    f->alt = (uint64_t*)c.tmp.p;
    f->xs = (float*)(f->alt + n);   // after the sort's ping-pong half
    f->scan = (uint32_t*)c.hist.p;
    f->bsum = (uint32_t*)c.bsum.p;
    f->dm = (pj_model*)c.model.p;
    f->ab[0] = (float2*)c.params.p;
    f->ab[1] = f->ab[0] + kLeaves;
    f->mom = (Moments*)c.moments.p;
    f->stat = (LeafStat*)(f->mom + kLeaves);
    f->res = (Resid*)c.resid.p;
    f->grad = (float*)(f->res + kLeaves);
    f->loss = (double*)c.loss.p;
    f->seg_lo = (uint64_t*)c.segs.p;
    f->seg_hi = f->seg_lo + kMaxSegs;
    f->seg_start = (uint32_t*)(f->seg_hi + kMaxSegs);
    return PJ_OK;
}

uint32_t grid_for(uint64_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }

// Root: segments of the sorted sample and their leaves. Returns false when
// the sample has more segments than the root can hold.
// With too_many, the root first estimates the build partitions a plan
// would need (each segment's keys at the LDS cost its sample density gives,
// at the highest fill, plus one boundary in every gap) and stops with
// *too_many set when they exceed kMaxFine by more than 10%: then neither the
// fit nor the training runs (a large, evenly spread key domain is declined
// this way).
int build_root(const FitBufs& f, uint64_t n, uint64_t n_keys, pj_model* M, bool* ok, hipStream_t st,
               bool* too_many = nullptr)
{
    *ok = false;
    const uint32_t T = 256;
    hipLaunchKernelGGL(k_seg_flags, dim3(grid_for(n, T)), dim3(T), 0, st, f.sorted, n, f.scan);
    int rc = scan_u32(f.scan, n, f.bsum, st);
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_seg_starts, dim3(grid_for(n, T)), dim3(T), 0, st, f.sorted, n, f.scan, kMaxSegs, f.seg_lo,
                       f.seg_hi, f.seg_start);
    PJ_HIP(hipGetLastError());
    uint32_t nseg = 0;
    PJ_HIP(hipMemcpyAsync(&nseg, f.bsum + scan_blocks(n), 4, hipMemcpyDeviceToHost, st));
    PJ_HIP(hipStreamSynchronize(st));
    if (nseg == 0 || nseg > kMaxSegs)
        return PJ_OK;
    std::vector<uint64_t> lo(nseg), hi(nseg);
    std::vector<uint32_t> start(nseg + 1);
    PJ_HIP(hipMemcpyAsync(lo.data(), f.seg_lo, nseg * 8, hipMemcpyDeviceToHost, st));
    PJ_HIP(hipMemcpyAsync(hi.data(), f.seg_hi, nseg * 8, hipMemcpyDeviceToHost, st));
    PJ_HIP(hipMemcpyAsync(start.data(), f.seg_start, nseg * 4, hipMemcpyDeviceToHost, st));
    PJ_HIP(hipStreamSynchronize(st));
    start[nseg] = (uint32_t)n;
    if (too_many) {
        const double scale = (double)n_keys / (double)n;   // keys per sample
        double cost = 0.0;
        for (uint32_t s = 0; s < nseg; ++s) {
            const double keys = (double)(start[s + 1] - start[s]) * scale;
            const double span = (double)(hi[s] - lo[s]) + 1.0;
            cost += keys / (keys / span >= kDenseDensity ? (double)kDenseCap : (double)kCodeCap);
        }
        // 10% headroom: the plan itself counts keys from the trained slopes,
        // which can come out a little below these sample counts
        *too_many = std::ceil(cost / kFillMax) + (double)(nseg - 1) > 1.1 * (double)kMaxFine;
        if (*too_many)
            return PJ_OK;
    }

    // leaves in proportion to sample counts, at least one per segment, one
    // per kLeafSamples samples in all (so a window holds enough samples)
    const uint64_t budget = std::min<uint64_t>(kLeaves, std::max<uint64_t>(1, n / kLeafSamples));
    std::vector<uint32_t> nl(nseg);
    uint64_t total = 0;
    for (uint32_t s = 0; s < nseg; ++s) {
        const uint64_t c = start[s + 1] - start[s];
        nl[s] = (uint32_t)std::max<uint64_t>(1, (uint64_t)((double)budget * (double)c / (double)n));
        total += nl[s];
    }
    while (total > std::max<uint64_t>(budget, nseg)) {   // trim the largest
        uint32_t big = 0;
        for (uint32_t s = 1; s < nseg; ++s)
            if (nl[s] > nl[big])
                big = s;
        nl[big]--;
        total--;
    }
    memset(M, 0, sizeof(*M));
    M->nseg = nseg;
    M->n = n;
    M->n_keys = n_keys;
    M->window = kWindow;
    M->lambda = kLambda;
    uint32_t leaf = 0;
    for (uint32_t s = 0; s < nseg; ++s) {
        const uint64_t span = hi[s] - lo[s];   // + 1 may overflow
        uint64_t lw = span / nl[s] + 1;        // ceil((span + 1) / nl)
        M->seg_lo[s] = lo[s];
        M->seg_hi[s] = hi[s];
        M->seg_lw[s] = lw;
        M->seg_leaf0[s] = leaf;
        for (uint32_t l = 0; l < nl[s]; ++l)
            M->leaf_seg[leaf + l] = (uint16_t)s;
        leaf += nl[s];
    }
    M->seg_leaf0[nseg] = leaf;
    M->nleaves = leaf;
    PJ_HIP(hipMemcpyAsync(f.dm, M, sizeof(*M), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_leaf_starts, dim3(grid_for(n, T)), dim3(T), 0, st, f.dm, f.sorted, n);
    PJ_HIP(hipGetLastError());
    PJ_HIP(hipMemcpyAsync(M->leaf_start, f.dm->leaf_start, sizeof(M->leaf_start), hipMemcpyDeviceToHost, st));
    *ok = true;
    return PJ_OK;
}

// One training step (forward, backward) from ab[cur]; eta > 0 writes the
// stepped parameters to ab[cur ^ 1].
int train_step(const FitBufs& f, uint32_t L, int cur, float eta, float* grad, hipStream_t st)
{
    hipLaunchKernelGGL(k_gd_forward, dim3(L), dim3(kFitThreads), 0, st, f.dm, f.xs, f.ab[cur], f.res);
    hipLaunchKernelGGL(k_gd_backward, dim3(grid_for(L, 256)), dim3(256), 0, st, f.dm, f.ab[cur], f.res, f.stat,
                       kLambda, eta, f.ab[cur ^ 1], grad, f.loss);
    PJ_HIP(hipGetLastError());
    return PJ_OK;
}

// Sorts the sample in f.sorted, builds the root, fits the leaves and runs
// `steps` training steps; M (host) receives the model.
int fit_model(const FitBufs& f, uint64_t n, uint64_t n_keys, uint32_t steps, pj_model* M, bool* ok,
              hipStream_t st, bool* too_many = nullptr)
{
    int rc;
    if ((rc = sort_keys(f.sorted, f.alt, n, f.scan, f.bsum, st)))
        return rc;
    if ((rc = build_root(f, n, n_keys, M, ok, st, too_many)) || !*ok)
        return rc;
    const uint32_t L = M->nleaves;
    hipLaunchKernelGGL(k_leaf_moments, dim3(L), dim3(kFitThreads), 0, st, f.dm, f.sorted, f.mom, f.xs);
    hipLaunchKernelGGL(k_leaf_fit, dim3(grid_for(L, 256)), dim3(256), 0, st, f.dm, f.mom, f.ab[0], f.stat);
    PJ_HIP(hipGetLastError());
    double losses[2] = {0.0, 0.0};
    int cur = 0;
    for (uint32_t t = 0; t <= steps; ++t) {
        // the last pass only measures the loss
        if (t == 0 || t == steps)
            PJ_HIP(hipMemsetAsync(f.loss, 0, 8, st));
        if ((rc = train_step(f, L, cur, t < steps ? kEta : 0.f, nullptr, st)))
            return rc;
        if (t == 0)
            PJ_HIP(hipMemcpyAsync(&losses[0], f.loss, 8, hipMemcpyDeviceToHost, st));
        if (t < steps)
            cur ^= 1;
    }
    PJ_HIP(hipMemcpyAsync(&losses[1], f.loss, 8, hipMemcpyDeviceToHost, st));
    std::vector<float2> ab(L);
    PJ_HIP(hipMemcpyAsync(ab.data(), f.ab[cur], L * sizeof(float2), hipMemcpyDeviceToHost, st));
    PJ_HIP(hipStreamSynchronize(st));
    for (uint32_t j = 0; j < L; ++j) {
        M->a[j] = ab[j].x;
        M->b[j] = ab[j].y;
    }
    M->iters = steps;
    M->loss_lsq = (float)losses[0];
    M->loss = (float)losses[1];
    return PJ_OK;
}

// A model handed in through the C ABI: every index the kernels follow must
// stay inside its arrays and the sample.
bool model_valid(const pj_model& M, uint64_t n)
{
    if (M.nleaves == 0 || M.nleaves > kLeaves || M.nseg == 0 || M.nseg > kMaxSegs || M.n != n)
        return false;
    if (M.seg_leaf0[0] != 0 || M.seg_leaf0[M.nseg] != M.nleaves)
        return false;
    for (uint32_t s = 0; s < M.nseg; ++s)
        if (M.seg_leaf0[s + 1] <= M.seg_leaf0[s] || M.seg_lw[s] == 0 || (s && M.seg_lo[s] <= M.seg_lo[s - 1]))
            return false;
    for (uint32_t j = 0; j < M.nleaves; ++j) {
        const uint32_t s = M.leaf_seg[j];
        if (s >= M.nseg || j < M.seg_leaf0[s] || j >= M.seg_leaf0[s + 1] || M.leaf_start[j + 1] < M.leaf_start[j])
            return false;
    }
    return M.leaf_start[0] == 0 && M.leaf_start[M.nleaves] == n;
}

uint64_t leaf_lo(const pj_model& M, uint32_t j)
{
    const uint32_t s = M.leaf_seg[j];
    return M.seg_lo[s] + (uint64_t)(j - M.seg_leaf0[s]) * M.seg_lw[s];
}

// Boundaries from the model: equal LDS cost per build partition plus one
// boundary in the middle of every gap between segments. A key costs
// 1/kDenseCap in a segment whose sample density says every key value is
// present (the whole segment's samples make that estimate far steadier
// than one leaf's), 1/kCodeCap elsewhere; a leaf's keys come from its slope.
void plan_bounds(const pj_model& M, uint32_t max_coarse, std::vector<uint64_t>& out, uint32_t* div_out)
{
    const uint32_t L = M.nleaves;
    const double scale = M.n ? (double)M.n_keys / (double)M.n : 0.0;   // keys per sample
    std::vector<double> pre(L + 1, 0.0);
    for (uint32_t j = 0; j < L; ++j) {
        const uint32_t s = M.leaf_seg[j];
        const double seg_keys = (double)(M.leaf_start[M.seg_leaf0[s + 1]] - M.leaf_start[M.seg_leaf0[s]]) * scale;
        const double seg_span = (double)(M.seg_hi[s] - M.seg_lo[s]) + 1.0;
        const double cost = seg_keys / seg_span >= kDenseDensity ? 1.0 / kDenseCap : 1.0 / kCodeCap;
        pre[j + 1] = pre[j] + cost * std::max(0.0, (double)M.b[j]) * scale;
    }
    const double C = pre[L];
    const uint32_t gaps = M.nseg - 1;
    // the fewest probe re-reads (div) at a fill of at most kFillMax, then as
    // many build partitions as that div allows, down to a fill of kFill
    const uint32_t nmin = (uint32_t)std::max(1.0, std::ceil(C / kFillMax));
    const uint32_t div = std::max<uint32_t>(1, (nmin + gaps + max_coarse - 1) / max_coarse);
    const uint32_t room = div * max_coarse > gaps ? div * max_coarse - gaps : 1;
    uint32_t ncost = std::max(nmin, (uint32_t)std::ceil(C / kFill));
    if (div > 1)
        ncost = std::max(ncost, room);
    ncost = std::max(nmin, std::min(ncost, room));
    out.clear();
    uint32_t j = 0;
    for (uint32_t t = 1; t < ncost; ++t) {
        const double T = C * (double)t / (double)ncost;
        while (j + 1 < L && pre[j + 1] <= T)
            ++j;
        const double mass = pre[j + 1] - pre[j];
        double frac = mass > 0.0 ? (T - pre[j]) / mass : 0.0;
        frac = std::min(1.0, std::max(0.0, frac));
        const uint64_t lw = M.seg_lw[M.leaf_seg[j]];
        out.push_back(leaf_lo(M, j) + (uint64_t)std::ceil(frac * (double)lw));
    }
    for (uint32_t s = 0; s + 1 < M.nseg; ++s)
        out.push_back(M.seg_hi[s] + (M.seg_lo[s + 1] - M.seg_hi[s]) / 2);
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    if (!out.empty() && out[0] == 0)
        out.erase(out.begin());
    *div_out = std::max<uint32_t>(1, ((uint32_t)out.size() + 1 + max_coarse - 1) / max_coarse);
}

// The learned partition function from the model (its layout: pj_internal.h).
// Every segment gets partitions of its own, as many as its predicted keys
// fill at the planned fill (a key costs 1/kDenseCap of a table in a segment
// whose sample density says every key value is present, 1/kCodeCap
// elsewhere), and leaf j's line in partition units continues the previous
// leaf's: c[j] is the cost of the segment's leaves before j, d[j] leaf j's
// own, both scaled to the segment's partitions. The fill is the lowest in
// [kFill, kFillMax] that keeps the fewest probe re-reads (div: build
// partitions a probe partition serves) within max_coarse probe partitions.
// Returns false with plan->why when the partitions do not fit.
bool plan_map(const pj_model& M, uint32_t max_coarse, RangePlan* plan)
{
    const uint32_t L = M.nleaves, S = M.nseg;
    const double scale = M.n ? (double)M.n_keys / (double)M.n : 0.0;   // keys per sample
    std::vector<double> cost(L, 0.0), segc(S, 0.0);                       // tables at a fill of 1
    for (uint32_t s = 0; s < S; ++s) {
        const double seg_keys = (double)(M.leaf_start[M.seg_leaf0[s + 1]] - M.leaf_start[M.seg_leaf0[s]]) * scale;
        const double span = (double)(M.seg_hi[s] - M.seg_lo[s]) + 1.0;
        const double cpk = seg_keys / span >= kDenseDensity ? 1.0 / kDenseCap : 1.0 / kCodeCap;
        for (uint32_t j = M.seg_leaf0[s]; j < M.seg_leaf0[s + 1]; ++j) {
            cost[j] = std::max(0.0, (double)M.b[j]) * scale * cpk;
            segc[s] += cost[j];
        }
    }
    auto parts = [&](double fill) {
        uint64_t t = 0;
        for (uint32_t s = 0; s < S; ++s)
            t += std::max<uint64_t>(1, (uint64_t)std::ceil(segc[s] / fill));
        return t;
    };
    const uint64_t nmin = parts(kFillMax);
    if (nmin > kMaxFine) {
        plan->why = "too many partitions";
        return false;
    }
    const uint32_t div = (uint32_t)std::max<uint64_t>(1, (nmin + max_coarse - 1) / max_coarse);
    const uint64_t room = std::min<uint64_t>((uint64_t)div * max_coarse, kMaxFine);
    double fill = kFill;
    while (fill < kFillMax && parts(fill) > room)
        fill += 0.005;
    if (parts(fill) > room)
        fill = kFillMax;
    plan->map.assign(kMapBytes, 0);
    uint8_t* m = plan->map.data();
    MapHdr* hdr = reinterpret_cast<MapHdr*>(m);
    uint64_t* seg_lo = reinterpret_cast<uint64_t*>(m + kMapOffSegLo);
    uint64_t* seg_cut = reinterpret_cast<uint64_t*>(m + kMapOffSegCut);
    float* seg_inv = reinterpret_cast<float*>(m + kMapOffSegInv);
    uint32_t* seg_meta = reinterpret_cast<uint32_t*>(m + kMapOffSegMeta);
    uint16_t* dir = reinterpret_cast<uint16_t*>(m + kMapOffDir);
    float* leaf = reinterpret_cast<float*>(m + kMapOffLeaf);
    uint32_t first = 0;
    for (uint32_t s = 0; s < S; ++s) {
        const uint32_t np = (uint32_t)std::max<uint64_t>(1, (uint64_t)std::ceil(segc[s] / fill));
        const uint32_t l0 = M.seg_leaf0[s], nl = M.seg_leaf0[s + 1] - l0;
        const double k = segc[s] > 0.0 ? (double)np / segc[s] : 0.0;   // partitions per table of cost
        double acc = 0.0;
        for (uint32_t j = l0; j < l0 + nl; ++j) {
            leaf[2 * j] = (float)(acc * k);
            leaf[2 * j + 1] = (float)(cost[j] * k);
            acc += cost[j];
        }
        seg_lo[s] = M.seg_lo[s];
        // the middle of the gap before the segment (segments are cut at gaps wider than 2^32)
        seg_cut[s] = s ? M.seg_hi[s - 1] + (M.seg_lo[s] - M.seg_hi[s - 1]) / 2 + 1 : M.seg_lo[0];
        seg_inv[s] = (float)(1.0 / (double)M.seg_lw[s]);
        seg_meta[2 * s] = l0 | (nl << 16);
        seg_meta[2 * s + 1] = first | (np << 16);
        first += np;
    }
    hdr->nseg = S;
    hdr->nleaves = L;
    hdr->nfine = first;
    hdr->lo0 = seg_cut[0];
    const uint64_t range = seg_cut[S - 1] - seg_cut[0];
    uint32_t sh = 0;
    while (sh < 63 && (range >> sh) >= kMapDir)
        ++sh;
    hdr->dshift = sh;
    for (uint32_t d = 0; d <= kMapDir; ++d) {   // the last segment starting at or below slice d's first key
        const uint64_t off = (uint64_t)d << sh;
        uint32_t s = S - 1;
        if ((off >> sh) == d && off <= range) {
            const uint64_t start = seg_cut[0] + off;
            s = (uint32_t)(std::upper_bound(seg_cut, seg_cut + S, start) - seg_cut) - 1;
        }
        dir[d] = (uint16_t)s;
    }
    plan->nfine = first;
    plan->div = div;
    plan->ncoarse = (first + div - 1) / div;
    if (plan->ncoarse > max_coarse || first > kMaxFine || first < 4) {
        plan->why = first < 4 ? "too few partitions" : "too many partitions";
        return false;
    }
    plan->coarse_of.resize(first);
    for (uint32_t f = 0; f < first; ++f)
        plan->coarse_of[f] = f / div;
    return true;
}

}  // namespace

int learned_plan(LearnedCtx& ctx, const void* d_build, uint64_t nb, const SideLayout& Ls, uint32_t max_coarse,
                 hipStream_t st, RangePlan* plan)
{
    *plan = RangePlan();
    const uint64_t ns = (nb + kSampleStride - 1) / kSampleStride;
    if (ns < 64) {
        plan->why = "too few build keys to sample";
        return PJ_OK;
    }
    FitBufs f;
    int rc;
    if ((rc = alloc_fit(ctx, ns, &f)))
        return rc;
    const bool ident = Ls.ident && ((uintptr_t)d_build % 16 == 0);
    if (ident)
        hipLaunchKernelGGL(k_sample<true>, dim3(grid_for(ns, 256)), dim3(256), 0, st, (const uint32_t*)d_build,
                           kSampleStride, ns, Ls, f.sorted);
    else
        hipLaunchKernelGGL(k_sample<false>, dim3(grid_for(ns, 256)), dim3(256), 0, st, (const uint32_t*)d_build,
                           kSampleStride, ns, Ls, f.sorted);
    PJ_HIP(hipGetLastError());
    ctx.host_model.resize(sizeof(pj_model));
    pj_model* M = (pj_model*)ctx.host_model.data();
    bool ok = false, too_many = false;
    if ((rc = fit_model(f, ns, nb, kSteps, M, &ok, st, &too_many)))
        return rc;
    if (too_many) {
        plan->why = "too many partitions";
        return PJ_OK;
    }
    if (!ok) {
        plan->why = "key domain too fragmented for the root";
        return PJ_OK;
    }
    if (!plan_map(*M, max_coarse, plan))
        return PJ_OK;
    plan->nseg = M->nseg;
    plan->nleaves = M->nleaves;
    plan->loss_lsq = M->loss_lsq;
    plan->loss = M->loss;
    plan->ok = true;
    return PJ_OK;
}

}  // namespace pj

using namespace pj;

extern "C" {

uint64_t pj_model_size(void) { return sizeof(pj_model); }

int pj_sort_u64(uint64_t* d_keys, uint64_t n, void* stream)
{
    if (n && !d_keys)
        return fail(PJ_ERR_INVALID, "null argument");
    if (n >= (1ull << 31))
        return fail(PJ_ERR_UNSUPPORTED, "sorts fewer than 2^31 keys");
    hipStream_t st = (hipStream_t)stream;
    DevBuf alt, hist, bsum;
    int rc;
    const uint64_t sw = sort_scratch_words(n);
    if ((rc = alt.ensure(n * 8 + 64)) || (rc = hist.ensure(sw * 4)) || (rc = bsum.ensure((scan_blocks(sw) + 2) * 4)))
        return rc;
    if ((rc = sort_keys(d_keys, (uint64_t*)alt.p, n, (uint32_t*)hist.p, (uint32_t*)bsum.p, st)))
        return rc;
    PJ_HIP(hipStreamSynchronize(st));
    return PJ_OK;
}

int pj_model_fit(const uint64_t* d_keys, uint64_t n, uint64_t n_keys, uint32_t iters, pj_model* model, void* stream)
{
    if (!d_keys || !model || n < 2)
        return fail(PJ_ERR_INVALID, "need a model and at least two keys");
    if (n >= (1ull << 31))
        return fail(PJ_ERR_UNSUPPORTED, "fits fewer than 2^31 keys");
    hipStream_t st = (hipStream_t)stream;
    LearnedCtx ctx;
    FitBufs f;
    int rc;
    if ((rc = alloc_fit(ctx, n, &f)))
        return rc;
    PJ_HIP(hipMemcpyAsync(f.sorted, d_keys, n * 8, hipMemcpyDeviceToDevice, st));
    bool ok = false;
    if ((rc = fit_model(f, n, n_keys, iters, model, &ok, st)))
        return rc;
    if (!ok)
        return fail(PJ_ERR_UNSUPPORTED, "more than %u key segments", kMaxSegs);
    return PJ_OK;
}

int pj_model_grad(const pj_model* model, const uint64_t* d_sorted, uint64_t n, float* grad, double* loss,
                  void* stream)
{
    if (!model || !d_sorted || !grad || !loss || !model_valid(*model, n))
        return fail(PJ_ERR_INVALID, "bad model or sample");
    hipStream_t st = (hipStream_t)stream;
    LearnedCtx ctx;
    FitBufs f;
    int rc;
    if ((rc = alloc_fit(ctx, n, &f)))
        return rc;
    const uint32_t L = model->nleaves;
    PJ_HIP(hipMemcpyAsync(f.sorted, d_sorted, n * 8, hipMemcpyDeviceToDevice, st));
    PJ_HIP(hipMemcpyAsync(f.dm, model, sizeof(*model), hipMemcpyHostToDevice, st));
    std::vector<float2> ab(L);
    for (uint32_t j = 0; j < L; ++j)
        ab[j] = make_float2(model->a[j], model->b[j]);
    PJ_HIP(hipMemcpyAsync(f.ab[0], ab.data(), L * sizeof(float2), hipMemcpyHostToDevice, st));
    PJ_HIP(hipMemsetAsync(f.loss, 0, 8, st));
    hipLaunchKernelGGL(k_leaf_moments, dim3(L), dim3(kFitThreads), 0, st, f.dm, f.sorted, f.mom, f.xs);
    if ((rc = train_step(f, L, 0, 0.f, f.grad, st)))
        return rc;
    PJ_HIP(hipMemcpyAsync(grad, f.grad, 2 * L * sizeof(float), hipMemcpyDeviceToHost, st));
    PJ_HIP(hipMemcpyAsync(loss, f.loss, 8, hipMemcpyDeviceToHost, st));
    PJ_HIP(hipStreamSynchronize(st));
    return PJ_OK;
}

int pj_model_bounds(const pj_model* model, uint32_t max_coarse, uint64_t* bounds, uint32_t cap, uint32_t* nbounds)
{
    if (!model || !nbounds || (cap && !bounds) || max_coarse == 0 || model->nleaves == 0 ||
        model->nleaves > PJ_MODEL_LEAVES || model->nseg == 0 || model->nseg > PJ_MODEL_MAX_SEGS)
        return fail(PJ_ERR_INVALID, "bad model");
    std::vector<uint64_t> b;
    uint32_t div = 1;
    plan_bounds(*model, max_coarse, b, &div);
    *nbounds = (uint32_t)b.size();
    if (b.size() > cap)
        return fail(PJ_ERR_CAPACITY, "%zu boundaries, room for %u", b.size(), cap);
    if (!b.empty())
        memcpy(bounds, b.data(), b.size() * 8);
    return PJ_OK;
}

}  // extern "C"
