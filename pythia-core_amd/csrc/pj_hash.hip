// pj_hash.hip -- the reference's TupleHasher bucket functions on the device:
// one bucket id per tuple, as TupleHasher::hash (hash.h:99-102) gives it for
// the function TupleHasher::create (hash.cpp:101-248) builds from a plan's
// `hash` node. Value hashers read one INTEGER (sign-extended) or LONG/DATE
// key (numericalize, hash.h:223-242); the byte hasher folds a byte range.
// The bucket ids are part of what a plan sees wherever the reference exposes
// them (PartitionOp's partitions, a hash-partitioned exchange), so the
// formulas are the reference's own.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pj_internal.h"

namespace pj {
namespace {

// Parameters of one hash function, as HashFunction's constructors set them.
struct HashParams {
    uint32_t fn;
    uint32_t buckets;     // HashFunction::buckets()
    uint32_t skipbits;
    uint32_t pad;
    int64_t k;            // a mask, a shift or a bucket count, by function
    int64_t min, max;
    int64_t bucketrange;  // exactrange
};

// getlogarithm (hash.cpp:49-58): the bits that number b buckets.
uint32_t log2_ceil(uint32_t b)
{
    uint32_t lg = 0;
    while (lg < 32 && (1ull << lg) < b)
        ++lg;
    return lg;
}

int hash_params(const pj_hash_spec* s, HashParams* p)
{
    *p = HashParams{};
    p->fn = s->fn;
    switch (s->fn) {
    case PJ_HASH_ALWAYSZERO:   // HashFunction(1), hash.h:130-136
        p->buckets = 1;
        return PJ_OK;
    case PJ_HASH_TPCHQ1MAGIC:  // HashFunction(4), hash.h:138-147
        p->k = 2;
        p->buckets = 4;
        return PJ_OK;
    default:
        break;
    }
    if (s->buckets == 0)
        return fail(PJ_ERR_INVALID, "Number of hash buckets cannot be zero.");   // hash.cpp:61-67
    const uint32_t lg = log2_ceil(s->buckets);
    if (lg > 31 && s->fn != PJ_HASH_EXACTRANGE)
        return fail(PJ_ERR_INVALID, "too many hash buckets");
    switch (s->fn) {
    case PJ_HASH_MODULO:
    case PJ_HASH_TPCHORDERKEY:
    case PJ_HASH_WILLIS:       // the bits as a mask (hash.h:277-281)
        p->k = (int64_t)((1ull << lg) - 1);
        p->buckets = 1u << lg;
        return PJ_OK;
    case PJ_HASH_PARAMMODULO:
    case PJ_HASH_KNUTH:        // the mask moved up by skipbits (hash.h:317-320)
        p->k = (int64_t)(((1ull << lg) - 1) << s->skipbits);
        p->min = (int64_t)s->offset;
        p->skipbits = s->skipbits;
        p->buckets = 1u << lg;
        return PJ_OK;
    case PJ_HASH_RANGE:
    case PJ_HASH_BYTES:        // the bit count itself
        if (s->fn == PJ_HASH_RANGE && (s->range_max < s->range_min ||
                                       (uint64_t)s->range_max - (uint64_t)s->range_min >= (uint64_t)INT64_MAX))
            return fail(PJ_ERR_INVALID, "range [%lld, %lld] is empty or spans every key", (long long)s->range_min,
                        (long long)s->range_max);
        p->k = lg;
        p->min = s->range_min;
        p->max = s->range_max;
        p->buckets = 1u << lg;
        return PJ_OK;
    case PJ_HASH_EXACTRANGE:   // hash.h:455-460
        if (s->range_max < s->range_min)
            return fail(PJ_ERR_INVALID, "range [%lld, %lld] is empty", (long long)s->range_min,
                        (long long)s->range_max);
        if ((uint64_t)s->range_max - (uint64_t)s->range_min > (uint64_t)INT64_MAX - s->buckets)
            return fail(PJ_ERR_INVALID, "range [%lld, %lld] over %u buckets overflows a LONG",
                        (long long)s->range_min, (long long)s->range_max, s->buckets);
        p->k = s->buckets;
        p->min = s->range_min;
        p->max = s->range_max;
        p->bucketrange = (p->max - p->min + p->k) / p->k;
        p->buckets = s->buckets;
        return PJ_OK;
    default:
        return fail(PJ_ERR_INVALID, "unknown hash function %u", s->fn);   // UnknownHashException
    }
}

// Differences of keys wrap (as the reference's int64 arithmetic does on the
// machines it runs on) instead of overflowing.
__device__ __forceinline__ uint32_t hash_value(const HashParams& h, int64_t v)
{
    switch (h.fn) {
    case PJ_HASH_MODULO:       // hash.h:284-287
        return (uint32_t)(v & h.k);
    case PJ_HASH_PARAMMODULO:  // hash.h:323-325
        return (uint32_t)((((uint64_t)v - (uint64_t)h.min) & (uint64_t)h.k) >> h.skipbits);
    case PJ_HASH_KNUTH:        // hash.h:369-371 (the offset is not used)
        return (uint32_t)((((uint64_t)v * 2654435761ull) & (uint64_t)h.k) >> h.skipbits);
    case PJ_HASH_RANGE: {      // hash.h:256-261
        const int64_t val = (int64_t)(((uint64_t)v - (uint64_t)h.min) << h.k);
        return (uint32_t)(val / (h.max - h.min + 1));
    }
    case PJ_HASH_EXACTRANGE:   // hash.h:462-466
        return (uint32_t)((int64_t)((uint64_t)v - (uint64_t)h.min) / h.bucketrange);
    case PJ_HASH_TPCHORDERKEY: // hash.h:398-400
        return (uint32_t)((((v >> 2) & ~7ll) | (v & 7)) & h.k);
    case PJ_HASH_WILLIS: {     // hash.h:413-424
        uint64_t l = (uint64_t)v;
        l = (~l) + (l << 21);
        l = l ^ (uint64_t)((int64_t)l >> 24);
        l = (l + (l << 3)) + (l << 8);
        l = l ^ (uint64_t)((int64_t)l >> 14);
        l = (l + (l << 2)) + (l << 4);
        l = l ^ (uint64_t)((int64_t)l >> 28);
        l = l + (l << 31);
        if (!((int64_t)l > 0))
            l = 0 - l;
        return (uint32_t)(l & (uint64_t)h.k);
    }
    default:
        return 0;
    }
}

__global__ __launch_bounds__(256) void k_tuple_hash(const uint8_t* __restrict__ in, uint64_t n, uint32_t tw,
                                                    uint32_t off, uint32_t size, HashParams h,
                                                    uint32_t* __restrict__ out)
{
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
        const uint8_t* p = in + i * tw + off;
        uint32_t b = 0;
        if (h.fn == PJ_HASH_ALWAYSZERO) {
            b = 0;
        } else if (h.fn == PJ_HASH_TPCHQ1MAGIC) {   // hash.h:142-146
            const uint32_t v = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
            b = (((v >> 4) | (v >> 16)) & 0x1u) | ((v >> 1) & 0x2u);
        } else if (h.fn == PJ_HASH_BYTES) {         // hash.h:165-197: FNV-1a over sign-extended bytes, folded
            uint64_t x = 14695981039346656037ull;
            for (uint32_t c = 0; c < size; ++c) {
                x ^= (uint64_t)(int64_t)(int8_t)p[c];
                x += (x << 1) + (x << 4) + (x << 5) + (x << 7) + (x << 8) + (x << 40);
            }
            const uint64_t sh = (uint64_t)((int64_t)x >> h.k);
            b = (uint32_t)((sh ^ x) & ((1ull << h.k) - 1));
        } else {                                    // numericalize: INTEGER sign-extended, LONG as is
            uint64_t v = 0;
            for (uint32_t c = 0; c < size; ++c)
                v |= (uint64_t)p[c] << (8 * c);
            const int64_t sv = size == 4 ? (int64_t)(int32_t)(uint32_t)v : (int64_t)v;
            b = hash_value(h, sv);
        }
        out[i] = b;
    }
}

}  // namespace
}  // namespace pj

using namespace pj;

extern "C" {

int pj_hash_buckets(const pj_hash_spec* spec, uint32_t* buckets)
{
    if (!spec || !buckets)
        return fail(PJ_ERR_INVALID, "null argument");
    HashParams p;
    const int rc = hash_params(spec, &p);
    if (rc)
        return rc;
    *buckets = p.buckets;
    return PJ_OK;
}

// ParameterizedModuloValueHasher::generate (hash.cpp:251-281): t = the bits
// of spec's buckets (getlogarithm(buckets() - 1)); passes - 1 functions of
// t / passes bits each, from the top of the mask down, then one of the bits
// left over at skipbits. The reference's unsigned loop bound passes - 1 has
// no meaning at passes = 0, and a one-bucket function has no bits to split
// (getlogarithm(0) is 32 there); both are refused.
int pj_hash_generate(const pj_hash_spec* spec, uint32_t passes, pj_hash_spec* out)
{
    if (!spec || !out)
        return fail(PJ_ERR_INVALID, "null argument");
    if (spec->fn != PJ_HASH_PARAMMODULO)
        return fail(PJ_ERR_INVALID, "generate is defined for parammodulo functions only");
    if (passes == 0)
        return fail(PJ_ERR_INVALID, "generate needs at least one pass");
    HashParams h;
    int rc = hash_params(spec, &h);
    if (rc)
        return rc;
    if (h.buckets < 2)
        return fail(PJ_ERR_INVALID, "a one-bucket function has no bits to split");
    const uint32_t total = log2_ceil(h.buckets - 1);
    uint32_t per = total / passes;
    for (uint32_t i = 0; i < passes; ++i) {
        out[i] = *spec;
        if (i + 1 < passes) {
            out[i].buckets = 1u << per;
            out[i].skipbits = spec->skipbits + total - (i + 1) * per;
        } else {
            out[i].buckets = 1u << (total - (passes - 1) * per);
            out[i].skipbits = spec->skipbits;
        }
    }
    return PJ_OK;
}

int pj_tuple_hash_device(const pj_hash_spec* spec, const void* d_tuples, uint64_t n, uint32_t tuple_width,
                         uint32_t key_offset, uint32_t key_size, uint32_t* d_buckets, void* stream)
{
    if (!spec || (n && (!d_tuples || !d_buckets)))
        return fail(PJ_ERR_INVALID, "null argument");
    HashParams p;
    int rc = hash_params(spec, &p);
    if (rc)
        return rc;
    const bool value = spec->fn != PJ_HASH_BYTES && spec->fn != PJ_HASH_ALWAYSZERO && spec->fn != PJ_HASH_TPCHQ1MAGIC;
    if (value && key_size != 4 && key_size != 8)
        return fail(PJ_ERR_UNSUPPORTED, "value hashers read a 4- or 8-byte key, not %u bytes", key_size);
    if (spec->fn == PJ_HASH_TPCHQ1MAGIC && key_size < 4)
        return fail(PJ_ERR_UNSUPPORTED, "tpchq1magic reads 4 bytes");
    if (spec->fn == PJ_HASH_BYTES && key_size == 0)
        return fail(PJ_ERR_INVALID, "the byte hasher needs at least one byte");
    if ((uint64_t)key_offset + key_size > tuple_width)
        return fail(PJ_ERR_INVALID, "key bytes [%u, %u) lie outside a %u-byte tuple", key_offset,
                    key_offset + key_size, tuple_width);
    if (n == 0)
        return PJ_OK;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_tuple_hash, dim3(grid), dim3(256), 0, (hipStream_t)stream, (const uint8_t*)d_tuples, n,
                       tuple_width, key_offset, key_size, p, d_buckets);
    PJ_HIP(hipGetLastError());
    return PJ_OK;
}

}  // extern "C"
