/*
 * oracle.c -- CPU restatement of Pythia's hash-join path (checker only).
 * Test infrastructure: never linked into or called by the product.
 * File:line citations are to the reference tree.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ schema */

/* schema.cpp:82-117: columns packed back to back, no alignment. */
int or_schema_init(or_schema* s, int ncols, const int* types, const int* widths)
{
    if (ncols < 0 || ncols > OR_MAX_COLS)
        return -1;
    s->ncols = ncols;
    s->size = 0;
    for (int i = 0; i < ncols; ++i) {
        int w;
        switch (types[i]) {
        case OR_CT_INTEGER: w = 4; break;
        case OR_CT_LONG: w = 8; break;
        case OR_CT_DECIMAL: w = 8; break;
        case OR_CT_DATE: w = 8; break;     /* CtDate is CtLong-sized */
        case OR_CT_POINTER: w = 8; break;
        case OR_CT_CHAR:
            w = widths ? widths[i] : 0;
            if (w < 2)
                return -1;                 /* schema.cpp:97-103 */
            break;
        default:
            return -1;
        }
        s->type[i] = types[i];
        s->width[i] = w;
        s->offset[i] = s->size;
        s->size += w;
    }
    return 0;
}

/* schema.inl:47-93 writeData: a typed copy; CHAR is strncpy. */
static void write_col(void* dst, int type, int width, const void* src)
{
    if (type == OR_CT_CHAR)
        strncpy((char*)dst, (const char*)src, (size_t)width);
    else if (type == OR_CT_INTEGER)
        memcpy(dst, src, 4);
    else
        memcpy(dst, src, 8);
}

/* ------------------------------------------------------------------ hashes */

/* hash.cpp:49-58: log2 of the next power of two (ceil(log2 k)). */
unsigned or_getlogarithm(unsigned k)
{
    --k;
    for (unsigned i = 0; i < 32; ++i)
        if ((0xFFFFFFFFu << (31 - i)) & k)
            return 32 - i;
    return 0;
}

int or_fn_from_name(const char* n)
{
    static const struct { const char* name; int fn; } tbl[] = {
        {"modulo", OR_FN_MODULO}, {"parammodulo", OR_FN_PARAMMODULO},
        {"knuth", OR_FN_KNUTH}, {"range", OR_FN_RANGE},
        {"exactrange", OR_FN_EXACTRANGE}, {"bytes", OR_FN_BYTES},
        {"tpchorderkey", OR_FN_TPCHORDERKEY}, {"willis", OR_FN_WILLIS},
        {"alwayszero", OR_FN_ALWAYSZERO}, {"tpchq1magic", OR_FN_TPCHQ1MAGIC},
    };
    for (size_t i = 0; i < sizeof(tbl) / sizeof(tbl[0]); ++i)
        if (strcmp(n, tbl[i].name) == 0)
            return tbl[i].fn;
    return -1;
}

/* Constructors of hash.h:104-487 as chosen by hash.cpp:101-248. */
int or_hasher_init(or_hasher* h, int fn, unsigned buckets, int64_t range_min,
                   int64_t range_max, unsigned offset_param, unsigned skipbits)
{
    memset(h, 0, sizeof(*h));
    h->fn = fn;
    if (fn == OR_FN_ALWAYSZERO) {           /* HashFunction(1): _k = 0 */
        h->k = 0;
        return 0;
    }
    if (fn == OR_FN_TPCHQ1MAGIC) {          /* HashFunction(4) */
        h->k = 2;
        return 0;
    }
    if (buckets == 0)
        return -1;                          /* hash.cpp:61-67 */
    int64_t lg = or_getlogarithm(buckets);
    switch (fn) {
    case OR_FN_BYTES:
    case OR_FN_RANGE:
        h->k = lg;
        h->min = range_min;
        h->max = range_max;
        return 0;
    case OR_FN_MODULO:
    case OR_FN_TPCHORDERKEY:
    case OR_FN_WILLIS:
        h->k = (int64_t)((1u << lg) - 1);   /* hash.h:277-281: _k as a mask */
        return 0;
    case OR_FN_PARAMMODULO:
    case OR_FN_KNUTH:
        h->k = (int64_t)((1u << lg) - 1);
        h->k <<= skipbits;                  /* hash.h:317-320 */
        h->min = (int64_t)offset_param;
        h->skipbits = skipbits;
        return 0;
    case OR_FN_EXACTRANGE:
        h->min = range_min;
        h->max = range_max;
        h->k = buckets;                     /* hash.h:455-460 */
        h->bucketrange = (h->max - h->min + h->k) / h->k;
        return 0;
    default:
        return -1;
    }
}

unsigned or_hasher_buckets(const or_hasher* h)
{
    switch (h->fn) {
    case OR_FN_MODULO:
    case OR_FN_TPCHORDERKEY:
    case OR_FN_WILLIS:
        return (unsigned)(h->k + 1);
    case OR_FN_PARAMMODULO:
    case OR_FN_KNUTH:
        return (unsigned)((h->k >> h->skipbits) + 1);
    case OR_FN_EXACTRANGE:
        return (unsigned)h->k;
    default:
        return 1u << h->k;
    }
}

unsigned or_hasher_hash_value(const or_hasher* h, int64_t v)
{
    switch (h->fn) {
    case OR_FN_MODULO:                      /* hash.h:284-287 */
        return (unsigned)(v & h->k);
    case OR_FN_PARAMMODULO:                 /* hash.h:323-325 */
        return (unsigned)(((uint64_t)(v - h->min) & (uint64_t)h->k) >> h->skipbits);
    case OR_FN_KNUTH:                       /* hash.h:369-371: offset unused */
        return (unsigned)((((uint64_t)v * 2654435761ull) & (uint64_t)h->k) >> h->skipbits);
    case OR_FN_RANGE: {                     /* hash.h:256-261 */
        int64_t val = v - h->min;
        val = (int64_t)((uint64_t)val << h->k);
        return (unsigned)(val / (h->max - h->min + 1));
    }
    case OR_FN_EXACTRANGE:                  /* hash.h:462-466 */
        return (unsigned)((v - h->min) / h->bucketrange);
    case OR_FN_TPCHORDERKEY:                /* hash.h:398-400 */
        return (unsigned)((((v >> 2) & ~7LL) | (v & 7)) & h->k);
    case OR_FN_WILLIS: {                    /* hash.h:413-424 */
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
        return (unsigned)(l & (uint64_t)h->k);
    }
    default:
        return 0;
    }
}

/* hash.h:223-242: 4-byte fields are sign-extended CtInt, 8-byte CtLong. */
static int numericalize(const void* p, size_t size, int64_t* out)
{
    if (size == 4) {
        int32_t v;
        memcpy(&v, p, 4);
        *out = v;
        return 0;
    }
    if (size == 8) {
        memcpy(out, p, 8);
        return 0;
    }
    return -1;
}

unsigned or_hasher_hash(const or_hasher* h, const void* start, size_t size)
{
    switch (h->fn) {
    case OR_FN_ALWAYSZERO:
        return 0;
    case OR_FN_TPCHQ1MAGIC: {               /* hash.h:142-146 */
        unsigned v;
        memcpy(&v, start, 4);
        return (((v >> 4) | (v >> 16)) & 0x1u) | ((v >> 1) & 0x2u);
    }
    case OR_FN_BYTES: {                     /* hash.h:165-197 */
        uint64_t hash = 14695981039346656037ull;
        const signed char* c = (const signed char*)start;
        const signed char* e = c + size;
        do {
            hash ^= (uint64_t)(int64_t)*c++;          /* sign-extended char */
            hash += (hash << 1) + (hash << 4) + (hash << 5) + (hash << 7) +
                    (hash << 8) + (hash << 40);
        } while (c < e);
        uint64_t sh = (uint64_t)((int64_t)hash >> h->k);  /* CtLong shift */
        hash = (sh ^ hash) & ((1ull << h->k) - 1);
        return (unsigned)hash;
    }
    default: {
        int64_t v;
        if (numericalize(start, size, &v) != 0)
            return 0xFFFFFFFFu;
        return or_hasher_hash_value(h, v);
    }
    }
}

int64_t or_exactrange_minimumforbucket(const or_hasher* h, unsigned bucket)
{
    if ((int64_t)bucket == h->k)
        return h->max + 1;
    return (int64_t)bucket * h->bucketrange + h->min;
}

/* hash.cpp:251-281. */
unsigned or_parammodulo_generate(int64_t offset, unsigned buckets, unsigned skipbits,
                                 unsigned passes, unsigned* out_buckets,
                                 unsigned* out_skipbits)
{
    (void)offset;
    or_hasher self;
    or_hasher_init(&self, OR_FN_PARAMMODULO, buckets, 0, 0, 0, skipbits);
    unsigned totalbits = or_getlogarithm(or_hasher_buckets(&self) - 1);
    unsigned per = totalbits / passes;
    unsigned i;
    for (i = 0; i + 1 < passes; ++i) {
        out_skipbits[i] = skipbits + totalbits - (i + 1) * per;
        out_buckets[i] = 1u << per;
    }
    per = totalbits - (passes - 1) * per;
    out_skipbits[i] = skipbits;
    out_buckets[i] = 1u << per;
    for (unsigned j = 0; j < passes; ++j) {  /* buckets() of each function */
        or_hasher t;
        or_hasher_init(&t, OR_FN_PARAMMODULO, out_buckets[j], 0, 0, 0, out_skipbits[j]);
        out_buckets[j] = or_hasher_buckets(&t);
    }
    return passes;
}

/* operators/join.cpp:340-359. */
unsigned or_find_max_contiguous(const uint16_t* v1, const uint16_t* v2, unsigned n)
{
    unsigned r = 0;
    while (r < n - 1 && v1[r] + 1 == v1[r + 1] && v2[r] + 1 == v2[r + 1])
        ++r;
    return r;
}

/* --------------------------------------------------------------- hash table */

/* util/hashtable.h:262-267: {Lock, ushort used, next}. The lock is not
 * needed single-threaded; used and next keep their meaning. */
typedef struct bucket {
    unsigned used;
    struct bucket* next;
    unsigned char data[];
} bucket;

typedef struct {
    unsigned nbuckets, bucksize, tuplesize;
    size_t stride;
    unsigned char* mem;
    uint64_t spills;
} htable;

static bucket* ht_header(htable* t, unsigned b)
{
    return (bucket*)(t->mem + (size_t)b * t->stride);
}

static int ht_init(htable* t, unsigned nbuckets, unsigned bucksize, unsigned tuplesize)
{
    t->nbuckets = nbuckets;
    t->bucksize = bucksize;
    t->tuplesize = tuplesize;
    t->stride = sizeof(bucket) + bucksize;
    t->spills = 0;
    t->mem = (unsigned char*)calloc(nbuckets, t->stride);
    return t->mem ? 0 : -1;
}

/* util/hashtable.cpp:254-297: first bucket in the chain with room, else a
 * new overflow bucket linked at the end. */
static void* ht_allocate(htable* t, unsigned b)
{
    bucket* last = NULL;
    for (bucket* h = ht_header(t, b); h; h = h->next) {
        if (h->used + t->tuplesize <= t->bucksize) {
            void* at = h->data + h->used;
            h->used += t->tuplesize;
            return at;
        }
        last = h;
    }
    t->spills++;
    bucket* nb = (bucket*)calloc(1, t->stride);
    if (!nb)
        return NULL;
    nb->used = t->tuplesize;
    last->next = nb;
    return nb->data;
}

static void ht_destroy(htable* t)
{
    for (unsigned b = 0; b < t->nbuckets; ++b) {
        bucket* h = ht_header(t, b)->next;
        while (h) {
            bucket* n = h->next;
            free(h);
            h = n;
        }
    }
    free(t->mem);
    t->mem = NULL;
}

/* util/hashtable.h:126-171: Iterator placed on a bucket, walking its chain. */
typedef struct {
    unsigned char *cur, *free;
    bucket* nxt;
    unsigned tuplesize;
} ht_iter;

static void ht_place(htable* t, ht_iter* it, unsigned b)
{
    bucket* h = ht_header(t, b);
    it->cur = h->data;
    it->free = h->data + h->used;
    it->nxt = h->next;
    it->tuplesize = t->tuplesize;
}

static void* ht_next(ht_iter* it)
{
    if (it->cur && it->cur < it->free) {
        void* r = it->cur;
        it->cur += it->tuplesize;
        return r;
    }
    if (it->nxt) {
        bucket* h = it->nxt;
        unsigned char* r = h->data;
        it->cur = r + it->tuplesize;
        it->free = r + h->used;
        it->nxt = h->next;
        return r < it->free ? r : NULL;
    }
    return NULL;
}

/* --------------------------------------------------------------- hash join */

/* rawcompfns.cpp:48-180: Int/Long equality, with the narrower operand
 * widened; DATE compares as CtLong; DECIMAL as double. */
static int key_equal(int t1, const void* a, int t2, const void* b)
{
    if ((t1 == OR_CT_DECIMAL) != (t2 == OR_CT_DECIMAL)) {
        double x, y;
        int64_t v;
        if (t1 == OR_CT_DECIMAL) {
            memcpy(&x, a, 8);
            if (t2 == OR_CT_INTEGER) { int32_t i; memcpy(&i, b, 4); y = i; }
            else { memcpy(&v, b, 8); y = (double)v; }
        } else {
            memcpy(&y, b, 8);
            if (t1 == OR_CT_INTEGER) { int32_t i; memcpy(&i, a, 4); x = i; }
            else { memcpy(&v, a, 8); x = (double)v; }
        }
        return x == y;
    }
    if (t1 == OR_CT_DECIMAL) {
        double x, y;
        memcpy(&x, a, 8);
        memcpy(&y, b, 8);
        return x == y;
    }
    if (t1 == OR_CT_CHAR || t2 == OR_CT_CHAR)
        return -1;
    int64_t x, y;
    if (t1 == OR_CT_INTEGER) { int32_t i; memcpy(&i, a, 4); x = i; } else memcpy(&x, a, 8);
    if (t2 == OR_CT_INTEGER) { int32_t i; memcpy(&i, b, 4); y = i; } else memcpy(&y, b, 8);
    return x == y;
}

int or_join_out_width(const or_join_desc* d)
{
    int w = 0;
    for (int j = 0; j < d->nproj; ++j) {
        const or_schema* s = d->proj_side[j] == OR_BUILD ? &d->build : &d->probe;
        if (d->proj_attr[j] < 0 || d->proj_attr[j] >= s->ncols)
            return -1;
        w += s->width[d->proj_attr[j]];
    }
    return w;
}

typedef struct {
    or_schema sbuild;        /* keys, then build-projected columns */
    or_schema out;
    or_hasher bh, ph;
    unsigned boff, bsize, poff, psize;
} join_plan;

/* operators/join.cpp:224-326 and 367-518. */
static int join_prepare(const or_join_desc* d, join_plan* p)
{
    int types[OR_MAX_COLS], widths[OR_MAX_COLS], n = 0;
    if (d->nkeys < 1 || d->nkeys > 8 || d->nproj < 0 || d->nproj > OR_MAX_COLS)
        return -1;
    for (int i = 0; i < d->nkeys; ++i) {
        if (d->buildjattr[i] >= d->build.ncols || d->probejattr[i] >= d->probe.ncols)
            return -1;
        types[n] = d->build.type[d->buildjattr[i]];
        widths[n++] = d->build.width[d->buildjattr[i]];
    }
    for (int j = 0; j < d->nproj; ++j) {
        if (d->proj_side[j] != OR_BUILD)
            continue;
        if (n >= OR_MAX_COLS)
            return -1;
        types[n] = d->build.type[d->proj_attr[j]];
        widths[n++] = d->build.width[d->proj_attr[j]];
    }
    if (or_schema_init(&p->sbuild, n, types, widths))
        return -1;
    n = 0;
    for (int j = 0; j < d->nproj; ++j) {
        const or_schema* s = d->proj_side[j] == OR_BUILD ? &d->build : &d->probe;
        if (d->proj_attr[j] < 0 || d->proj_attr[j] >= s->ncols)
            return -1;
        types[n] = s->type[d->proj_attr[j]];
        widths[n++] = s->width[d->proj_attr[j]];
    }
    if (or_schema_init(&p->out, n, types, widths))
        return -1;

    /* join.cpp:406-447: bytes hashes the maximal contiguous key range; every
     * other fn hashes the first key column only. */
    unsigned last = 0;
    if (d->hash_fn == OR_FN_BYTES)
        last = or_find_max_contiguous(d->buildjattr, d->probejattr, (unsigned)d->nkeys);
    int fields[2][2] = {{d->buildjattr[0], d->buildjattr[last]},
                        {d->probejattr[0], d->probejattr[last]}};
    const or_schema* sch[2] = {&d->build, &d->probe};
    or_hasher* hs[2] = {&p->bh, &p->ph};
    unsigned* offs[2] = {&p->boff, &p->poff};
    unsigned* sizes[2] = {&p->bsize, &p->psize};
    for (int side = 0; side < 2; ++side) {
        if (or_hasher_init(hs[side], d->hash_fn, d->buckets, d->range_min, d->range_max,
                           d->hash_offset, d->skipbits))
            return -1;
        int fmin = fields[side][0], fmax = fields[side][1];
        if (d->hash_fn == OR_FN_ALWAYSZERO) {
            *offs[side] = 0;
            *sizes[side] = 0;
            continue;
        }
        if (d->hash_fn != OR_FN_BYTES && d->hash_fn != OR_FN_TPCHQ1MAGIC) {
            int t = sch[side]->type[fmin];    /* hash.cpp:164-180 */
            if (fmin != fmax || !(t == OR_CT_INTEGER || t == OR_CT_LONG || t == OR_CT_DATE))
                return -1;
        }
        *offs[side] = (unsigned)sch[side]->offset[fmin];
        unsigned sz = 0;
        for (int i = fmin; i <= fmax; ++i)
            sz += (unsigned)sch[side]->width[i];
        *sizes[side] = sz;
    }
    if (or_hasher_buckets(&p->bh) != or_hasher_buckets(&p->ph))
        return -1;
    return 0;
}

static int keys_match(const or_join_desc* d, const join_plan* p, const unsigned char* stored,
                      const unsigned char* ptup)
{
    for (int i = 0; i < d->nkeys; ++i) {
        int pa = d->probejattr[i];
        int e = key_equal(p->sbuild.type[i], stored + p->sbuild.offset[i], d->probe.type[pa],
                          ptup + d->probe.offset[pa]);
        if (e != 1)
            return 0;
    }
    return 1;
}

/* join.cpp:648-679: build columns come from the stored tuple after the keys,
 * probe columns from the probe tuple. */
static void construct_output(const or_join_desc* d, const join_plan* p,
                             const unsigned char* stored, const unsigned char* ptup,
                             unsigned char* out)
{
    int battr = d->nkeys;
    for (int j = 0; j < d->nproj; ++j) {
        const void* src;
        if (d->proj_side[j] == OR_BUILD) {
            src = stored + p->sbuild.offset[battr];
            battr++;
        } else {
            src = ptup + d->probe.offset[d->proj_attr[j]];
        }
        write_col(out + p->out.offset[j], p->out.type[j], p->out.width[j], src);
    }
}

int or_hashjoin(const or_join_desc* d, const void* build, uint64_t nbuild,
                const void* probe, uint64_t nprobe, void* out, uint64_t out_cap,
                or_join_stats* st)
{
    join_plan p;
    if (join_prepare(d, &p))
        return -1;
    unsigned ow = (unsigned)p.out.size;
    if (ow == 0 || d->buffsize < ow || d->tuplesperbucket == 0)
        return -1;

    htable t;
    unsigned nb = or_hasher_buckets(&p.bh);
    if (ht_init(&t, nb, d->tuplesperbucket * (unsigned)p.sbuild.size, (unsigned)p.sbuild.size))
        return -1;

    /* buildFromPage, join.cpp:838-884. */
    const unsigned char* B = (const unsigned char*)build;
    for (uint64_t i = 0; i < nbuild; ++i) {
        const unsigned char* tup = B + i * (uint64_t)d->build.size;
        unsigned h = or_hasher_hash(&p.bh, tup + p.boff, p.bsize);
        unsigned char* target = (unsigned char*)ht_allocate(&t, h);
        if (!target) {
            ht_destroy(&t);
            return -1;
        }
        int col = 0;
        for (; col < d->nkeys; ++col) {
            int a = d->buildjattr[col];
            write_col(target + p.sbuild.offset[col], p.sbuild.type[col], p.sbuild.width[col],
                      tup + d->build.offset[a]);
        }
        for (int j = 0; j < d->nproj; ++j) {
            if (d->proj_side[j] != OR_BUILD)
                continue;
            int a = d->proj_attr[j];
            write_col(target + p.sbuild.offset[col], p.sbuild.type[col], p.sbuild.width[col],
                      tup + d->build.offset[a]);
            col++;
        }
    }

    /* scanStart positioning + getNext, join.cpp:619-757: pages of buffsize
     * bytes; a page is returned when it cannot hold another tuple, and the
     * iterator resumes on the next call. */
    const unsigned char* P = (const unsigned char*)probe;
    unsigned char* O = (unsigned char*)out;
    uint64_t produced = 0, pages = 0;
    uint64_t page_cap = d->buffsize / ow, in_page = 0;
    int rc = 0;
    ht_iter it;
    memset(&it, 0, sizeof(it));
    for (uint64_t i = 0; i < nprobe && rc == 0; ++i) {
        const unsigned char* ptup = P + i * (uint64_t)d->probe.size;
        ht_place(&t, &it, or_hasher_hash(&p.ph, ptup + p.poff, p.psize));
        const unsigned char* stored;
        while ((stored = (const unsigned char*)ht_next(&it))) {
            if (!keys_match(d, &p, stored, ptup))
                continue;
            if (produced >= out_cap) {
                rc = -2;
                break;
            }
            construct_output(d, &p, stored, ptup, O + produced * ow);
            produced++;
            if (++in_page == page_cap) {   /* !canStoreTuple -> Ready */
                pages++;
                in_page = 0;
            }
        }
    }
    pages++;                               /* the Finished page */
    if (st) {
        st->out_tuples = produced;
        st->pages = pages;
        st->spills = t.spills;
        st->out_tuple_size = ow;
    }
    ht_destroy(&t);
    return rc;
}

int or_nested_loop_join(const or_join_desc* d, const void* build, uint64_t nbuild,
                        const void* probe, uint64_t nprobe, void* out, uint64_t out_cap,
                        uint64_t* out_n)
{
    join_plan p;
    if (join_prepare(d, &p))
        return -1;
    unsigned ow = (unsigned)p.out.size;
    unsigned char* stored = (unsigned char*)malloc((size_t)p.sbuild.size + 1);
    const unsigned char* B = (const unsigned char*)build;
    const unsigned char* P = (const unsigned char*)probe;
    uint64_t n = 0;
    for (uint64_t j = 0; j < nprobe; ++j) {
        const unsigned char* ptup = P + j * (uint64_t)d->probe.size;
        for (uint64_t i = 0; i < nbuild; ++i) {
            const unsigned char* btup = B + i * (uint64_t)d->build.size;
            int col = 0;
            for (; col < d->nkeys; ++col)
                write_col(stored + p.sbuild.offset[col], p.sbuild.type[col],
                          p.sbuild.width[col], btup + d->build.offset[d->buildjattr[col]]);
            for (int k = 0; k < d->nproj; ++k)
                if (d->proj_side[k] == OR_BUILD) {
                    write_col(stored + p.sbuild.offset[col], p.sbuild.type[col],
                              p.sbuild.width[col], btup + d->build.offset[d->proj_attr[k]]);
                    col++;
                }
            if (!keys_match(d, &p, stored, ptup))
                continue;
            if (n >= out_cap) {
                free(stored);
                return -2;
            }
            construct_output(d, &p, stored, ptup, (unsigned char*)out + n * ow);
            n++;
        }
    }
    free(stored);
    *out_n = n;
    return 0;
}

/* -------------------------------------------------------------- partition */

/* operators/partition.cpp:166-200 (histogram), 288-321 (prefix across
 * threads per partition), 249-269 (scatter). */
int or_partition(const or_schema* s, int attr, int64_t range_min, int64_t range_max,
                 unsigned threads, const void* in, const uint64_t* counts,
                 void* out, uint64_t* part_sizes)
{
    if (threads == 0 || threads >= 128 || attr < 0 || attr >= s->ncols)
        return -1;
    int t = s->type[attr];
    if (!(t == OR_CT_INTEGER || t == OR_CT_LONG || t == OR_CT_DATE))
        return -1;
    or_hasher h;
    if (or_hasher_init(&h, OR_FN_EXACTRANGE, threads, range_min, range_max, 0, 0))
        return -1;
    unsigned B = or_hasher_buckets(&h);
    unsigned w = (unsigned)s->size, off = (unsigned)s->offset[attr], sz = (unsigned)s->width[attr];
    uint64_t* hist = (uint64_t*)calloc((size_t)threads * B, sizeof(uint64_t));
    uint64_t* idx = (uint64_t*)calloc((size_t)threads * B, sizeof(uint64_t));
    const unsigned char* src = (const unsigned char*)in;
    uint64_t base = 0;
    for (unsigned th = 0; th < threads; ++th) {
        for (uint64_t i = 0; i < counts[th]; ++i) {
            unsigned b = or_hasher_hash(&h, src + (base + i) * w + off, sz);
            if (b >= B) {
                free(hist);
                free(idx);
                return -1;
            }
            hist[(size_t)th * B + b]++;
        }
        base += counts[th];
    }
    /* idxstart[i][j] = idxstart[i-1][j] + hist[i-1][j]; partition j's size
     * is the last thread's start plus its count. */
    uint64_t pstart = 0;
    for (unsigned j = 0; j < B; ++j) {
        for (unsigned th = 1; th < threads; ++th)
            idx[(size_t)th * B + j] = idx[(size_t)(th - 1) * B + j] + hist[(size_t)(th - 1) * B + j];
        part_sizes[j] = idx[(size_t)(threads - 1) * B + j] + hist[(size_t)(threads - 1) * B + j];
        for (unsigned th = 0; th < threads; ++th)
            idx[(size_t)th * B + j] += pstart;
        pstart += part_sizes[j];
    }
    unsigned char* dst = (unsigned char*)out;
    base = 0;
    for (unsigned th = 0; th < threads; ++th) {
        for (uint64_t i = 0; i < counts[th]; ++i) {
            const unsigned char* tup = src + (base + i) * w;
            unsigned b = or_hasher_hash(&h, tup + off, sz);
            memcpy(dst + idx[(size_t)th * B + b] * w, tup, w);
            idx[(size_t)th * B + b]++;
        }
        base += counts[th];
    }
    free(hist);
    free(idx);
    return 0;
}

/* --------------------------------------------------------------- checksums */

static uint64_t splitmix_fin(uint64_t x)
{
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

uint64_t or_tuple_mix(const void* t, unsigned width)
{
    const unsigned char* b = (const unsigned char*)t;
    uint64_t h = 0x9e3779b97f4a7c15ull ^ width;
    for (unsigned o = 0; o < width; o += 8) {
        uint64_t c = 0;
        unsigned n = width - o < 8 ? width - o : 8;
        memcpy(&c, b + o, n);
        h = splitmix_fin(h ^ c) + 0x9e3779b97f4a7c15ull;
    }
    return splitmix_fin(h);
}

uint64_t or_checksum(const void* tuples, uint64_t n, unsigned width)
{
    const unsigned char* b = (const unsigned char*)tuples;
    uint64_t s = 0;
    for (uint64_t i = 0; i < n; ++i)
        s += or_tuple_mix(b + i * width, width);
    return s;
}

uint32_t or_xor32(const void* tuples, uint64_t n, unsigned width)
{
    const unsigned char* b = (const unsigned char*)tuples;
    uint32_t x = 0;
    for (uint64_t i = 0; i < n; ++i)
        for (unsigned k = 0; k + 4 <= width; k += 4) {
            uint32_t v;
            memcpy(&v, b + i * width + k, 4);
            x ^= v;
        }
    return x;
}
