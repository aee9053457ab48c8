// ref_harness.cpp -- drives the reference HashJoinOp and the aggregates
// over it, compiled from the reference tree by oracle/ref/Makefile, on
// in-memory inputs.
//
// Checker and CPU baseline only: tests/, smoke() and bench.py's cpu_baseline
// load the resulting oracle/_ref/libpythiaref.so; the product never does.
// The plan is wired in code, as the reference's unit tests wire theirs
// (unit_tests/queryhashjoin.cpp:116-202): MergeOp(threads = T) over a
// HashJoinOp whose build and probe children hand out the input as pages.
#include <pthread.h>
#include <stdint.h>
#include <string.h>
#include <time.h>

#include <string>
#include <vector>

#include "libconfig.h++"
#include "hash.h"
#include "operators/operators.h"
#include "query.h"

using namespace libconfig;

namespace {

// Serves an in-memory relation of packed tuples as buffsize-byte pages;
// threads take pages by fetch-and-add, as ParallelScanOp takes them from a
// table (operators/parallelscan.cpp:176-190).
class MemSourceOp : public ZeroInputOp {
  public:
    MemSourceOp() : data(0), n(0), cursor(0), per_page(1) {}

    void setup(const int* types, const int* widths, int ncols, const void* d, uint64_t count)
    {
        coltypes.assign(types, types + ncols);
        colwidths.assign(widths, widths + ncols);
        data = static_cast<const char*>(d);
        n = count;
    }

    virtual void init(Config& root, Setting& node)
    {
        Operator::init(root, node);
        for (size_t i = 0; i < coltypes.size(); ++i) {
            if (coltypes[i] == CT_CHAR)
                schema.add(CT_CHAR, colwidths[i]);
            else if (coltypes[i] == CT_DATE)   // DATE columns carry a format string (schema.cpp:62-74)
                schema.add(CT_DATE, std::string("%Y-%m-%d"));
            else
                schema.add(static_cast<ColumnType>(coltypes[i]));
        }
        const unsigned w = schema.getTupleSize();
        per_page = buffsize / w ? buffsize / w : 1;
        pages.assign(MAX_THREADS, Page(static_cast<void*>(0), 0, static_cast<void*>(0), w));
    }

    virtual ResultCode scanStart(unsigned short, Page*, Schema&)
    {
        return Ready;
    }

    virtual GetNextResultT getNext(unsigned short tid)
    {
        const unsigned w = schema.getTupleSize();
        const uint64_t start = __sync_fetch_and_add(&cursor, per_page);
        if (start >= n) {
            pages[tid] = Page(static_cast<void*>(0), 0, static_cast<void*>(0), w);
            return std::make_pair(Finished, &pages[tid]);
        }
        const uint64_t cnt = n - start < per_page ? n - start : per_page;
        pages[tid] = Page(const_cast<char*>(data) + start * w, cnt * w, NULL, w);
        return std::make_pair(Ready, &pages[tid]);
    }

    void rewind() { cursor = 0; }

  private:
    std::vector<int> coltypes, colwidths;
    const char* data;
    uint64_t n;
    volatile uint64_t cursor;
    uint64_t per_page;
    std::vector<Page> pages;
};

double now()
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

}  // namespace

extern "C" {

// Runs one hash join. Returns 0, -1 on a reference error, -2 when out_cap
// tuples are too few (out_n then holds the count).
int ref_hashjoin_run(int threads, int nbcols, const int* btypes, const int* bwidths, int npcols,
                     const int* ptypes, const int* pwidths, int nkeys, const int* bjattr,
                     const int* pjattr, int nproj, const char* const* proj, const char* hashfn,
                     int buckets, int tuplesperbucket, int buffsize, int range_min, int range_max,
                     int prefetch, const void* build, uint64_t nbuild, const void* probe, uint64_t nprobe,
                     void* out, uint64_t out_cap, uint64_t* out_n, double* t_build,
                     double* t_total)
{
    try {
        Query q;
        MemSourceOp src_build, src_probe;
        HashJoinOp join;
        MergeOp merge;
        src_build.setup(btypes, bwidths, nbcols, build, nbuild);
        src_probe.setup(ptypes, pwidths, npcols, probe, nprobe);

        Config cfg;
        cfg.getRoot().add("path", Setting::TypeString) = "./";
        cfg.getRoot().add("buffsize", Setting::TypeInt) = buffsize;
        Setting& sb = cfg.getRoot().add("srcbuild", Setting::TypeGroup);
        Setting& sp = cfg.getRoot().add("srcprobe", Setting::TypeGroup);
        Setting& jn = cfg.getRoot().add("join", Setting::TypeGroup);
        Setting& hn = jn.add("hash", Setting::TypeGroup);
        hn.add("fn", Setting::TypeString) = hashfn;
        hn.add("buckets", Setting::TypeInt) = buckets;
        Setting& rn = hn.add("range", Setting::TypeArray);
        rn.add(Setting::TypeInt) = range_min;
        rn.add(Setting::TypeInt) = range_max;
        Setting& tg = jn.add("threadgroups", Setting::TypeList);
        Setting& g0 = tg.add(Setting::TypeArray);
        for (int i = 0; i < threads; ++i)
            g0.add(Setting::TypeInt) = i;
        jn.add("tuplesperbucket", Setting::TypeInt) = tuplesperbucket;
        jn.add("allocpolicy", Setting::TypeString) = "local";
        if (prefetch >= 0) {   // HashJoinOp::init, join.cpp:386-404 (absent: no prefetch)
            jn.add("probeprefetchstep", Setting::TypeInt) = prefetch;
            jn.add("buildprefetchstep", Setting::TypeInt) = prefetch;
        }
        if (nkeys == 1) {
            jn.add("buildjattr", Setting::TypeInt) = bjattr[0];
            jn.add("probejattr", Setting::TypeInt) = pjattr[0];
        } else {
            Setting& bj = jn.add("buildjattr", Setting::TypeArray);
            Setting& pj = jn.add("probejattr", Setting::TypeArray);
            for (int i = 0; i < nkeys; ++i) {
                bj.add(Setting::TypeInt) = bjattr[i];
                pj.add(Setting::TypeInt) = pjattr[i];
            }
        }
        Setting& pn = jn.add("projection", Setting::TypeList);
        for (int i = 0; i < nproj; ++i)
            pn.add(Setting::TypeString) = proj[i];
        Setting& mn = cfg.getRoot().add("merge", Setting::TypeGroup);
        mn.add("threads", Setting::TypeInt) = threads;

        q.tree = &merge;
        merge.nextOp = &join;
        join.buildOp = &src_build;
        join.probeOp = &src_probe;
        src_build.init(cfg, sb);
        src_probe.init(cfg, sp);
        join.init(cfg, jn);
        merge.init(cfg, mn);

        const unsigned ow = q.getOutSchema().getTupleSize();
        q.threadInit();
        const double t0 = now();
        if (q.scanStart() == Operator::Error)
            return -1;
        const double t1 = now();
        uint64_t produced = 0;
        char* o = static_cast<char*>(out);
        Operator::GetNextResultT res;
        res.first = Operator::Ready;
        while (res.first == Operator::Ready) {
            res = q.getNext();
            if (res.first == Operator::Error)
                return -1;
            Operator::Page::Iterator it = res.second->createIterator();
            void* tup;
            while ((tup = it.next())) {
                if (produced < out_cap)
                    memcpy(o + produced * ow, tup, ow);
                produced++;
            }
        }
        if (q.scanStop() == Operator::Error)
            return -1;
        const double t2 = now();
        q.threadClose();
        q.destroynofree();
        *out_n = produced;
        *t_build = t1 - t0;
        *t_total = t2 - t0;
        return produced > out_cap ? -2 : 0;
    } catch (...) {
        return -1;
    }
}

// The headline plan of the CPU baseline: MergeOp(threads) -> ConsumeOp ->
// HashJoinOp (operators/consume.cpp:61-110): every worker thread drains its
// own join output and XORs its 32-bit words, so no row is copied and the
// timed region is the join. *xor32 gets the XOR over all threads (the
// order-independent check of the output), *t_build the scanStart time.
int ref_hashjoin_consume(int threads, int nbcols, const int* btypes, const int* bwidths, int npcols,
                         const int* ptypes, const int* pwidths, int nkeys, const int* bjattr, const int* pjattr,
                         int nproj, const char* const* proj, const char* hashfn, int buckets,
                         int tuplesperbucket, int buffsize, int prefetch, const void* build, uint64_t nbuild,
                         const void* probe, uint64_t nprobe, uint32_t* xor32, double* t_build, double* t_total)
{
    try {
        Query q;
        MemSourceOp src_build, src_probe;
        HashJoinOp join;
        ConsumeOp consume;
        MergeOp merge;
        src_build.setup(btypes, bwidths, nbcols, build, nbuild);
        src_probe.setup(ptypes, pwidths, npcols, probe, nprobe);

        Config cfg;
        cfg.getRoot().add("path", Setting::TypeString) = "./";
        cfg.getRoot().add("buffsize", Setting::TypeInt) = buffsize;
        Setting& sb = cfg.getRoot().add("srcbuild", Setting::TypeGroup);
        Setting& sp = cfg.getRoot().add("srcprobe", Setting::TypeGroup);
        Setting& jn = cfg.getRoot().add("join", Setting::TypeGroup);
        Setting& hn = jn.add("hash", Setting::TypeGroup);
        hn.add("fn", Setting::TypeString) = hashfn;
        hn.add("buckets", Setting::TypeInt) = buckets;
        Setting& tg = jn.add("threadgroups", Setting::TypeList);
        Setting& g0 = tg.add(Setting::TypeArray);
        for (int i = 0; i < threads; ++i)
            g0.add(Setting::TypeInt) = i;
        jn.add("tuplesperbucket", Setting::TypeInt) = tuplesperbucket;
        jn.add("allocpolicy", Setting::TypeString) = "local";
        if (prefetch >= 0) {
            jn.add("probeprefetchstep", Setting::TypeInt) = prefetch;
            jn.add("buildprefetchstep", Setting::TypeInt) = prefetch;
        }
        if (nkeys == 1) {
            jn.add("buildjattr", Setting::TypeInt) = bjattr[0];
            jn.add("probejattr", Setting::TypeInt) = pjattr[0];
        } else {
            Setting& bj = jn.add("buildjattr", Setting::TypeArray);
            Setting& pj = jn.add("probejattr", Setting::TypeArray);
            for (int i = 0; i < nkeys; ++i) {
                bj.add(Setting::TypeInt) = bjattr[i];
                pj.add(Setting::TypeInt) = pjattr[i];
            }
        }
        Setting& pn = jn.add("projection", Setting::TypeList);
        for (int i = 0; i < nproj; ++i)
            pn.add(Setting::TypeString) = proj[i];
        Setting& cn = cfg.getRoot().add("consume", Setting::TypeGroup);
        Setting& mn = cfg.getRoot().add("merge", Setting::TypeGroup);
        mn.add("threads", Setting::TypeInt) = threads;

        q.tree = &merge;
        merge.nextOp = &consume;
        consume.nextOp = &join;
        join.buildOp = &src_build;
        join.probeOp = &src_probe;
        src_build.init(cfg, sb);
        src_probe.init(cfg, sp);
        join.init(cfg, jn);
        consume.init(cfg, cn);
        merge.init(cfg, mn);

        q.threadInit();
        const double t0 = now();
        if (q.scanStart() == Operator::Error)
            return -1;
        const double t1 = now();
        uint32_t x = 0;
        Operator::GetNextResultT res;
        res.first = Operator::Ready;
        while (res.first == Operator::Ready) {
            res = q.getNext();
            if (res.first == Operator::Error)
                return -1;
            Operator::Page::Iterator it = res.second->createIterator();
            void* tup;
            while ((tup = it.next())) {
                long long v;
                memcpy(&v, tup, sizeof(v));
                x ^= (uint32_t)v;
            }
        }
        if (q.scanStop() == Operator::Error)
            return -1;
        const double t2 = now();
        q.threadClose();
        q.destroynofree();
        *xor32 = x;
        *t_build = t1 - t0;
        *t_total = t2 - t0;
        return 0;
    } catch (...) {
        return -1;
    }
}

// Runs AggregateSum (op 0, over column sumfield) or AggregateCount (op 1),
// grouped by nfields columns (0: the whole input), in the reference's global
// mode: MergeOp(threads) over one aggregate whose hash table all threads
// share (operators/genericaggregate.cpp:130-165). Grouping hashes with
// `hashfn` over the first grouping column, or with "bytes" over the range
// of grouping columns. Returns 0, -1 on a reference error, -2 when out_cap tuples are too
// few (out_n then holds the count).
int ref_aggregate_run(int threads, int op, int ncols, const int* types, const int* widths, int nfields,
                      const int* fields, int sumfield, const char* hashfn, int buckets, const void* data,
                      uint64_t n, void* out, uint64_t out_cap, uint64_t* out_n, unsigned* out_width,
                      double* t_total)
{
    try {
        Query q;
        MemSourceOp src;
        AggregateSum asum;
        AggregateCount acount;
        GenericAggregate* agg = op == 0 ? static_cast<GenericAggregate*>(&asum) : static_cast<GenericAggregate*>(&acount);
        MergeOp merge;
        src.setup(types, widths, ncols, data, n);

        Config cfg;
        cfg.getRoot().add("path", Setting::TypeString) = "./";
        cfg.getRoot().add("buffsize", Setting::TypeInt) = 1 << 20;
        Setting& sn = cfg.getRoot().add("src", Setting::TypeGroup);
        Setting& an = cfg.getRoot().add("agg", Setting::TypeGroup);
        Setting& fl = an.add("fields", Setting::TypeArray);
        for (int i = 0; i < nfields; ++i)
            fl.add(Setting::TypeInt) = fields[i];
        if (op == 0)
            an.add("sumfield", Setting::TypeInt) = sumfield;
        if (nfields > 0) {
            Setting& hn = an.add("hash", Setting::TypeGroup);
            hn.add("fn", Setting::TypeString) = hashfn;
            hn.add("buckets", Setting::TypeInt) = buckets;
            if (strcmp(hashfn, "bytes") != 0) {
                hn.add("field", Setting::TypeInt) = fields[0];
            } else {
                Setting& fr = hn.add("fieldrange", Setting::TypeArray);
                fr.add(Setting::TypeInt) = fields[0];
                fr.add(Setting::TypeInt) = fields[nfields - 1];
            }
        }
        an.add("global", Setting::TypeBoolean) = true;
        an.add("threads", Setting::TypeInt) = threads;
        Setting& mn = cfg.getRoot().add("merge", Setting::TypeGroup);
        mn.add("threads", Setting::TypeInt) = threads;

        q.tree = &merge;
        merge.nextOp = agg;
        agg->nextOp = &src;
        src.init(cfg, sn);
        agg->init(cfg, an);
        merge.init(cfg, mn);

        const unsigned ow = q.getOutSchema().getTupleSize();
        *out_width = ow;
        q.threadInit();
        const double t0 = now();
        if (q.scanStart() == Operator::Error)
            return -1;
        uint64_t produced = 0;
        char* o = static_cast<char*>(out);
        Operator::GetNextResultT res;
        res.first = Operator::Ready;
        while (res.first == Operator::Ready) {
            res = q.getNext();
            if (res.first == Operator::Error)
                return -1;
            Operator::Page::Iterator it = res.second->createIterator();
            void* tup;
            while ((tup = it.next())) {
                if (produced < out_cap)
                    memcpy(o + produced * ow, tup, ow);
                produced++;
            }
        }
        if (q.scanStop() == Operator::Error)
            return -1;
        *t_total = now() - t0;
        q.threadClose();
        q.destroynofree();
        *out_n = produced;
        return produced > out_cap ? -2 : 0;
    } catch (...) {
        return -1;
    }
}

// TupleHasher::create over a one-column tuple of the given type, then hash.
unsigned ref_tuple_hash(const char* fn, int buckets, int coltype, const void* tuple, int range_min,
                        int range_max, int offset, int skipbits)
{
    Schema s;
    s.add(static_cast<ColumnType>(coltype));
    Config cfg;
    Setting& hn = cfg.getRoot().add("hash", Setting::TypeGroup);
    hn.add("fn", Setting::TypeString) = fn;
    hn.add("buckets", Setting::TypeInt) = buckets;
    hn.add("field", Setting::TypeInt) = 0;
    Setting& rn = hn.add("range", Setting::TypeArray);
    rn.add(Setting::TypeInt) = range_min;
    rn.add(Setting::TypeInt) = range_max;
    hn.add("offset", Setting::TypeInt) = offset;
    hn.add("skipbits", Setting::TypeInt) = skipbits;
    TupleHasher h = TupleHasher::create(s, hn);
    unsigned v = h.hash(const_cast<void*>(tuple));
    h.destroy();
    return v;
}

unsigned ref_tuple_hash_bytes(int buckets, const void* bytes, int size)
{
    // ByteHasher over `size` bytes: a schema of LONG columns covering them,
    // hashed with fieldrange over all of them (hash.cpp:123-150).
    Schema s;
    const int ncols = size / 8;
    for (int i = 0; i < ncols; ++i)
        s.add(CT_LONG);
    Config cfg;
    Setting& hn = cfg.getRoot().add("hash", Setting::TypeGroup);
    hn.add("fn", Setting::TypeString) = "bytes";
    hn.add("buckets", Setting::TypeInt) = buckets;
    Setting& fr = hn.add("fieldrange", Setting::TypeArray);
    fr.add(Setting::TypeInt) = 0;
    fr.add(Setting::TypeInt) = ncols - 1;
    TupleHasher h = TupleHasher::create(s, hn);
    unsigned v = h.hash(const_cast<void*>(bytes));
    h.destroy();
    return v;
}

}  // extern "C"
