// operators.cpp -- HashJoinOp, PartitionOp, MergeOp, the scans and the plan
// builder of the host operator layer, calling libpjoin's C ABI.
#include <fcntl.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <functional>
#include <mutex>
#include <thread>

#include "pjh.h"

namespace pjh {

static std::string last_pj_error() { return pj_last_error(); }

void Operator::init(const Setting& root, const Setting& node)
{
    (void)node;
    buffsize = (unsigned)root["buffsize"].asInt();   // operators/base.cpp:37-40
}

// ---------------------------------------------------------------- scans

// Drains op on thread tid into `to`. An Error from it, or an exception it
// throws (whose message becomes the thread's library error), gives Error;
// either way the caller goes on to its group's barrier.
static Operator::ResultCode drain(Operator* op, unsigned short tid, std::vector<char>& to)
{
    Operator::ResultCode rc = Operator::Ready;
    try {
        if (op->scanStart(tid) == Operator::Error)
            rc = Operator::Error;
        Operator::GetNextResultT r(Operator::Ready, nullptr);
        while (rc != Operator::Error && r.first == Operator::Ready) {
            r = op->getNext(tid);
            if (r.first == Operator::Error)
                rc = Operator::Error;
            else if (r.second && r.second->getUsedSpace())
                to.insert(to.end(), r.second->data(), r.second->data() + r.second->getUsedSpace());
        }
        if (op->scanStop(tid) == Operator::Error)
            rc = Operator::Error;
    } catch (const std::exception& e) {
        pj_set_error(e.what());
        rc = Operator::Error;
    }
    return rc;
}


void MemScanOp::init(const Setting& root, const Setting& node)
{
    Operator::init(root, node);
    const Setting& sc = node["schema"];
    for (int i = 0; i < sc.getLength(); ++i) {
        unsigned w;
        ColumnType t = Schema::parse(sc[i].asString(), &w);
        schema.add(t, w);
    }
    const unsigned w = schema.getTupleSize();
    per_page = buffsize / w ? buffsize / w : 1;
    pages.assign(MAX_THREADS, Page());
    served.assign(MAX_THREADS, 0);
    if (node.exists("errorthread")) {
        error_thread = (int)node["errorthread"].asInt();
        error_after = node.exists("errorafter") ? (unsigned)node["errorafter"].asInt() : 0;
    }
}

void MemScanOp::threadInit(unsigned short tid)
{
    (void)tid;
}

Operator::GetNextResultT MemScanOp::getNext(unsigned short tid)
{
    const unsigned w = schema.getTupleSize();
    if ((int)tid == error_thread && served[tid]++ >= error_after)
        return {Error, nullptr};
    const uint64_t start = __sync_fetch_and_add(&cursor, per_page);
    if (start >= count) {
        pages[tid] = Page(nullptr, 0, w, true);
        return {Finished, &pages[tid]};
    }
    const uint64_t n = std::min<uint64_t>(per_page, count - start);
    pages[tid] = Page(const_cast<char*>(data) + start * w, n * w, w, true);
    return {Ready, &pages[tid]};
}

BinaryScanOp::~BinaryScanOp()
{
    if (map)
        munmap(map, mapsize);
}

void BinaryScanOp::init(const Setting& root, const Setting& node)
{
    MemScanOp::init(root, node);
    std::string path = root.exists("path") ? root["path"].asString() : "";
    std::string file = path + node["file"].asString();
    int fd = open(file.c_str(), O_RDONLY);
    if (fd < 0)
        throw InvalidParameter("cannot open " + file);
    struct stat stt;
    fstat(fd, &stt);
    mapsize = (uint64_t)stt.st_size;
    if (mapsize) {
        map = mmap(nullptr, mapsize, PROT_READ, MAP_PRIVATE, fd, 0);
        if (map == MAP_FAILED) {
            map = nullptr;
            close(fd);
            throw InvalidParameter("cannot map " + file);
        }
    }
    close(fd);
    setData(map, mapsize / getOutSchema().getTupleSize());
}

// Prepends the thread id as an INTEGER column
// (operators/threadidprepend.cpp).
class ThreadIdPrependOp : public Operator {
  public:
    Operator* nextOp = nullptr;
    void init(const Setting& root, const Setting& node) override
    {
        Operator::init(root, node);
        schema.add(CT_INTEGER);
        Schema& in = nextOp->getOutSchema();
        for (unsigned c = 0; c < in.columns(); ++c)
            schema.add(in.getColumnType(c), in.getColumnWidth(c));
        out.resize(MAX_THREADS);
        inpage.resize(MAX_THREADS, nullptr);
        inpos.assign(MAX_THREADS, 0);
        prev.assign(MAX_THREADS, Ready);
    }
    void threadInit(unsigned short tid) override { out[tid] = Page(buffsize, schema.getTupleSize()); }
    ResultCode scanStart(unsigned short tid) override { return nextOp->scanStart(tid); }
    ResultCode scanStop(unsigned short tid) override { return nextOp->scanStop(tid); }
    std::vector<Operator*> children() override { return {nextOp}; }
    GetNextResultT getNext(unsigned short tid) override
    {
        Page& o = out[tid];
        o.clear();
        const unsigned iw = nextOp->getOutSchema().getTupleSize();
        for (;;) {
            Page* in = inpage[tid];
            while (in && inpos[tid] < in->getNumTuples()) {
                if (!o.canStoreTuple())
                    return {Ready, &o};
                char* t = static_cast<char*>(o.allocateTuple());
                int32_t id = tid;
                memcpy(t, &id, 4);
                memcpy(t + 4, in->data() + inpos[tid] * iw, iw);
                inpos[tid]++;
            }
            if (prev[tid] == Finished)
                return {Finished, &o};
            GetNextResultT r = nextOp->getNext(tid);
            if (r.first == Error)
                return {Error, &o};
            prev[tid] = r.first;
            inpage[tid] = r.second;
            inpos[tid] = 0;
        }
    }

  private:
    std::vector<Page> out;
    std::vector<Page*> inpage;
    std::vector<uint64_t> inpos;
    std::vector<ResultCode> prev;
};

// ---------------------------------------------------------------- HashJoinOp

namespace {

// "B$i" / "P$j" (operators/join.cpp:142-183).
std::pair<int, unsigned> parse_projection(const std::string& s)
{
    size_t l = s.find('$');
    if (l == std::string::npos || l == 0)
        throw InvalidParameter("projection " + s);
    char c = s.at(l - 1);
    std::string rest = s.substr(l + 1);
    if (rest.empty() || rest.find_first_not_of("0123456789") != std::string::npos)
        throw InvalidParameter("projection " + s);
    unsigned a = (unsigned)atoi(rest.c_str());
    if (c == 'B')
        return {PJ_SIDE_BUILD, a};
    if (c == 'P')
        return {PJ_SIDE_PROBE, a};
    throw InvalidParameter("projection " + s);
}

std::vector<unsigned short> join_attrs(const Setting& n)
{
    std::vector<unsigned short> r;
    if (n.isAggregate())
        for (int i = 0; i < n.getLength(); ++i)
            r.push_back((unsigned short)n[i].asInt());
    else
        r.push_back((unsigned short)n.asInt());
    return r;
}

}  // namespace

bool HashJoinOp::isLeftThread(unsigned short tid) const
{
    if (leftthreads.empty())
        return true;
    return std::find(leftthreads.begin(), leftthreads.end(), tid) != leftthreads.end();
}

// JoinOp::init + HashJoinOp::init (operators/join.cpp:224-326, 367-518).
void HashJoinOp::init(const Setting& root, const Setting& node)
{
    Operator::init(root, node);
    pj_join_desc d;
    memset(&d, 0, sizeof(d));
    Schema& bs = buildOp->getOutSchema();
    Schema& ps = probeOp->getOutSchema();
    bs.to_c(&d.build);
    ps.to_c(&d.probe);

    std::vector<unsigned short> bj = join_attrs(node["buildjattr"]);
    std::vector<unsigned short> pj = join_attrs(node["probejattr"]);
    if (bj.size() != pj.size() || bj.empty() || bj.size() > PJ_MAX_KEYS)
        throw InvalidParameter("buildjattr and probejattr differ in length");
    d.nkeys = (uint32_t)bj.size();
    for (size_t i = 0; i < bj.size(); ++i) {
        if (bj[i] >= bs.columns() || pj[i] >= ps.columns())
            throw InvalidParameter("join attribute out of range");
        d.buildjattr[i] = bj[i];
        d.probejattr[i] = pj[i];
    }
    const Setting& proj = node["projection"];
    if (!proj.isAggregate() || proj.getLength() == 0 || proj.getLength() > PJ_MAX_COLS)
        throw InvalidParameter("projection");
    d.nproj = (uint32_t)proj.getLength();
    for (int i = 0; i < proj.getLength(); ++i) {
        auto pr = parse_projection(proj[i].asString());
        Schema& s = pr.first == PJ_SIDE_BUILD ? bs : ps;
        if (pr.second >= s.columns())
            throw InvalidParameter("projection " + proj[i].asString() + " out of range");
        d.proj_side[i] = (uint8_t)pr.first;
        d.proj_attr[i] = (uint16_t)pr.second;
        schema.add(s.getColumnType(pr.second), s.getColumnWidth(pr.second));
    }

    // hash: the reference's checks (hash.cpp:101-248 through
    // join.cpp:406-447). A value hasher takes the first key column, which
    // must be INTEGER, LONG or DATE.
    hash = make_hashspec(node["hash"]);
    if (hash.fn != "bytes" && hash.fn != "alwayszero" && hash.fn != "tpchq1magic") {
        for (int side = 0; side < 2; ++side) {
            ColumnType t = (side ? ps : bs).getColumnType(side ? pj[0] : bj[0]);
            if (!(t == CT_INTEGER || t == CT_LONG || t == CT_DATE))
                throw IllegalSchemaDeclarationException("value hash on a non-numeric key");
        }
    }
    if (hash.fn == "bytes") {
        // ByteHasher compares raw key bytes, so the reference joins only keys
        // of equal byte width; a plan mixing widths would give a different
        // result here.
        for (size_t i = 0; i < bj.size(); ++i)
            if (bs.getColumnWidth(bj[i]) != ps.getColumnWidth(pj[i]))
                throw InvalidParameter("bytes hash over key columns of different widths");
    }
    const int64_t tpb = node["tuplesperbucket"].asInt();
    if (tpb <= 0)
        throw InvalidParameter("tuplesperbucket");
    tuplesperbucket = (unsigned)tpb;
    std::string policy = node["allocpolicy"].asString();
    if (policy != "local" && policy != "striped")
        throw InvalidParameter("allocpolicy " + policy);

    // thread groups (join.cpp:264-304)
    group_of.assign(MAX_THREADS, -1);
    const Setting& tg = node["threadgroups"];
    for (int g = 0; g < tg.getLength(); ++g) {
        const Setting& grp = tg[g];
        if (grp.getLength() == 0)
            throw InvalidParameter("empty thread group");
        for (int j = 0; j < grp.getLength(); ++j) {
            int64_t t = grp[j].asInt();
            if (t < 0 || t >= MAX_THREADS || group_of[t] != -1)
                throw InvalidParameter("thread " + std::to_string(t) + " in two groups");
            group_of[t] = g;
        }
        leader_of_group.push_back((int)grp[0].asInt());
        groupsize.push_back(grp.getLength());
        barriers.emplace_back(new Barrier((unsigned)grp.getLength()));
    }
    if (node.exists("leftthreads")) {
        const Setting& lt = node["leftthreads"];
        for (int j = 0; j < lt.getLength(); ++j)
            leftthreads.push_back((unsigned short)lt[j].asInt());
    }
    if (node.exists("partitioner")) {
        std::string p = node["partitioner"].asString();
        d.partitioner = p == "learned" ? PJ_PART_LEARNED : p == "auto" ? PJ_PART_AUTO : PJ_PART_RADIX;
    }
    d.radix_bits = node.exists("radixbits") ? (int32_t)node["radixbits"].asInt() : -1;
    if (node.exists("probebatch"))
        probe_batch_tuples = (uint64_t)node["probebatch"].asInt();

    desc_ = d;
    // one device hash table per thread group, as the reference keeps one
    // HashTable per group
    for (size_t g = 0; g < leader_of_group.size(); ++g) {
        pj_join_t h = nullptr;
        int rc = pj_join_create(&d, &h);
        if (rc == PJ_ERR_UNSUPPORTED)
            throw IllegalSchemaDeclarationException(last_pj_error());
        if (rc)
            throw InvalidParameter(last_pj_error());
        handles.push_back(h);
    }
    state.resize(MAX_THREADS);
    build_rc.assign(leader_of_group.size(), PJ_OK);
    build_failed.reset(leader_of_group.size());
}

bool HashJoinOp::oneGroupOf(unsigned n) const
{
    if (groupsize.size() != 1 || (unsigned)groupsize[0] != n || !leftthreads.empty())
        return false;
    for (unsigned t = 0; t < n; ++t)
        if (group_of[t] != 0)
            return false;
    return true;
}

// HashJoinOp::threadInit (join.cpp:520-543): per-thread state and output page.
void HashJoinOp::threadInit(unsigned short tid)
{
    if (group_of.at(tid) < 0)
        throw InvalidParameter("thread " + std::to_string(tid) + " is in no thread group");
    state[tid].reset(new ThreadState());
    state[tid]->out = Page(buffsize, schema.getTupleSize());
}

// HashJoinOp::scanStart (join.cpp:548-637): the build. Every build thread
// stages its pages; after the group barrier the group leader uploads and
// partitions the build side on the device; a second barrier releases the
// group to probe.
Operator::ResultCode HashJoinOp::scanStart(unsigned short tid)
{
    const int g = group_of.at(tid);
    pj_join_t h = handles.at(g);
    ResultCode rc = Ready;
    if (isLeftThread(tid)) {
        bool failed = false;
        try {   // a child's exception fails the build as its Error does
            failed = buildOp->scanStart(tid) == Error;
            GetNextResultT r(Ready, nullptr);
            while (!failed && r.first == Ready) {
                r = buildOp->getNext(tid);
                if (r.first == Error) {
                    failed = true;
                    break;
                }
                if (r.second && r.second->getNumTuples())
                    if (pj_join_build_append(h, r.second->data(), r.second->getNumTuples()))
                        failed = true;
            }
            if (!failed && buildOp->scanStop(tid) == Error)
                failed = true;
        } catch (const std::exception& e) {
            pj_set_error(e.what());
            failed = true;
        }
        if (failed)
            build_failed.raise(g);
    }
    try {
        if (probeOp->scanStart(tid) == Error)
            rc = Error;
    } catch (const std::exception& e) {
        pj_set_error(e.what());
        rc = Error;
    }
    barriers[g]->arrive();   // every build page of the group is staged, or a build thread failed
    if (tid == leader_of_group[g])
        build_rc[g] = build_failed.raised(g) ? PJ_ERR_STATE : pj_join_build_finish(h);
    barriers[g]->arrive();
    if (rc == Error || build_failed.raised(g) || build_rc[g] != PJ_OK)
        return Error;
    // position on the first probe page; an empty probe finishes here
    // (join.cpp:619-633)
    ThreadState& st = *state[tid];
    GetNextResultT r = probeOp->getNext(tid);
    if (r.first == Error)
        return Error;
    if (r.second && r.second->getNumTuples())
        st.batch.insert(st.batch.end(), r.second->data(), r.second->data() + r.second->getUsedSpace());
    st.probe_done = r.first == Finished;
    if (st.probe_done && st.batch.empty())
        return Finished;
    return Ready;
}

// Fills the thread's output page from the cursor, probing a new batch when
// the cursor is drained (join.cpp:681-800).
Operator::ResultCode HashJoinOp::fillPage(unsigned short tid, ThreadState& st)
{
    pj_join_t h = handles.at(group_of[tid]);
    const unsigned ow = schema.getTupleSize();
    const unsigned pw = probeOp->getOutSchema().getTupleSize();
    Page& out = st.out;
    for (;;) {
        uint64_t room = (out.maxsize() - out.getUsedSpace()) / ow;
        if (st.cur.pending) {
            uint64_t got = 0;
            if (pj_join_probe(h, nullptr, 0, &st.cur, out.data() + out.getUsedSpace(), room, &got))
                throw QueryExecutionError(last_pj_error());
            for (uint64_t i = 0; i < got; ++i)
                out.allocateTuple();
            if (!out.canStoreTuple())
                return Ready;   // page full: resume from the cursor next call
            continue;
        }
        if (st.batch.empty() && st.probe_done)
            return Finished;
        while (!st.probe_done && st.batch.size() / pw < probe_batch_tuples) {
            GetNextResultT r = probeOp->getNext(tid);
            if (r.first == Error)
                throw QueryExecutionError("probe side returned Error");   // join.cpp:787-789
            if (r.second && r.second->getNumTuples())
                st.batch.insert(st.batch.end(), r.second->data(), r.second->data() + r.second->getUsedSpace());
            st.probe_done = r.first == Finished;
        }
        if (st.batch.empty())
            continue;
        uint64_t got = 0;
        if (pj_join_probe(h, st.batch.data(), st.batch.size() / pw, &st.cur, out.data() + out.getUsedSpace(), room,
                          &got))
            throw QueryExecutionError(last_pj_error());
        st.batch.clear();
        for (uint64_t i = 0; i < got; ++i)
            out.allocateTuple();
        if (!out.canStoreTuple())
            return Ready;
    }
}

Operator::GetNextResultT HashJoinOp::getNext(unsigned short tid)
{
    ThreadState& st = *state.at(tid);
    st.out.clear();
    ResultCode rc = fillPage(tid, st);
    if (rc == Ready && !st.cur.pending && st.batch.empty() && st.probe_done)
        rc = Finished;
    return {rc, &st.out};
}

Operator::ResultCode HashJoinOp::scanStop(unsigned short tid)
{
    return probeOp->scanStop(tid);
}

// HashJoinOp::threadClose (join.cpp:802-826): barrier, then the leader drops
// the group's table.
void HashJoinOp::threadClose(unsigned short tid)
{
    const int g = group_of.at(tid);
    if (state[tid]) {
        pj_cursor_release(&state[tid]->cur);
        state[tid].reset();
    }
    barriers[g]->arrive();
    if (tid == leader_of_group[g])
        pj_join_reset(handles[g]);
    barriers[g]->arrive();
}

void HashJoinOp::destroy()
{
    for (pj_join_t h : handles)
        pj_join_destroy(h);
    handles.clear();
}

// ---------------------------------------------------------------- PartitionOp

// PartitionOp::init (operators/partition.cpp:44-105).
void PartitionOp::init(const Setting& root, const Setting& node)
{
    Operator::init(root, node);
    Schema& in = nextOp->getOutSchema();
    for (unsigned c = 0; c < in.columns(); ++c)
        schema.add(in.getColumnType(c), in.getColumnWidth(c));
    attr = (unsigned)node["attr"].asInt();
    if (attr >= schema.columns())
        throw InvalidParameter("attr out of range");
    ColumnType t = schema.getColumnType(attr);
    if (!(t == CT_INTEGER || t == CT_LONG || t == CT_DATE))
        throw IllegalSchemaDeclarationException("partition attribute must be numeric");
    const Setting& r = node["range"];
    range_min = r[0].asInt();
    range_max = r[1].asInt();
    const int64_t b = node["buckets"].asInt();
    if (b <= 0)
        throw MissingParameterException("Number of hash buckets cannot be zero.");
    if (b >= MAX_THREADS)
        throw InvalidParameter("buckets must be below 128");   // partition.cpp:63-64
    buckets = (unsigned)b;
    if (!node.exists("maxtuples") && !node.exists("maxtuplesinM"))
        throw MissingParameterException("maxtuples");
    sortoutput = node["sort"].asString() == "yes";
    sortattr = node.exists("sortattr") ? (unsigned)node["sortattr"].asInt() : attr;
    barrier.reset(new Barrier(buckets));
    input_failed.reset(1);
    staged.assign(MAX_THREADS, {});
    outputloc.assign(MAX_THREADS, 0);
    overlay.assign(MAX_THREADS, Page());
}

void PartitionOp::threadInit(unsigned short tid)
{
    if (tid >= buckets)
        throw InvalidParameter("PartitionOp runs one thread a partition");
}

// PartitionOp::scanStart (partition.cpp:271-359): every thread stages its
// input; thread 0 partitions the staged tuples on the device; each thread
// then serves its own partition.
Operator::ResultCode PartitionOp::scanStart(unsigned short tid)
{
    const ResultCode rc = drain(nextOp, tid, staged[tid]);
    outputloc[tid] = 0;
    if (rc == Error)
        input_failed.raise(0);
    barrier->arrive();
    if (tid == 0 && input_failed.raised(0)) {
        for (unsigned t = 0; t < buckets; ++t)
            std::vector<char>().swap(staged[t]);
        part_size.assign(buckets, 0);
        part_start.assign(buckets + 1, 0);
        part_rc = PJ_ERR_STATE;
    } else if (tid == 0) {
        std::vector<char> all;
        for (unsigned t = 0; t < buckets; ++t) {
            all.insert(all.end(), staged[t].begin(), staged[t].end());
            std::vector<char>().swap(staged[t]);
        }
        const unsigned w = schema.getTupleSize();
        const uint64_t n = all.size() / w;
        partitioned.assign(all.size(), 0);
        part_size.assign(buckets, 0);
        pj_schema cs;
        schema.to_c(&cs);
        int prc = pj_partition_exactrange(&cs, attr, range_min, range_max, buckets, all.data(), n,
                                          partitioned.data(), part_size.data());
        part_start.assign(buckets + 1, 0);
        for (unsigned p = 0; p < buckets; ++p)
            part_start[p + 1] = part_start[p] + part_size[p];
        if (prc == PJ_OK && sortoutput) {
            // sortAllInPage (partition.cpp:222-246): each partition by sortattr
            const unsigned off = schema.getColumnOffset(sortattr);
            const bool is_int = schema.getColumnType(sortattr) == CT_INTEGER;
            for (unsigned p = 0; p < buckets; ++p) {
                std::vector<uint64_t> idx(part_size[p]);
                for (uint64_t i = 0; i < idx.size(); ++i)
                    idx[i] = i;
                char* base = partitioned.data() + part_start[p] * w;
                auto key = [&](uint64_t i) -> int64_t {
                    if (is_int) {
                        int32_t v;
                        memcpy(&v, base + i * w + off, 4);
                        return v;
                    }
                    int64_t v;
                    memcpy(&v, base + i * w + off, 8);
                    return v;
                };
                std::stable_sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) { return key(a) < key(b); });
                std::vector<char> tmp(part_size[p] * w);
                for (uint64_t i = 0; i < idx.size(); ++i)
                    memcpy(tmp.data() + i * w, base + idx[i] * w, w);
                memcpy(base, tmp.data(), tmp.size());
            }
        }
        part_rc = prc;
    }
    barrier->arrive();
    if (rc == Error || input_failed.raised(0) || part_rc != PJ_OK)
        return Error;
    return Ready;
}

// PartitionOp::getNext (partition.cpp:376-404): buffsize slices of this
// thread's partition, Finished with the last one.
Operator::GetNextResultT PartitionOp::getNext(unsigned short tid)
{
    const unsigned w = schema.getTupleSize();
    const uint64_t maxout = std::max<uint64_t>(buffsize / w, 1);
    const uint64_t remaining = part_size[tid] - outputloc[tid];
    char* start = partitioned.data() + (part_start[tid] + outputloc[tid]) * w;
    if (remaining <= maxout) {
        overlay[tid] = Page(start, remaining * w, w, true);
        outputloc[tid] += remaining;
        return {Finished, &overlay[tid]};
    }
    overlay[tid] = Page(start, maxout * w, w, true);
    outputloc[tid] += maxout;
    return {Ready, &overlay[tid]};
}

Operator::ResultCode PartitionOp::scanStop(unsigned short tid)
{
    (void)tid;
    barrier->arrive();
    return Ready;
}

// ---------------------------------------------------------------- SortMergeJoinOp

// JoinOp::init + SortMergeJoinOp::init (operators/join.cpp:224-326,
// 905-1020): key columns, projection, thread groups, and the staging sizes
// the reference requires (maxbuildtuples / maxprobetuples, or the ...inM
// forms); its presorted / prepartitioned hints change nothing here.
void SortMergeJoinOp::init(const Setting& root, const Setting& node)
{
    Operator::init(root, node);
    pj_join_desc d;
    memset(&d, 0, sizeof(d));
    Schema& bs = buildOp->getOutSchema();
    Schema& ps = probeOp->getOutSchema();
    bs.to_c(&d.build);
    ps.to_c(&d.probe);
    std::vector<unsigned short> bj = join_attrs(node["buildjattr"]);
    std::vector<unsigned short> pj = join_attrs(node["probejattr"]);
    if (bj.size() != pj.size() || bj.empty() || bj.size() > PJ_MAX_KEYS)
        throw InvalidParameter("buildjattr and probejattr differ in length");
    d.nkeys = (uint32_t)bj.size();
    for (size_t i = 0; i < bj.size(); ++i) {
        if (bj[i] >= bs.columns() || pj[i] >= ps.columns())
            throw InvalidParameter("join attribute out of range");
        d.buildjattr[i] = bj[i];
        d.probejattr[i] = pj[i];
    }
    const Setting& proj = node["projection"];
    if (!proj.isAggregate() || proj.getLength() == 0 || proj.getLength() > PJ_MAX_COLS)
        throw InvalidParameter("projection");
    d.nproj = (uint32_t)proj.getLength();
    for (int i = 0; i < proj.getLength(); ++i) {
        auto pr = parse_projection(proj[i].asString());
        Schema& s = pr.first == PJ_SIDE_BUILD ? bs : ps;
        if (pr.second >= s.columns())
            throw InvalidParameter("projection " + proj[i].asString() + " out of range");
        d.proj_side[i] = (uint8_t)pr.first;
        d.proj_attr[i] = (uint16_t)pr.second;
        schema.add(s.getColumnType(pr.second), s.getColumnWidth(pr.second));
    }
    if (!node.exists("maxbuildtuples") && !node.exists("maxbuildtuplesinM"))
        throw MissingParameterException("maxbuildtuples");
    if (!node.exists("maxprobetuples") && !node.exists("maxprobetuplesinM"))
        throw MissingParameterException("maxprobetuples");
    group_of.assign(MAX_THREADS, -1);
    const Setting& tg = node["threadgroups"];
    for (int g = 0; g < tg.getLength(); ++g) {
        const Setting& grp = tg[g];
        if (grp.getLength() == 0)
            throw InvalidParameter("empty thread group");
        members.emplace_back();
        for (int j = 0; j < grp.getLength(); ++j) {
            const int64_t t = grp[j].asInt();
            if (t < 0 || t >= MAX_THREADS || group_of[t] != -1)
                throw InvalidParameter("thread " + std::to_string(t) + " in two groups");
            group_of[t] = g;
            members.back().push_back((unsigned short)t);
        }
        leader_of_group.push_back((int)grp[0].asInt());
        barriers.emplace_back(new Barrier((unsigned)grp.getLength()));
        pj_smj_t h = nullptr;
        const int rc = pj_smj_create(&d, &h);
        if (rc == PJ_ERR_UNSUPPORTED)
            throw IllegalSchemaDeclarationException(last_pj_error());
        if (rc)
            throw InvalidParameter(last_pj_error());
        handles.push_back(h);
    }
    build_staged.assign(MAX_THREADS, {});
    probe_staged.assign(MAX_THREADS, {});
    rows.assign(leader_of_group.size(), {});
    join_rc.assign(leader_of_group.size(), PJ_OK);
    input_failed.reset(leader_of_group.size());
    outputloc.assign(MAX_THREADS, 0);
    outend.assign(MAX_THREADS, 0);
    overlay.assign(MAX_THREADS, Page());
}

void SortMergeJoinOp::threadInit(unsigned short tid)
{
    if (group_of.at(tid) < 0)
        throw InvalidParameter("thread " + std::to_string(tid) + " is in no thread group");
}


// SortMergeJoinOp::scanStart (join.cpp:1234-1292, BufferAndSort
// :1181-1232): every thread stages both inputs; the group's leader joins
// them on the device; each thread then serves its share of the rows.
Operator::ResultCode SortMergeJoinOp::scanStart(unsigned short tid)
{
    const int g = group_of.at(tid);
    ResultCode rc = drain(buildOp, tid, build_staged[tid]);
    if (drain(probeOp, tid, probe_staged[tid]) == Error)
        rc = Error;
    if (rc == Error)
        input_failed.raise(g);
    barriers[g]->arrive();
    if (tid == leader_of_group[g] && input_failed.raised(g)) {
        for (unsigned short t : members[g]) {
            std::vector<char>().swap(build_staged[t]);
            std::vector<char>().swap(probe_staged[t]);
            outputloc[t] = outend[t] = 0;
        }
        join_rc[g] = PJ_ERR_STATE;
    } else if (tid == leader_of_group[g]) {
        std::vector<char> b, p;
        for (unsigned short t : members[g]) {
            b.insert(b.end(), build_staged[t].begin(), build_staged[t].end());
            p.insert(p.end(), probe_staged[t].begin(), probe_staged[t].end());
            std::vector<char>().swap(build_staged[t]);
            std::vector<char>().swap(probe_staged[t]);
        }
        const uint64_t bw = buildOp->getOutSchema().getTupleSize(), pw = probeOp->getOutSchema().getTupleSize();
        const uint64_t ow = schema.getTupleSize(), nb = bw ? b.size() / bw : 0, np = pw ? p.size() / pw : 0;
        uint64_t cap = std::max<uint64_t>(np, 1), n = 0;
        rows[g].resize(cap * ow);
        int jrc = pj_smj_run_host(handles[g], b.data(), nb, p.data(), np, rows[g].data(), cap, &n);
        if (jrc == PJ_ERR_CAPACITY && n > cap) {   // many-to-many: room for every row, once more
            cap = n;
            rows[g].resize(cap * ow);
            jrc = pj_smj_run_host(handles[g], b.data(), nb, p.data(), np, rows[g].data(), cap, &n);
        }
        rows[g].resize(jrc == PJ_OK ? n * ow : 0);
        join_rc[g] = jrc;
        const uint64_t k = members[g].size();
        for (uint64_t i = 0; i < k; ++i) {
            const unsigned short t = members[g][i];
            outputloc[t] = (jrc == PJ_OK ? n : 0) * i / k;
            outend[t] = (jrc == PJ_OK ? n : 0) * (i + 1) / k;
        }
    }
    barriers[g]->arrive();
    return rc == Error || input_failed.raised(g) || join_rc[g] != PJ_OK ? Error : Ready;
}

// SortMergeJoinOp::getNext (join.cpp:1294-1411): this thread's rows in
// buffsize slices, Finished with the last one.
Operator::GetNextResultT SortMergeJoinOp::getNext(unsigned short tid)
{
    std::vector<char>& r = rows[group_of.at(tid)];
    const unsigned ow = schema.getTupleSize();
    const uint64_t maxout = std::max<uint64_t>(buffsize / ow, 1);
    const uint64_t remaining = outend[tid] - outputloc[tid];
    char* start = r.data() + outputloc[tid] * ow;
    if (remaining <= maxout) {
        overlay[tid] = Page(start, remaining * ow, ow, true);
        outputloc[tid] += remaining;
        return {Finished, &overlay[tid]};
    }
    overlay[tid] = Page(start, maxout * ow, ow, true);
    outputloc[tid] += maxout;
    return {Ready, &overlay[tid]};
}

Operator::ResultCode SortMergeJoinOp::scanStop(unsigned short tid)
{
    barriers[group_of.at(tid)]->arrive();   // no thread frees the rows while another serves them
    return Ready;
}

void SortMergeJoinOp::destroy()
{
    for (pj_smj_t h : handles)
        pj_smj_destroy(h);
    handles.clear();
}

// ---------------------------------------------------------------- AggregateOp

// GenericAggregate::init (operators/genericaggregate.cpp:51-180) with
// AggregateSum::foldinit (aggregatesum.cpp:37-51) or AggregateCount::foldinit
// (aggregatecount.cpp:37-42): the grouping columns ("field" or "fields"),
// then the sum (the summed column's type) or the count (LONG).
void AggregateOp::init(const Setting& root, const Setting& node)
{
    Operator::init(root, node);
    Schema& in = nextOp->getOutSchema();
    std::vector<unsigned> fields;
    if (node.exists("fields")) {
        const Setting& f = node["fields"];
        if (!f.isAggregate())
            throw InvalidParameter("fields must be a list of column numbers");
        for (int i = 0; i < f.getLength(); ++i)
            fields.push_back((unsigned)f[i].asInt());
    } else if (node.exists("field")) {
        fields.push_back((unsigned)node["field"].asInt());
    } else {
        throw MissingParameterException("field");
    }
    for (unsigned f : fields)
        if (f >= in.columns())
            throw InvalidParameter("aggregation field out of range");
    if (fields.size() > 2)
        throw IllegalSchemaDeclarationException("at most two grouping columns on the device");
    unsigned buckets = 1024;
    if (!fields.empty()) {
        const HashSpec hs = make_hashspec(node["hash"]);   // TupleHasher::create's checks
        buckets = std::max(hs.rounded, 1u);
    }
    unsigned sumfield = 0;
    if (!count) {
        sumfield = (unsigned)node["sumfield"].asInt();
        if (sumfield >= in.columns())
            throw InvalidParameter("sumfield out of range");
        const ColumnType t = in.getColumnType(sumfield);
        if (!(t == CT_INTEGER || t == CT_LONG || t == CT_DECIMAL))
            throw InvalidParameter("sumfield must be a numeric column");
    }
    for (unsigned f : fields)
        schema.add(in.getColumnType(f), in.getColumnWidth(f));
    schema.add(count ? CT_LONG : in.getColumnType(sumfield));
    global = node.exists("global");
    if (global) {
        const int64_t t = node["threads"].asInt();
        if (t <= 0 || t > MAX_THREADS)
            throw InvalidParameter("threads");
        threads = (unsigned)t;
        barrier.reset(new Barrier(threads));
    }
    input_failed.reset(1);
    const uint64_t slots = node.exists("slotperbucket") ? (uint64_t)node["slotperbucket"].asInt() : 6;
    in.to_c(&desc.in);
    desc.op = count ? PJ_AGG_COUNT : PJ_AGG_SUM;
    desc.nfields = (uint32_t)fields.size();
    for (size_t i = 0; i < fields.size(); ++i)
        desc.fields[i] = (uint16_t)fields[i];
    desc.sumfield = (uint16_t)sumfield;
    desc.max_groups = std::max<uint64_t>((uint64_t)buckets * std::max<uint64_t>(slots, 1), 1024);
    const int rc = pj_agg_create(&desc, &handle);
    if (rc == PJ_ERR_UNSUPPORTED)
        throw IllegalSchemaDeclarationException(pj_last_error());
    if (rc)
        throw InvalidParameter(pj_last_error());
    // fusable: a global count whose groups are the join's build key columns, in order
    if (auto* j = dynamic_cast<HashJoinOp*>(nextOp)) {
        const pj_join_desc& jd = j->descriptor();
        bool same = count && global && !fields.empty() && fields.size() == jd.nkeys && j->oneGroupOf(threads);
        for (size_t k = 0; same && k < fields.size(); ++k)
            same = jd.proj_side[fields[k]] == PJ_SIDE_BUILD && jd.proj_attr[fields[k]] == jd.buildjattr[k];
        if (same)
            fused = j;
    }
    staged.assign(MAX_THREADS, {});
    groups.assign(MAX_THREADS, {});
    outputloc.assign(MAX_THREADS, 0);
    outend.assign(MAX_THREADS, 0);
    overlay.assign(MAX_THREADS, Page());
}

void AggregateOp::threadInit(unsigned short tid)
{
    if (global && tid >= threads)
        throw InvalidParameter("a global aggregation runs on its own `threads` threads");
}

// The device aggregation of `in`; the handle's table grows to the input's
// size when the plan's buckets were too few (the reference's table chains
// instead).
int AggregateOp::aggregate(const std::vector<char>& in, std::vector<char>& out, uint64_t* ngroups)
{
    std::lock_guard<std::mutex> g(mu);
    const uint64_t w = nextOp->getOutSchema().getTupleSize(), ow = schema.getTupleSize();
    const uint64_t n = w ? in.size() / w : 0, cap = std::max<uint64_t>(n, 1);
    out.resize(cap * ow);
    int rc = pj_agg_run_host(handle, in.data(), n, out.data(), cap, ngroups);
    if (rc == PJ_ERR_CAPACITY && desc.max_groups < cap) {
        pj_agg_destroy(handle);
        handle = nullptr;
        desc.max_groups = std::min<uint64_t>(cap, 1ull << 30);
        if ((rc = pj_agg_create(&desc, &handle)) == PJ_OK)
            rc = pj_agg_run_host(handle, in.data(), n, out.data(), cap, ngroups);
    }
    out.resize(rc == PJ_OK ? *ngroups * ow : 0);
    return rc;
}

// The fused count: every thread stages its share of the join's two inputs,
// thread 0 counts per build key inside the join, and each thread serves an
// even share of the groups.
Operator::ResultCode AggregateOp::scanStartFused(unsigned short tid)
{
    std::vector<char> b, p;
    ResultCode rc = drain(fused->buildOp, tid, staged[tid]);
    if (drain(fused->probeOp, tid, groups[tid]) == Error)   // groups[tid] holds the probe pages until the count
        rc = Error;
    outputloc[tid] = outend[tid] = 0;
    if (rc == Error)
        input_failed.raise(0);
    barrier->arrive();
    if (tid == 0 && input_failed.raised(0)) {
        for (unsigned t = 0; t < threads; ++t) {
            std::vector<char>().swap(staged[t]);
            std::vector<char>().swap(groups[t]);
        }
        agg_rc = PJ_ERR_STATE;
    } else if (tid == 0) {
        for (unsigned t = 0; t < threads; ++t) {
            b.insert(b.end(), staged[t].begin(), staged[t].end());
            p.insert(p.end(), groups[t].begin(), groups[t].end());
            std::vector<char>().swap(staged[t]);
            std::vector<char>().swap(groups[t]);
        }
        const uint64_t bw = fused->buildOp->getOutSchema().getTupleSize();
        const uint64_t pw = fused->probeOp->getOutSchema().getTupleSize();
        const uint64_t nb = bw ? b.size() / bw : 0, np = pw ? p.size() / pw : 0;
        const uint64_t ow = schema.getTupleSize(), cap = std::max<uint64_t>(nb, 1);
        uint64_t ng = 0;
        groups[0].resize(cap * ow);
        std::lock_guard<std::mutex> g(mu);
        agg_rc = pj_join_count_by_key_host(fused->handleOf(0), b.data(), nb, p.data(), np, groups[0].data(), cap, &ng);
        groups[0].resize(agg_rc == PJ_OK ? ng * ow : 0);
        if (agg_rc != PJ_OK)
            ng = 0;
        for (unsigned t = 0; t < threads; ++t) {
            outputloc[t] = ng * t / threads;
            outend[t] = ng * (t + 1) / threads;
        }
    }
    barrier->arrive();
    return rc == Error || input_failed.raised(0) || agg_rc != PJ_OK ? Error : Ready;
}

// GenericAggregate::scanStart (genericaggregate.cpp:291-424): drain the
// input, then aggregate it (global mode: all threads' input, once).
Operator::ResultCode AggregateOp::scanStart(unsigned short tid)
{
    if (fused)
        return scanStartFused(tid);
    const ResultCode rc = drain(nextOp, tid, staged[tid]);
    outputloc[tid] = outend[tid] = 0;
    if (!global) {
        uint64_t ng = 0;
        const int arc = rc == Error ? PJ_OK : aggregate(staged[tid], groups[tid], &ng);
        std::vector<char>().swap(staged[tid]);
        outend[tid] = ng;
        return rc == Error || arc != PJ_OK ? Error : Ready;
    }
    if (rc == Error)
        input_failed.raise(0);
    barrier->arrive();
    if (tid == 0 && input_failed.raised(0)) {
        for (unsigned t = 0; t < threads; ++t)
            std::vector<char>().swap(staged[t]);
        agg_rc = PJ_ERR_STATE;
    } else if (tid == 0) {
        std::vector<char> all;
        for (unsigned t = 0; t < threads; ++t) {
            all.insert(all.end(), staged[t].begin(), staged[t].end());
            std::vector<char>().swap(staged[t]);
        }
        uint64_t ng = 0;
        agg_rc = aggregate(all, groups[0], &ng);
        for (unsigned t = 0; t < threads; ++t) {   // an even share of the groups a thread
            outputloc[t] = ng * t / threads;
            outend[t] = ng * (t + 1) / threads;
        }
    }
    barrier->arrive();
    return rc == Error || input_failed.raised(0) || agg_rc != PJ_OK ? Error : Ready;
}

// GenericAggregate::getNext (genericaggregate.cpp:426-495): this thread's
// groups in buffsize slices, Finished with the last one.
Operator::GetNextResultT AggregateOp::getNext(unsigned short tid)
{
    std::vector<char>& g = global ? groups[0] : groups[tid];
    const unsigned ow = schema.getTupleSize();
    const uint64_t maxout = std::max<uint64_t>(buffsize / ow, 1);
    const uint64_t remaining = outend[tid] - outputloc[tid];
    char* start = g.data() + outputloc[tid] * ow;
    if (remaining <= maxout) {
        overlay[tid] = Page(start, remaining * ow, ow, true);
        outputloc[tid] += remaining;
        return {Finished, &overlay[tid]};
    }
    overlay[tid] = Page(start, maxout * ow, ow, true);
    outputloc[tid] += maxout;
    return {Ready, &overlay[tid]};
}

Operator::ResultCode AggregateOp::scanStop(unsigned short tid)
{
    if (global)
        barrier->arrive();   // no thread frees the groups while another serves them
    else
        std::vector<char>().swap(groups[tid]);
    return Ready;
}

void AggregateOp::destroy()
{
    if (handle)
        pj_agg_destroy(handle);
    handle = nullptr;
}

// ---------------------------------------------------------------- MergeOp

void MergeOp::init(const Setting& root, const Setting& node)
{
    Operator::init(root, node);
    const int64_t t = node["threads"].asInt();
    if (t <= 0 || t > MAX_THREADS)
        throw InvalidParameter("threads");
    threads = (unsigned)t;
    Schema& in = nextOp->getOutSchema();
    for (unsigned c = 0; c < in.columns(); ++c)
        schema.add(in.getColumnType(c), in.getColumnWidth(c));
}

namespace {
void collect(Operator* op, std::vector<Operator*>& v)
{
    if (!op)
        return;
    v.push_back(op);
    for (Operator* c : op->children())
        collect(c, v);
}
}  // namespace

bool MergeOp::run(const std::function<void(const Page&)>& consume)
{
    std::vector<Operator*> sub;
    collect(nextOp, sub);
    for (Operator* op : sub)   // a global aggregation's barrier waits for exactly its `threads`
        if (auto* a = dynamic_cast<AggregateOp*>(op))
            if (a->globalThreads() && a->globalThreads() != threads)
                throw InvalidParameter("a global aggregation's threads must equal the merge's");
    std::mutex mu;
    std::atomic<unsigned> failed_threads(0);
    std::vector<std::string> errors;
    std::string first_error;
    auto worker = [&](unsigned short tid) {
        bool failed = false;
        size_t inited = 0, closed = 0;   // operators of sub whose threadInit / threadClose ran
        bool stopped = true;             // no scanStop owed
        try {
            for (Operator* op : sub) {
                op->threadInit(tid);
                ++inited;
            }
            stopped = false;
            ResultCode rc = nextOp->scanStart(tid);
            if (rc == Error)
                failed = true;
            GetNextResultT r(rc, nullptr);
            while (r.first == Ready) {
                r = nextOp->getNext(tid);
                if (r.first == Error) {
                    failed = true;
                    break;
                }
                if (r.second && r.second->getUsedSpace()) {
                    std::lock_guard<std::mutex> g(mu);
                    consume(*r.second);
                }
            }
            stopped = true;
            if (nextOp->scanStop(tid) == Error)
                failed = true;
            while (closed < inited)
                sub[closed++]->threadClose(tid);
        } catch (const std::exception& e) {
            {
                std::lock_guard<std::mutex> g(mu);
                errors.push_back(e.what());
            }
            failed = true;
            // finish the thread's part of the protocol, so that the threads of
            // its groups waiting at a barrier in scanStop or threadClose go on
            try {
                if (!stopped)
                    nextOp->scanStop(tid);
            } catch (const std::exception&) {
            }
            while (closed < inited) {
                try {
                    sub[closed++]->threadClose(tid);
                } catch (const std::exception&) {
                }
            }
        }
        if (failed) {
            ++failed_threads;
            const char* why = pj_last_error();   // this thread's library failure, if it had one
            std::lock_guard<std::mutex> g(mu);
            if (why && *why && first_error.empty())
                first_error = why;
        }
    };
    std::vector<std::thread> th;
    for (unsigned t = 0; t < threads; ++t)
        th.emplace_back(worker, (unsigned short)t);
    for (auto& x : th)
        x.join();
    nfailed = failed_threads.load();
    failure = first_error;
    if (!errors.empty())
        throw QueryExecutionError(errors.front());
    return nfailed == 0;
}

// ---------------------------------------------------------------- Query

Query::~Query()
{
    for (auto& op : ops)
        op->destroy();
}

void Query::create(const std::string& plan_text, const std::map<std::string, Source>& sources)
{
    cfg = parse_config(plan_text);
    srcs = sources;
    if (!cfg->exists("buffsize"))
        throw MissingParameterException("buffsize");
    tree = build((*cfg)["treeroot"]);
}

// constructsubtree (query.cpp:198-442): children first, then init.
Operator* Query::build(const Setting& sub)
{
    const std::string name = sub["name"].asString();
    const Setting& node = (*cfg)[name];
    const std::string type = node["type"].asString();
    Operator* op = nullptr;
    if (type == "hashjoin") {
        auto* j = new HashJoinOp();
        ops.emplace_back(j);
        j->buildOp = build(sub["build"]);
        j->probeOp = build(sub["probe"]);
        op = j;
    } else if (type == "partition") {
        auto* p = new PartitionOp();
        ops.emplace_back(p);
        p->nextOp = build(sub["input"]);
        op = p;
    } else if (type == "sortmergejoin" || type == "mpsmjoin" || type == "newmpsmjoin") {
        auto* j = new SortMergeJoinOp();
        ops.emplace_back(j);
        j->buildOp = build(sub["build"]);
        j->probeOp = build(sub["probe"]);
        op = j;
    } else if (type == "aggregate_sum" || type == "aggregate_count") {
        auto* a = new AggregateOp();
        ops.emplace_back(a);
        a->count = type == "aggregate_count";
        a->nextOp = build(sub["input"]);
        op = a;
    } else if (type == "merge") {
        auto* m = new MergeOp();
        ops.emplace_back(m);
        m->nextOp = build(sub["input"]);
        op = m;
    } else if (type == "threadidprepend") {
        auto* t = new ThreadIdPrependOp();
        ops.emplace_back(t);
        t->nextOp = build(sub["input"]);
        op = t;
    } else if (type == "memscan") {
        auto* s = new MemScanOp();
        ops.emplace_back(s);
        auto it = srcs.find(name);
        if (it == srcs.end())
            throw MissingParameterException("input buffer for memscan node " + name);
        s->setData(it->second.data, it->second.tuples);
        op = s;
    } else if (type == "scan" || type == "parallelscan") {
        if (node.exists("filetype") && node["filetype"].asString() != "binary")
            throw InvalidParameter("only binary tables are read here (filetype = \"binary\")");
        auto* s = new BinaryScanOp();
        ops.emplace_back(s);
        op = s;
    } else {
        throw PlanError("unknown operator type " + type + " (node " + name + ")");
    }
    op->init(*cfg, node);
    return op;
}

}  // namespace pjh

// ---------------------------------------------------------------- C ABI

extern "C" int pj_query_run(const char* plan_text, uint32_t nsources, const char* const* names,
                            const void* const* data, const uint64_t* tuples, void* out, uint64_t out_cap_bytes,
                            uint64_t* out_bytes, uint32_t* out_tuple_width)
{
    using namespace pjh;
    if (!plan_text || !out_bytes || (nsources && (!names || !data || !tuples)))
        return -1;
    *out_bytes = 0;
    try {
        std::map<std::string, Query::Source> srcs;
        for (uint32_t i = 0; i < nsources; ++i)
            srcs[names[i]] = Query::Source{data[i], tuples[i]};
        Query q;
        q.create(plan_text, srcs);
        if (out_tuple_width)
            *out_tuple_width = q.root()->getOutSchema().getTupleSize();
        MergeOp* m = q.merge();
        if (!m)
            throw PlanError("the plan's treeroot must be a merge node");
        uint64_t used = 0;
        bool ok = m->run([&](const Page& p) {
            const uint64_t n = p.getUsedSpace();
            if (used + n <= out_cap_bytes && out)
                memcpy(static_cast<char*>(out) + used, p.data(), n);
            used += n;
        });
        *out_bytes = used;
        if (!ok) {
            std::string msg = std::to_string(m->failedThreads()) + " of " + std::to_string(m->threads) +
                              " threads returned Error";
            if (!m->failure.empty())
                msg += ": " + m->failure;
            pj_set_error(msg.c_str());
            return PJ_ERR_STATE;
        }
        return used > out_cap_bytes ? PJ_ERR_CAPACITY : PJ_OK;
    } catch (const pjh::IllegalSchemaDeclarationException& e) {
        pj_set_error(e.what());
        return PJ_ERR_UNSUPPORTED;
    } catch (const pjh::PlanError& e) {
        pj_set_error(e.what());
        return PJ_ERR_INVALID;
    } catch (const std::exception& e) {
        pj_set_error(e.what());
        return PJ_ERR_STATE;
    }
}
