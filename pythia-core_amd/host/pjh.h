// pjh.h -- the Operator API of Pythia's hash-join path, on the C ABI of
// libpjoin (include/pjoin.h).
//
// A query plan written for the reference selects these operators by the
// same node types and parameters: `hashjoin` (operators/operators.h:2965-3132,
// operators/join.cpp), `partition` (operators/operators.h:4250-4332,
// operators/partition.cpp) and `merge` (operators/merge.cpp), with the same
// call protocol (init once after the children, then per thread threadInit,
// scanStart, getNext until Finished, scanStop, threadClose; destroy once) and
// the same Page, Schema and result-code conventions. Paths are relative to
// the reference tree.
#pragma once

#include <pthread.h>
#include <stdint.h>

#include <atomic>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/pjoin.h"

namespace pjh {

static const unsigned short MAX_THREADS = 128;  // operators/operators.h:94

// ------------------------------------------------------------ exceptions
// The reference's exception types (exceptions.h) that init and the plan
// builder throw.
struct PlanError : std::runtime_error {
    explicit PlanError(const std::string& s) : std::runtime_error(s) {}
};
struct MissingParameterException : PlanError {
    explicit MissingParameterException(const std::string& s) : PlanError("missing parameter: " + s) {}
};
struct InvalidParameter : PlanError {
    explicit InvalidParameter(const std::string& s) : PlanError("invalid parameter: " + s) {}
};
struct UnknownHashException : PlanError {
    explicit UnknownHashException(const std::string& s) : PlanError("unknown hash function: " + s) {}
};
struct IllegalSchemaDeclarationException : PlanError {
    explicit IllegalSchemaDeclarationException(const std::string& s) : PlanError("illegal schema: " + s) {}
};
struct QueryExecutionError : std::runtime_error {
    explicit QueryExecutionError(const std::string& s) : std::runtime_error(s) {}
};

// ------------------------------------------------------------ plan config
// A parsed libconfig setting: group, list, array or scalar. The parser takes
// the libconfig 1.x syntax the reference's plans use: name = value or
// name : value, optional ; or , terminators, groups { }, lists ( ), arrays
// [ ], integers (decimal or hex, 64-bit here; libconfig 1.2 truncates to 32),
// floats, booleans, "strings" (adjacent ones concatenate) and #, //, /* */
// comments.
class Setting {
  public:
    enum Type { None, Int, Float, Bool, String, Group, Array, List };

    Type type = None;
    std::string name;
    int64_t ival = 0;
    double fval = 0;
    std::string sval;
    std::vector<std::unique_ptr<Setting>> children;

    bool isGroup() const { return type == Group; }
    bool isAggregate() const { return type == Group || type == Array || type == List; }
    bool isNumber() const { return type == Int || type == Float; }
    int getLength() const { return (int)children.size(); }
    bool exists(const std::string& n) const;
    const Setting& operator[](const std::string& n) const;
    const Setting& operator[](int i) const;
    int64_t asInt() const;
    std::string asString() const;
};

// Parses plan text; throws PlanError with the line of a syntax error.
std::unique_ptr<Setting> parse_config(const std::string& text);

// ------------------------------------------------------------ schema, page
enum ColumnType { CT_INTEGER = 0, CT_LONG, CT_DECIMAL, CT_CHAR, CT_DATE, CT_POINTER };

// Packed columns, no alignment (schema.cpp:82-117).
class Schema {
  public:
    void add(ColumnType ct, unsigned width = 0);
    unsigned columns() const { return (unsigned)types.size(); }
    unsigned getTupleSize() const { return size; }
    ColumnType getColumnType(unsigned c) const { return types.at(c); }
    unsigned getColumnWidth(unsigned c) const { return widths.at(c); }
    unsigned getColumnOffset(unsigned c) const { return offsets.at(c); }
    void* calcOffset(void* tuple, unsigned c) const { return static_cast<char*>(tuple) + offsets.at(c); }
    // Parses "int", "long", "dec", "date", "pointer", "char(n)".
    static ColumnType parse(const std::string& s, unsigned* width);
    void to_c(pj_schema* out) const;

  private:
    std::vector<ColumnType> types;
    std::vector<unsigned> widths, offsets;
    unsigned size = 0;
};

// A fixed-capacity buffer of packed tuples (util/buffer.h:134-354), owning its
// memory or overlaying someone else's.
class Page {
  public:
    Page() {}
    Page(uint64_t bytes, unsigned tuplesize);                       // owns
    Page(void* data, uint64_t bytes, unsigned tuplesize, bool full);  // overlay
    unsigned tuplesize() const { return tsize; }
    uint64_t maxsize() const { return cap; }
    uint64_t getUsedSpace() const { return used; }
    uint64_t getNumTuples() const { return tsize ? used / tsize : 0; }
    bool canStoreTuple() const { return used + tsize <= cap; }
    void* allocateTuple();
    void clear() { used = 0; }
    char* data() const { return base; }
    void* getTupleOffset(uint64_t i) const { return i * tsize < used ? base + i * tsize : nullptr; }

  private:
    std::shared_ptr<std::vector<char>> own;
    char* base = nullptr;
    uint64_t cap = 0, used = 0;
    unsigned tsize = 0;
};

// ------------------------------------------------------------ hashing
// TupleHasher::create's validation (hash.cpp:101-248): the function names,
// field rules and bucket rounding of the reference. The device join
// partitions on its own hash, so the host keeps these for checking a plan
// and for PartitionOp's exactrange bounds.
struct HashSpec {
    std::string fn;
    unsigned buckets = 0;   // as given in the plan
    unsigned rounded = 0;   // HashFunction::buckets() after rounding
    int64_t range_min = 0, range_max = 0;
    unsigned offset = 0, skipbits = 0;
};
unsigned getlogarithm(unsigned k);                                   // hash.cpp:49-58
HashSpec make_hashspec(const Setting& node);                         // throws

// ------------------------------------------------------------ operators
class Barrier {
  public:
    explicit Barrier(unsigned n = 1) { pthread_barrier_init(&b, nullptr, n); }
    ~Barrier() { pthread_barrier_destroy(&b); }
    void arrive() { pthread_barrier_wait(&b); }

  private:
    pthread_barrier_t b;
};

class Operator {
  public:
    enum ResultCode { Ready = 0, Finished, Error };
    typedef std::pair<ResultCode, Page*> GetNextResultT;

    virtual ~Operator() {}
    virtual void init(const Setting& root, const Setting& node);
    virtual void threadInit(unsigned short tid) { (void)tid; }
    virtual ResultCode scanStart(unsigned short tid) = 0;
    virtual GetNextResultT getNext(unsigned short tid) = 0;
    virtual ResultCode scanStop(unsigned short tid) = 0;
    virtual void threadClose(unsigned short tid) { (void)tid; }
    virtual void destroy() {}
    virtual std::vector<Operator*> children() { return {}; }
    Schema& getOutSchema() { return schema; }

  protected:
    Schema schema;
    unsigned buffsize = 0;
};

// Serves host memory registered with the query (the build's stand-in for
// the reference's scans; the table is handed out a page at a time by
// fetch-and-add as ParallelScanOp does, operators/parallelscan.cpp:176-190).
class MemScanOp : public Operator {
  public:
    void setData(const void* d, uint64_t n) { data = static_cast<const char*>(d); count = n; }
    void init(const Setting& root, const Setting& node) override;
    void threadInit(unsigned short tid) override;
    ResultCode scanStart(unsigned short) override { return Ready; }
    GetNextResultT getNext(unsigned short tid) override;
    ResultCode scanStop(unsigned short) override { return Ready; }

  private:
    const char* data = nullptr;
    uint64_t count = 0;
    uint64_t per_page = 1;
    volatile uint64_t cursor = 0;
    std::vector<Page> pages;
    // Failure injection for tests of error handling: thread error_thread
    // returns Error from its getNext call after error_after pages.
    int error_thread = -1;
    unsigned error_after = 0;
    std::vector<unsigned> served;   // pages handed to each thread
};

// A raw binary file of packed tuples with no header, the layout of the
// reference's MemMappedTable (operators/loaders/table.cpp:170-196).
class BinaryScanOp : public MemScanOp {
  public:
    ~BinaryScanOp() override;
    void init(const Setting& root, const Setting& node) override;

  private:
    void* map = nullptr;
    uint64_t mapsize = 0;
};

// One flag per thread group, raised before the group's barrier by a thread
// whose input returned Error and read after it: the staged input of the
// group is then incomplete, so the group's device work is skipped and every
// thread of the group returns Error (the reference's thread returns Error
// without arriving, and its peers wait at the barrier for good,
// join.cpp:545-547, 586-588).
class GroupFlags {
  public:
    void reset(size_t n)
    {
        f.reset(new std::atomic<bool>[n]);
        for (size_t i = 0; i < n; ++i)
            f[i].store(false);
    }
    void raise(size_t g) { f[g].store(true); }
    bool raised(size_t g) const { return f[g].load(); }

  private:
    std::unique_ptr<std::atomic<bool>[]> f;
};

class HashJoinOp : public Operator {
  public:
    Operator* buildOp = nullptr;
    Operator* probeOp = nullptr;

    void init(const Setting& root, const Setting& node) override;
    void threadInit(unsigned short tid) override;
    ResultCode scanStart(unsigned short tid) override;
    GetNextResultT getNext(unsigned short tid) override;
    ResultCode scanStop(unsigned short tid) override;
    void threadClose(unsigned short tid) override;
    void destroy() override;
    std::vector<Operator*> children() override { return {buildOp, probeOp}; }

    // Probe tuples gathered per device call.
    uint64_t probe_batch_tuples = 1 << 22;

    const pj_join_desc& descriptor() const { return desc_; }
    pj_join_t handleOf(unsigned g) const { return handles.at(g); }
    // Whether threads 0 .. n-1 form the only thread group and all of them build.
    bool oneGroupOf(unsigned n) const;

  private:
    struct ThreadState {
        pj_cursor cur = {nullptr, 0};
        Page out;
        std::vector<char> batch;
        bool probe_done = false;
        bool started = false;
    };
    ResultCode fillPage(unsigned short tid, ThreadState& st);

    std::vector<pj_join_t> handles;   // one device table per thread group
    pj_join_desc desc_ = {};
    std::vector<int> build_rc;
    GroupFlags build_failed;
    HashSpec hash;
    unsigned tuplesperbucket = 0;
    std::vector<int> group_of, leader_of_group, groupsize;
    std::vector<std::unique_ptr<Barrier>> barriers;
    std::vector<std::unique_ptr<ThreadState>> state;
    std::vector<unsigned short> leftthreads;
    bool isLeftThread(unsigned short tid) const;
};

class PartitionOp : public Operator {
  public:
    Operator* nextOp = nullptr;

    void init(const Setting& root, const Setting& node) override;
    void threadInit(unsigned short tid) override;
    ResultCode scanStart(unsigned short tid) override;
    GetNextResultT getNext(unsigned short tid) override;
    ResultCode scanStop(unsigned short tid) override;
    std::vector<Operator*> children() override { return {nextOp}; }

  private:
    unsigned attr = 0, buckets = 0;
    int64_t range_min = 0, range_max = 0;
    bool sortoutput = false;
    unsigned sortattr = 0;
    std::unique_ptr<Barrier> barrier;
    GroupFlags input_failed;                      // one group: every thread
    std::vector<std::vector<char>> staged;       // per thread
    std::vector<char> partitioned;                // all partitions, back to back
    std::vector<uint64_t> part_start, part_size;
    std::vector<uint64_t> outputloc;
    std::vector<Page> overlay;
    int part_rc = 0;
};

// sortmergejoin / mpsmjoin (SortMergeJoinOp, MPSMJoinOp,
// operators/join.cpp:905-1791) on the device: the threads of a group stage
// their build and probe pages, the group's leader joins them with the device
// sort-merge join, and each thread of the group serves an even share of the
// rows. Same rows as a hashjoin node with the same projection.
class SortMergeJoinOp : public Operator {
  public:
    Operator* buildOp = nullptr;
    Operator* probeOp = nullptr;

    void init(const Setting& root, const Setting& node) override;
    void threadInit(unsigned short tid) override;
    ResultCode scanStart(unsigned short tid) override;
    GetNextResultT getNext(unsigned short tid) override;
    ResultCode scanStop(unsigned short tid) override;
    void destroy() override;
    std::vector<Operator*> children() override { return {buildOp, probeOp}; }

  private:
    std::vector<pj_smj_t> handles;   // one a thread group
    std::vector<int> group_of, leader_of_group;
    std::vector<std::vector<unsigned short>> members;
    std::vector<std::unique_ptr<Barrier>> barriers;
    std::vector<std::vector<char>> build_staged, probe_staged, rows;   // per thread / per group
    std::vector<int> join_rc;
    GroupFlags input_failed;
    std::vector<uint64_t> outputloc, outend;
    std::vector<Page> overlay;
};

// aggregate_sum / aggregate_count (GenericAggregate,
// operators/genericaggregate.cpp:51-515, with AggregateSum's or
// AggregateCount's fold) on the device. Global mode (a "global" key): every
// thread stages its input, the last one to arrive has the device aggregate
// all of it, and each of the `threads` threads serves an even share of the
// groups. Thread-local mode: each thread aggregates its own input and serves
// its own groups (a key can then appear once per thread, as in the
// reference).
class AggregateOp : public Operator {
  public:
    Operator* nextOp = nullptr;
    bool count = false;   // aggregate_count; else aggregate_sum

    void init(const Setting& root, const Setting& node) override;
    void threadInit(unsigned short tid) override;
    ResultCode scanStart(unsigned short tid) override;
    GetNextResultT getNext(unsigned short tid) override;
    ResultCode scanStop(unsigned short tid) override;
    void destroy() override;
    std::vector<Operator*> children() override { return {nextOp}; }
    unsigned globalThreads() const { return global ? threads : 0u; }

  private:
    int aggregate(const std::vector<char>& in, std::vector<char>& out, uint64_t* ngroups);
    ResultCode scanStartFused(unsigned short tid);

    // A count grouped by the build key over a hashjoin child that runs as one
    // thread group: counted inside the join (pj_join_count_by_key), without
    // the join's rows.
    HashJoinOp* fused = nullptr;

    pj_agg_desc desc = {};
    pj_agg_t handle = nullptr;
    std::mutex mu;                 // one device aggregation at a time
    bool global = false;
    unsigned threads = 1;
    std::unique_ptr<Barrier> barrier;
    std::vector<std::vector<char>> staged;    // per thread
    std::vector<std::vector<char>> groups;    // per thread (global mode: all in [0])
    std::vector<uint64_t> outputloc, outend;
    std::vector<Page> overlay;
    int agg_rc = 0;
    GroupFlags input_failed;       // global and fused modes: one group
    std::atomic<unsigned> arrived{0};
};

// Runs its subtree on T threads and hands their pages to the single
// consumer thread (operators/merge.cpp:304-649).
class MergeOp : public Operator {
  public:
    Operator* nextOp = nullptr;
    unsigned threads = 1;

    void init(const Setting& root, const Setting& node) override;
    ResultCode scanStart(unsigned short) override { return Ready; }
    GetNextResultT getNext(unsigned short) override { return {Error, nullptr}; }
    ResultCode scanStop(unsigned short) override { return Ready; }
    std::vector<Operator*> children() override { return {nextOp}; }
    // Drives threadInit .. threadClose on T threads; consume() sees every
    // output page (serialised). Returns false if a thread saw Error;
    // failedThreads() then counts them.
    bool run(const std::function<void(const Page&)>& consume);
    unsigned failedThreads() const { return nfailed; }
    std::string failure;   // the first library error a failed thread saw, if any

  private:
    unsigned nfailed = 0;
};

// Query::create (query.cpp:198-461): builds the operator tree of a plan.
// Sources of type "memscan" take the buffers registered under their node
// name.
class Query {
  public:
    struct Source {
        const void* data;
        uint64_t tuples;
    };
    void create(const std::string& plan_text, const std::map<std::string, Source>& sources);
    Operator* root() { return tree; }
    MergeOp* merge() { return dynamic_cast<MergeOp*>(tree); }
    ~Query();

  private:
    Operator* build(const Setting& subtree);
    std::unique_ptr<Setting> cfg;
    std::vector<std::unique_ptr<Operator>> ops;
    std::map<std::string, Source> srcs;
    Operator* tree = nullptr;
};

}  // namespace pjh
