// pj_batch.hip -- many independent joins of one descriptor in one launch.
//
// No counterpart in the reference, which runs one HashJoinOp per plan
// (query.cpp:301-334). A batch runs in one of these modes:
//   PJ_BATCH_GRAPH / PJ_BATCH_PLAIN  the batched pipeline (pj_join.hip,
//       MBatch): every phase one kernel for all joins of the batch -- the
//       tile histograms of every side, one scan, the tile-sorted scatter
//       into one record buffer, each side's partition starts, each join's
//       work items counted and scanned on the device, and one persistent
//       join kernel over all joins' items (LDS bucket tables per build
//       partition, as the single join's). GRAPH replays the pipeline
//       captured into a hipGraph at pj_batch_prepare: one launch a batch.
//       The last item of each join stamps the device clock, which gives
//       every join's completion time within the batch.
//   PJ_BATCH_SEQUENTIAL / CONCURRENT / CONCURRENT_GRAPH  every join a whole
//       single radix join (pj_join_run_device's kernels), one after another,
//       on four streams, or those four streams captured into one graph.
// The tuner picks, per size class (floor(log2(build tuples))), the radix
// bits and the probe records of a work item by timing candidates on joins
// of that class from the batch itself, through the pipeline.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "pj_device.h"
#include "pj_internal.h"
#include "pj_scan.h"

namespace pj {
namespace {

constexpr uint32_t kBMinBits = 1, kBMaxBits = 12;
constexpr uint32_t kBTile = 32768;   // tuples a tile of the pipeline's partition passes
constexpr uint32_t kPieceUnit = 4096;   // pj_batch_cfg.piece counts probe records a work item in these
constexpr uint32_t kClasses = 32;
constexpr uint32_t kLanes = 4;       // streams of the concurrent mode (4 hardware queues a process)

// The batched pipeline is switched off: its first device run (a three-join
// batch, prepared and captured into a graph, then launched directly) did
// not finish, and no cause has been found in its code (a CPU replay of its
// plan and index math keeps every access in bounds and every loop bounded).
// The one-kernel-per-phase pipeline that came before it stopped the same way
// at its first run. PJ_BATCH_PLAIN and PJ_BATCH_GRAPH return
// PJ_ERR_UNSUPPORTED, pj_batch_prepare only keeps the list, and
// pj_batch_tune times the sequential mode.
constexpr bool kPipelineOn = false;

uint32_t size_class(uint64_t n)
{
    uint32_t c = 0;
    while (c + 1 < kClasses && (2ull << c) <= n)
        ++c;
    return c;
}

}  // namespace
}  // namespace pj

using namespace pj;

// ======================================================================
// handle
// ======================================================================

struct pj_batch_s {
    pj_join_desc desc;
    Layout lay;
    std::mutex mu;
    pj_batch_cfg cfg[kClasses];
    bool tuned[kClasses];
    // the prepared batch
    std::vector<pj_batch_join> joins;
    MBatch mb;
    std::vector<uint32_t> bits;         // radix bits of each join in the pipeline
    hipGraphExec_t exec = nullptr;      // the pipeline captured for the prepared batch
    bool prepared = false;
    bool ran_pipeline = false;
    // sequential mode: the joins run one by one through pj_join_run_device,
    // one join handle per radix-bits value
    // one join handle per (radix bits, lane): the lanes of the concurrent
    // mode each run on a stream of their own, with buffers of their own
    std::map<std::pair<uint32_t, uint32_t>, pj_join_t> seq;
    // largest sides a handle has run: build, probe of the appended layout,
    // probe of the contiguous one
    struct Warm {
        uint64_t build = 0, probe = 0, probe_small = 0;
    };
    std::map<std::pair<uint32_t, uint32_t>, Warm> seq_warm;
    std::vector<hipStream_t> lanes;   // created for the concurrent mode, default priority
    hipGraphExec_t lane_exec = nullptr;     // the concurrent mode captured (PJ_BATCH_CONCURRENT_GRAPH)
    std::vector<uint32_t> lane_graph_bits;  // radix bits of each join in that capture
    std::vector<hipEvent_t> seq_ev;   // two a join, reused across runs
    uint32_t* seq_slots = nullptr;    // pinned: 8 counter words a join
    size_t seq_nslots = 0;
    bool last_sequential = false;
    std::vector<uint64_t> seq_counts;
    std::vector<double> seq_done_us, seq_run_us;
};

namespace {

// Radix bits of a join of nb build tuples under a configuration: the
// configuration's, else the fewest whose mean partition fills at most
// load_factor of an LDS table, else (load factor 0) the single join's choice.
uint32_t cfg_bits(const pj_batch_cfg& c, uint64_t nb)
{
    uint32_t b;
    if (c.bits) {
        b = c.bits;
    } else if (c.load_factor > 0.f) {
        b = 0;
        while (b < kBMaxBits && (double)nb / (double)(1u << b) > (double)c.load_factor * kTableCap)
            ++b;
    } else {
        b = choose_build_bits(nb, -1);
    }
    return std::max(kBMinBits, std::min(kBMaxBits, b));
}

// Probe records of a work item: the configuration's, else about 1/8 of a
// CU's share of the join's probe side, at least 32768.
uint64_t cfg_piece(const pj_batch_cfg& c, uint64_t np, uint32_t cus)
{
    if (c.piece)
        return (uint64_t)c.piece * kPieceUnit;
    return std::max<uint64_t>((np + (uint64_t)cus * 8 - 1) / ((uint64_t)cus * 8), 32768);
}

pj_batch_cfg default_cfg()
{
    pj_batch_cfg c;
    c.bits = 0;   // the single join's choice
    c.passes = 1;
    c.piece = 0;
    c.load_factor = 0.f;
    return c;
}

int cfg_valid(const pj_batch_cfg& c)
{
    return c.bits <= kBMaxBits && c.passes == 1 && c.load_factor >= 0.f && c.load_factor <= 1.f &&
           !(c.bits == 0 && c.load_factor > 0.f && c.load_factor < 0.05f);
}

std::vector<pj_batch_cfg> cfgs_for(pj_batch_s* h, const std::vector<pj_batch_join>& joins)
{
    std::vector<pj_batch_cfg> c(joins.size());
    for (size_t i = 0; i < joins.size(); ++i) {
        const uint32_t cls = size_class(joins[i].nbuild);
        c[i] = h->tuned[cls] ? h->cfg[cls] : default_cfg();
    }
    return c;
}

int device_cus()
{
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            cus = prop.multiProcessorCount;
        if (cus <= 0)
            cus = 256;
    }
    return cus;
}

void drop_pipeline_graph(pj_batch_s* h)
{
    if (h->exec) {
        (void)hipGraphExecDestroy(h->exec);
        h->exec = nullptr;
    }
}

// Plans the pipeline for a list of joins under per-join configurations:
// the sides' tiles, histogram counters, record ranges and partition starts,
// the joins' plan views and the bound on their work items; checks every
// count against the 32-bit offsets the kernels use, then uploads the
// descriptors.
int plan_pipeline(pj_batch_s* h, const std::vector<pj_batch_join>& joins, const std::vector<pj_batch_cfg>& cfgs,
                  hipStream_t st)
{
    const uint32_t nj = (uint32_t)joins.size();
    MBatch& b = h->mb;
    std::vector<MSide> sides;
    std::vector<MJoin> mj(nj);
    std::vector<uint64_t> obase(nj);
    std::vector<uint32_t> pstart(2 * nj);   // first partition start of each join's build and probe side
    h->bits.assign(nj, 0);
    uint64_t blocks = 0, nhist = 0, nrec = 0, nstarts = 0, noffs = 0, max_items = 0;
    uint32_t maxF = 1;
    bool ident = h->lay.build.ident && h->lay.probe.ident;
    const uint32_t cus = (uint32_t)device_cus();
    for (uint32_t j = 0; j < nj; ++j) {
        const pj_batch_join& q = joins[j];
        const uint32_t bits = cfg_bits(cfgs[j], q.nbuild);
        const uint32_t F = 1u << bits;
        const uint64_t piece = cfg_piece(cfgs[j], q.nprobe, cus);
        h->bits[j] = bits;
        maxF = std::max(maxF, F);
        for (uint32_t side = 0; side < 2; ++side) {
            const uint64_t n = side ? q.nprobe : q.nbuild;
            const void* in = side ? q.probe : q.build;
            if (n >= (1ull << 31))
                return fail(PJ_ERR_UNSUPPORTED, "join %u: a side of 2^31 tuples or more", j);
            if ((uintptr_t)in % 16)
                ident = false;
            MSide S = {};
            S.in = (const uint32_t*)in;
            S.n = (uint32_t)n;
            S.bits = bits;
            S.probe = side;
            S.nblk = (uint32_t)((n + kBTile - 1) / kBTile);
            S.chunk = kBTile;
            S.blk0 = (uint32_t)blocks;
            S.hbase = (uint32_t)nhist;
            S.rbase = (uint32_t)nrec;
            S.pbase = (uint32_t)nstarts;
            pstart[2 * j + side] = S.pbase;
            blocks += S.nblk;
            nhist += (uint64_t)F * S.nblk;
            nrec += n;
            nstarts += F + 1;
            sides.push_back(S);
            if (blocks >= (1ull << 31) || nhist >= (1ull << 32) || nrec >= (1ull << 32) || nstarts >= (1ull << 32))
                return fail(PJ_ERR_UNSUPPORTED, "a batch holds fewer than 2^32 tuples, histogram counters and "
                                                "partition starts");
        }
        MJoin& J = mj[j];
        J.v = PlanView{};
        J.v.chunk = piece;
        J.v.FR = F;
        J.v.shift = 0;
        J.out = (uint32_t*)q.out;
        J.out_cap = q.out_cap;
        J.bits = bits;
        obase[j] = noffs;
        noffs += F + 1;
        // items: table splits times probe pieces, summed over partitions
        const uint64_t splits = (q.nbuild + kTableCap - 1) / kTableCap + F;
        const uint64_t pieces = (q.nprobe + piece - 1) / piece + F;
        max_items += splits * pieces;
    }
    if (max_items >= (1ull << 32))
        return fail(PJ_ERR_UNSUPPORTED, "a batch holds fewer than 2^32 work items");
    b.njoins = nj;
    b.nsides = (uint32_t)sides.size();
    b.nblocks = (uint32_t)blocks;
    b.maxF = maxF;
    b.nhist = nhist;
    b.nrec = nrec;
    b.nstarts = nstarts;
    b.noffs = noffs;
    b.max_items = max_items;
    b.rm = ident ? REC_IDENT : REC_WORDS;
    b.lay = h->lay;
    const uint32_t nsb = (uint32_t)((nhist + kScanSeg - 1) / kScanSeg);
    int rc;
    if ((rc = b.sides.ensure(sides.size() * sizeof(MSide))) || (rc = b.joins.ensure(nj * sizeof(MJoin))) ||
        (rc = b.hist.ensure(std::max<uint64_t>(nhist, 1) * 4)) || (rc = b.bsum.ensure(((size_t)nsb + 2) * 4)) ||
        (rc = b.recs.ensure(std::max<uint64_t>(nrec, 1) * 16)) || (rc = b.starts.ensure(nstarts * 4)) ||
        (rc = b.offs.ensure(noffs * 4)) || (rc = b.ibase.ensure(((size_t)nj + 1) * 4)) ||
        (rc = b.state.ensure(nj * sizeof(MState))) || (rc = b.misc.ensure(64)))
        return rc;
    for (uint32_t j = 0; j < nj; ++j) {
        MJoin& J = mj[j];
        J.v.offs = (const uint32_t*)b.offs.p + obase[j];
        J.v.rst = (const uint32_t*)b.starts.p + pstart[2 * j];
        J.v.sst = (const uint32_t*)b.starts.p + pstart[2 * j + 1];
    }
    PJ_HIP(hipMemcpyAsync(b.sides.p, sides.data(), sides.size() * sizeof(MSide), hipMemcpyHostToDevice, st));
    PJ_HIP(hipMemcpyAsync(b.joins.p, mj.data(), nj * sizeof(MJoin), hipMemcpyHostToDevice, st));
    PJ_HIP(hipStreamSynchronize(st));   // the host arrays are locals
    return mbatch_setup(b);
}

// Plans the pipeline and captures it into a graph.
int prepare_pipeline(pj_batch_s* h, const std::vector<pj_batch_join>& joins, const std::vector<pj_batch_cfg>& cfgs,
                     hipStream_t st)
{
    drop_pipeline_graph(h);
    h->prepared = false;
    int rc = plan_pipeline(h, joins, cfgs, st);
    if (rc)
        return rc;
    hipStream_t cs;
    PJ_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        rc = mbatch_enqueue(h->mb, cs);
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(cs, &g);
        if (rc == PJ_OK && e == hipSuccess)
            e = hipGraphInstantiate(&h->exec, g, nullptr, nullptr, 0);
        if (g)
            (void)hipGraphDestroy(g);
    }
    (void)hipStreamDestroy(cs);
    if (rc)
        return rc;
    if (e != hipSuccess) {
        h->exec = nullptr;
        return fail(PJ_ERR_DEVICE, "graph capture of the batch pipeline failed: %s", hipGetErrorString(e));
    }
    h->joins = joins;
    h->prepared = true;
    return PJ_OK;
}

// Rows, completion times and the error flag of the last pipeline run.
int pipeline_results(pj_batch_s* h, uint64_t* out_counts, double* done_us, hipStream_t st)
{
    const uint32_t nj = h->mb.njoins;
    std::vector<MState> s(nj);
    uint32_t misc[4] = {0, 0, 0, 0};
    PJ_HIP(hipMemcpyAsync(s.data(), h->mb.state.p, nj * sizeof(MState), hipMemcpyDeviceToHost, st));
    PJ_HIP(hipMemcpyAsync(misc, h->mb.misc.p, sizeof(misc), hipMemcpyDeviceToHost, st));
    PJ_HIP(hipStreamSynchronize(st));
    if (misc[1])
        return fail(PJ_ERR_DEVICE, "the batch's work items passed the planned bound; no join ran");
    const unsigned long long t0 = (unsigned long long)misc[2] | ((unsigned long long)misc[3] << 32);
    int rc = PJ_OK;
    for (uint32_t i = 0; i < nj; ++i) {
        out_counts[i] = s[i].rows;
        if (done_us)   // the device clock of s_memrealtime runs at 100 MHz
            done_us[i] = s[i].t_done ? (double)(s[i].t_done - t0) / 100.0 : 0.0;
        if (s[i].rows > h->joins[i].out_cap && rc == PJ_OK)
            rc = fail(PJ_ERR_CAPACITY, "join %u produced %llu rows, its output holds %llu", i,
                      (unsigned long long)s[i].rows, (unsigned long long)h->joins[i].out_cap);
    }
    return rc;
}

int seq_handle(pj_batch_s* h, uint32_t bits, uint32_t lane, pj_join_t* out)
{
    auto it = h->seq.find({bits, lane});
    if (it != h->seq.end()) {
        *out = it->second;
        return PJ_OK;
    }
    pj_join_desc d = h->desc;
    d.partitioner = PJ_PART_RADIX;
    d.radix_bits = (int32_t)bits;
    d.flags = 0;
    int rc = pj_join_create(&d, out);
    if (rc)
        return rc;
    h->seq[{bits, lane}] = *out;
    return PJ_OK;
}

// Radix bits of a sequential-mode join: its size class's tuned bits, else
// the bits the single-join path picks by itself.
uint32_t seq_bits(const pj_batch_s* h, uint64_t nb)
{
    const uint32_t c = size_class(nb);
    return h->tuned[c] ? cfg_bits(h->cfg[c], nb) : choose_build_bits(nb, -1);
}

// Drops the captured lane graph. Its nodes hold the addresses of the join
// handles' device buffers, which any other run or the tuner may grow.
void drop_lane_graph(pj_batch_s* h)
{
    if (h->lane_exec) {
        (void)hipGraphExecDestroy(h->lane_exec);
        h->lane_exec = nullptr;
    }
    h->lane_graph_bits.clear();
}

// Sets up the per-join state of the lane modes: lane streams, events, one
// pinned counter slot a join; then every handle (radix bits, lane) runs the
// largest build and probe sides it will meet, so that no buffer grows (and
// waits for the device, or allocates inside a graph capture) later.
int prepare_lanes(pj_batch_s* h, const std::vector<pj_batch_join>& joins, const std::vector<uint32_t>& bits,
                  uint32_t nlanes, hipStream_t st)
{
    const size_t nj = joins.size();
    while (h->lanes.size() < (nlanes > 1 ? nlanes : 0)) {
        hipStream_t ls;
        PJ_HIP(hipStreamCreateWithFlags(&ls, hipStreamNonBlocking));
        h->lanes.push_back(ls);
    }
    while (h->seq_ev.size() < 2 * nj + 1 + nlanes) {
        hipEvent_t e;
        PJ_HIP(hipEventCreate(&e));
        h->seq_ev.push_back(e);
    }
    if (h->seq_nslots < nj) {
        if (h->seq_slots)
            PJ_HIP(hipHostFree(h->seq_slots));
        h->seq_slots = nullptr;
        h->seq_nslots = 0;
        PJ_HIP(hipHostMalloc((void**)&h->seq_slots, nj * 8 * sizeof(uint32_t), hipHostMallocDefault));
        h->seq_nslots = nj;
    }
    int rc;
    // Per handle (radix bits, lane): the joins with the most build tuples,
    // the most probe tuples, and the most probe tuples below kAppendMin. The
    // graph mode captures the handles after these runs, which size every
    // buffer a handle's joins need: the probe sides of kAppendMin tuples or
    // more take the appended layout, the smaller ones the contiguous one,
    // and each layout's buffers grow with its largest probe side.
    struct Big {
        size_t build, probe, probe_small;
    };
    constexpr size_t kNone = ~(size_t)0;
    std::map<std::pair<uint32_t, uint32_t>, Big> big;
    for (size_t i = 0; i < nj; ++i) {
        const std::pair<uint32_t, uint32_t> key{bits[i], (uint32_t)(i % nlanes)};
        const bool small = joins[i].nprobe < kAppendMin;
        auto it = big.find(key);
        if (it == big.end()) {
            big[key] = {i, i, small ? i : kNone};
            continue;
        }
        Big& g = it->second;
        if (joins[i].nbuild > joins[g.build].nbuild)
            g.build = i;
        if (joins[i].nprobe > joins[g.probe].nprobe)
            g.probe = i;
        if (small && (g.probe_small == kNone || joins[i].nprobe > joins[g.probe_small].nprobe))
            g.probe_small = i;
    }
    bool waited = false;
    for (const auto& kv : big) {
        pj_join_t jh;
        if ((rc = seq_handle(h, kv.first.first, kv.first.second, &jh)))
            return rc;
        auto& w = h->seq_warm[kv.first];
        const hipStream_t ls = nlanes > 1 ? h->lanes[kv.first.second] : st;
        for (size_t i : {kv.second.build, kv.second.probe, kv.second.probe_small}) {
            if (i == kNone)
                continue;
            const pj_batch_join& q = joins[i];
            if (q.nbuild <= w.build && q.nprobe <= (q.nprobe < kAppendMin ? w.probe_small : w.probe))
                continue;
            if (!waited) {   // the inputs may still be in flight on st
                PJ_HIP(hipStreamSynchronize(st));
                waited = true;
            }
            uint64_t n = 0;
            rc = pj_join_run_device(jh, q.build, q.nbuild, q.probe, q.nprobe, q.out, q.out_cap, &n, ls);
            if (rc && rc != PJ_ERR_CAPACITY)
                return rc;
            w.build = std::max<uint64_t>(w.build, q.nbuild);
            if (q.nprobe < kAppendMin)
                w.probe_small = std::max<uint64_t>(w.probe_small, q.nprobe);
            else
                w.probe = std::max<uint64_t>(w.probe, q.nprobe);
        }
    }
    return PJ_OK;
}

// Enqueues every join as a whole radix join (partition both sides, join)
// with `bits` radix bits, nothing waiting in between: with one lane one after
// another on st; with more, join i on lane i % nlanes, each lane a stream of
// its own that starts where st is and that st waits for at the end. Each
// join's counters land in its own pinned slot. With `timed`, HIP events
// around each join give its own device time and its completion time within
// the batch (not inside a graph capture).
int enqueue_lanes(pj_batch_s* h, const std::vector<pj_batch_join>& joins, const std::vector<uint32_t>& bits,
                  uint32_t nlanes, hipStream_t st, bool timed, size_t* ran)
{
    const size_t nj = joins.size();
    hipEvent_t* ev = h->seq_ev.data();
    hipEvent_t* lane_end = ev + 2 * nj + 1;
    PJ_HIP(hipEventRecord(ev[0], st));
    if (nlanes > 1)
        for (uint32_t l = 0; l < nlanes; ++l)
            PJ_HIP(hipStreamWaitEvent(h->lanes[l], ev[0], 0));
    int rc = PJ_OK;
    *ran = 0;
    for (size_t i = 0; i < nj && rc == PJ_OK; ++i) {
        const pj_batch_join& q = joins[i];
        const uint32_t l = (uint32_t)(i % nlanes);
        const hipStream_t ls = nlanes > 1 ? h->lanes[l] : st;
        pj_join_t jh;
        if ((rc = seq_handle(h, bits[i], l, &jh)))
            break;
        if (timed)
            (void)hipEventRecord(ev[2 * i + 1], ls);
        rc = join_enqueue(jh, q.build, q.nbuild, q.probe, q.nprobe, q.out, q.out_cap, h->seq_slots + 8 * i, ls);
        if (timed)
            (void)hipEventRecord(ev[2 * i + 2], ls);
        *ran = i + 1;
    }
    if (nlanes > 1)
        for (uint32_t l = 0; l < nlanes; ++l) {
            PJ_HIP(hipEventRecord(lane_end[l], h->lanes[l]));
            PJ_HIP(hipStreamWaitEvent(st, lane_end[l], 0));
        }
    return rc;
}

// Counts from the slots, and with `timed` the events' times.
void collect_lanes(pj_batch_s* h, size_t nj, size_t ran, bool timed, std::vector<uint64_t>& counts,
                   std::vector<double>& done_us, std::vector<double>& run_us)
{
    counts.assign(nj, 0);
    done_us.assign(nj, 0.0);
    run_us.assign(nj, 0.0);
    hipEvent_t* ev = h->seq_ev.data();
    for (size_t i = 0; i < ran; ++i) {
        const uint32_t* c = h->seq_slots + 8 * i;
        counts[i] = (uint64_t)c[2] | ((uint64_t)c[3] << 32);
        if (!timed)
            continue;
        float ms = 0.f, t = 0.f;
        (void)hipEventElapsedTime(&ms, ev[2 * i + 1], ev[2 * i + 2]);
        (void)hipEventElapsedTime(&t, ev[0], ev[2 * i + 2]);
        run_us[i] = ms * 1e3;
        done_us[i] = t * 1e3;
    }
}

// The sequential and concurrent modes: prepare, enqueue, wait once, collect.
int run_lanes(pj_batch_s* h, const std::vector<pj_batch_join>& joins, const std::vector<uint32_t>& bits,
              uint32_t nlanes, hipStream_t st, std::vector<uint64_t>& counts, std::vector<double>& done_us,
              std::vector<double>& run_us)
{
    drop_lane_graph(h);
    int rc = prepare_lanes(h, joins, bits, nlanes, st);
    if (rc)
        return rc;
    size_t ran = 0;
    rc = enqueue_lanes(h, joins, bits, nlanes, st, true, &ran);
    PJ_HIP(hipStreamSynchronize(st));
    collect_lanes(h, joins.size(), ran, true, counts, done_us, run_us);
    return rc;
}

// The concurrent mode captured into one graph (lanes as parallel branches),
// captured again only when the joins or their radix bits change; one
// hipGraphLaunch runs the batch. No per-join times: done_us is the whole
// batch's time for every join.
int run_lanes_graph(pj_batch_s* h, const std::vector<pj_batch_join>& joins, const std::vector<uint32_t>& bits,
                    hipStream_t st, std::vector<uint64_t>& counts, std::vector<double>& done_us,
                    std::vector<double>& run_us)
{
    int rc;
    if (!h->lane_exec || h->lane_graph_bits != bits) {
        drop_lane_graph(h);
        if ((rc = prepare_lanes(h, joins, bits, kLanes, st)))
            return rc;
        PJ_HIP(hipStreamSynchronize(st));
        hipStream_t cs;
        PJ_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        size_t ran = 0;
        rc = PJ_OK;
        if (e == hipSuccess) {
            rc = enqueue_lanes(h, joins, bits, kLanes, cs, false, &ran);
            hipGraph_t g = nullptr;
            e = hipStreamEndCapture(cs, &g);
            if (rc == PJ_OK && e == hipSuccess)
                e = hipGraphInstantiate(&h->lane_exec, g, nullptr, nullptr, 0);
            if (g)
                (void)hipGraphDestroy(g);
        }
        (void)hipStreamDestroy(cs);
        if (rc)
            return rc;
        if (e != hipSuccess) {
            h->lane_exec = nullptr;
            return fail(PJ_ERR_DEVICE, "graph capture of the batch failed: %s", hipGetErrorString(e));
        }
        if (ran != joins.size())
            return fail(PJ_ERR_DEVICE, "graph capture stopped after %zu joins", ran);
        h->lane_graph_bits = bits;
    }
    hipEvent_t* ev = h->seq_ev.data();
    PJ_HIP(hipEventRecord(ev[0], st));
    PJ_HIP(hipGraphLaunch(h->lane_exec, st));
    PJ_HIP(hipEventRecord(ev[1], st));
    PJ_HIP(hipStreamSynchronize(st));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
    collect_lanes(h, joins.size(), joins.size(), false, counts, done_us, run_us);
    for (auto& d : done_us)
        d = ms * 1e3;
    return PJ_OK;
}

}  // namespace

extern "C" {

int pj_batch_create(const pj_join_desc* desc, pj_batch_t* out)
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
                    "batched joins take one INTEGER / LONG / DATE key or two INTEGER keys, at most 8 payload bytes "
                    "a side and output tuples of at most 32 bytes without CHAR columns");
    pj_batch_s* h = new (std::nothrow) pj_batch_s();
    if (!h)
        return fail(PJ_ERR_INVALID, "out of host memory");
    h->desc = *desc;
    h->lay = L;
    for (uint32_t c = 0; c < kClasses; ++c) {
        h->cfg[c] = default_cfg();
        h->tuned[c] = false;
    }
    *out = h;
    return PJ_OK;
}

int pj_batch_destroy(pj_batch_t h)
{
    if (!h)
        return PJ_OK;
    for (auto& kv : h->seq)
        pj_join_destroy(kv.second);
    if (h->lane_exec)
        (void)hipGraphExecDestroy(h->lane_exec);
    drop_pipeline_graph(h);
    for (hipStream_t ls : h->lanes)
        (void)hipStreamDestroy(ls);
    for (hipEvent_t e : h->seq_ev)
        (void)hipEventDestroy(e);
    if (h->seq_slots)
        (void)hipHostFree(h->seq_slots);
    delete h;
    return PJ_OK;
}

int pj_batch_set_config(pj_batch_t h, uint32_t size_class, const pj_batch_cfg* cfg)
{
    if (!h || !cfg || size_class >= kClasses)
        return fail(PJ_ERR_INVALID, "bad size class or configuration");
    if (cfg->passes != 1)
        return fail(PJ_ERR_UNSUPPORTED, "one partition pass: the pipeline partitions a side into at most 2^%u "
                                        "partitions in one tile-sorted pass", kBMaxBits);
    if (!cfg_valid(*cfg))
        return fail(PJ_ERR_INVALID, "bad configuration: radix bits 0..%u (0: from the load factor), load factor "
                                    "0.05..1 or 0", kBMaxBits);
    std::lock_guard<std::mutex> g(h->mu);
    h->cfg[size_class] = *cfg;
    h->tuned[size_class] = true;
    return PJ_OK;
}

int pj_batch_get_config(pj_batch_t h, uint32_t size_class, pj_batch_cfg* cfg)
{
    if (!h || !cfg || size_class >= kClasses)
        return fail(PJ_ERR_INVALID, "bad size class");
    *cfg = h->cfg[size_class];
    return PJ_OK;
}

int pj_batch_prepare(pj_batch_t h, uint32_t njoins, const pj_batch_join* joins, void* stream)
{
    if (!h || (njoins && !joins) || njoins == 0)
        return fail(PJ_ERR_INVALID, "need at least one join");
    for (uint32_t i = 0; i < njoins; ++i) {
        const pj_batch_join& q = joins[i];
        if ((q.nbuild && !q.build) || (q.nprobe && !q.probe) || (q.out_cap && !q.out) || ((uintptr_t)q.out % 16))
            return fail(PJ_ERR_INVALID, "join %u: null or misaligned buffer", i);
    }
    std::lock_guard<std::mutex> g(h->mu);
    hipStream_t st = (hipStream_t)stream;
    std::vector<pj_batch_join> js(joins, joins + njoins);
    h->last_sequential = false;
    h->ran_pipeline = false;
    h->seq_counts.clear();
    drop_lane_graph(h);   // captured for the previous list
    if (!kPipelineOn) {   // only the modes of whole single joins run: keep the list
        h->joins = js;
        h->prepared = true;
        return PJ_OK;
    }
    return prepare_pipeline(h, js, cfgs_for(h, js), st);
}

int pj_batch_run(pj_batch_t h, int mode, void* stream)
{
    if (!h || mode < PJ_BATCH_PLAIN || mode > PJ_BATCH_CONCURRENT_GRAPH)
        return fail(PJ_ERR_INVALID, "null handle or unknown mode");
    std::lock_guard<std::mutex> g(h->mu);
    if (!h->prepared)
        return fail(PJ_ERR_STATE, "pj_batch_run before pj_batch_prepare");
    hipStream_t st = (hipStream_t)stream;
    if (mode == PJ_BATCH_SEQUENTIAL || mode == PJ_BATCH_CONCURRENT || mode == PJ_BATCH_CONCURRENT_GRAPH) {
        std::vector<uint32_t> bits(h->joins.size());
        for (size_t i = 0; i < bits.size(); ++i)
            bits[i] = seq_bits(h, h->joins[i].nbuild);
        const uint32_t nlanes = mode == PJ_BATCH_SEQUENTIAL ? 1u : kLanes;
        h->ran_pipeline = false;
        const int rc = mode == PJ_BATCH_CONCURRENT_GRAPH
                           ? run_lanes_graph(h, h->joins, bits, st, h->seq_counts, h->seq_done_us, h->seq_run_us)
                           : run_lanes(h, h->joins, bits, nlanes, st, h->seq_counts, h->seq_done_us, h->seq_run_us);
        h->last_sequential = rc == PJ_OK;
        return rc;
    }
    if (!kPipelineOn)
        return fail(PJ_ERR_UNSUPPORTED, "the batched pipeline is switched off: its first device run did not finish "
                                         "and the cause is not known (PJ_BATCH_SEQUENTIAL, PJ_BATCH_CONCURRENT and "
                                         "PJ_BATCH_CONCURRENT_GRAPH run)");
    h->last_sequential = false;
    h->ran_pipeline = true;
    if (mode == PJ_BATCH_GRAPH) {
        if (!h->exec)
            return fail(PJ_ERR_STATE, "no captured graph");
        PJ_HIP(hipGraphLaunch(h->exec, st));
        return PJ_OK;
    }
    return mbatch_enqueue(h->mb, st);
}

int pj_batch_results(pj_batch_t h, uint64_t* out_counts, double* done_us, void* stream)
{
    if (!h || !out_counts)
        return fail(PJ_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(h->mu);
    if (!h->prepared)
        return fail(PJ_ERR_STATE, "no prepared batch");
    const uint32_t nj = (uint32_t)h->joins.size();
    if (h->ran_pipeline)
        return pipeline_results(h, out_counts, done_us, (hipStream_t)stream);
    if (!h->last_sequential || h->seq_counts.size() != nj)
        return fail(PJ_ERR_STATE, "no completed run");
    int rc = PJ_OK;
    for (uint32_t i = 0; i < nj; ++i) {
        out_counts[i] = h->seq_counts[i];
        if (done_us)
            done_us[i] = h->seq_done_us[i];
        if (h->seq_counts[i] > h->joins[i].out_cap && rc == PJ_OK)
            rc = fail(PJ_ERR_CAPACITY, "join %u produced %llu rows, its output holds %llu", i,
                      (unsigned long long)h->seq_counts[i], (unsigned long long)h->joins[i].out_cap);
    }
    return rc;
}

int pj_batch_run_times(pj_batch_t h, double* run_us)
{
    if (!h || !run_us)
        return fail(PJ_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(h->mu);
    if (!h->last_sequential || h->seq_run_us.size() != h->joins.size())
        return fail(PJ_ERR_STATE, "no completed sequential or concurrent run");
    for (size_t i = 0; i < h->seq_run_us.size(); ++i)
        run_us[i] = h->seq_run_us[i];
    return PJ_OK;
}

// Times the candidate configurations of every size class present in the
// batch on up to `trial` of its joins (the batch's own buffers) and keeps
// the fastest. Through the pipeline (when it is on): radix bits from one
// below to two above the single join's choice, each with the default probe
// records a work item and with 32K and 256K.
int pj_batch_tune(pj_batch_t h, uint32_t njoins, const pj_batch_join* joins, uint32_t trial, void* stream)
{
    if (!h || !joins || njoins == 0 || trial == 0)
        return fail(PJ_ERR_INVALID, "bad arguments");
    hipStream_t st = (hipStream_t)stream;
    std::lock_guard<std::mutex> g(h->mu);
    drop_lane_graph(h);   // a later lane run captures again
    std::vector<std::vector<pj_batch_join>> by(kClasses);
    for (uint32_t i = 0; i < njoins; ++i) {
        const uint32_t c = size_class(joins[i].nbuild);
        if (by[c].size() < trial)
            by[c].push_back(joins[i]);
    }
    if (!kPipelineOn) {
        // the pipeline is off: tune the radix bits of the sequential mode,
        // from the bits the single join picks by itself (at least 4) to three
        // more, on the class's joins run one after another
        for (uint32_t c = 0; c < kClasses; ++c) {
            if (by[c].empty())
                continue;
            uint64_t nmax = 0;
            for (const auto& q : by[c])
                nmax = std::max<uint64_t>(nmax, q.nbuild);
            const uint32_t b0 = std::max<uint32_t>(4, std::min(kBMaxBits, choose_build_bits(nmax, -1)));
            double best = 1e300;
            pj_batch_cfg bestc = default_cfg();
            for (uint32_t b = b0; b <= std::min(kBMaxBits, b0 + 3); ++b) {
                std::vector<uint32_t> bits(by[c].size(), b);
                std::vector<uint64_t> cnt;
                std::vector<double> done, run;
                double t = 1e300;
                for (int rep = 0; rep < 3; ++rep) {   // the first run warms up
                    int rc = run_lanes(h, by[c], bits, 1, st, cnt, done, run);
                    if (rc)
                        return rc;
                    if (rep)
                        t = std::min(t, done.back());
                }
                if (t < best) {
                    best = t;
                    bestc.bits = (uint8_t)b;
                }
            }
            h->cfg[c] = bestc;
            h->tuned[c] = true;
        }
        return PJ_OK;
    }
    const uint16_t pieces[3] = {0, 32768 / kPieceUnit, 262144 / kPieceUnit};
    for (uint32_t c = 0; c < kClasses; ++c) {
        if (by[c].empty())
            continue;
        uint64_t nmax = 0;
        for (const auto& q : by[c])
            nmax = std::max<uint64_t>(nmax, q.nbuild);
        const uint32_t b0 = std::max(kBMinBits, std::min(kBMaxBits, choose_build_bits(nmax, -1)));
        double best = 1e300;
        pj_batch_cfg bestc = default_cfg();
        for (uint32_t b = b0 > kBMinBits ? b0 - 1 : b0; b <= std::min(kBMaxBits, b0 + 2); ++b)
            for (uint16_t pc : pieces) {
                pj_batch_cfg cand = default_cfg();
                cand.bits = (uint8_t)b;
                cand.piece = pc;
                std::vector<pj_batch_cfg> cc(by[c].size(), cand);
                int rc = plan_pipeline(h, by[c], cc, st);
                if (rc)
                    return rc;
                double t = 1e300;
                for (int rep = 0; rep < 3; ++rep) {   // the first run warms up
                    PJ_HIP(hipStreamSynchronize(st));
                    const auto a = std::chrono::steady_clock::now();
                    if ((rc = mbatch_enqueue(h->mb, st)))
                        return rc;
                    PJ_HIP(hipStreamSynchronize(st));
                    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
                    if (rep)
                        t = std::min(t, dt);
                }
                if (t < best) {
                    best = t;
                    bestc = cand;
                }
            }
        h->cfg[c] = bestc;
        h->tuned[c] = true;
    }
    drop_pipeline_graph(h);
    h->prepared = false;   // the trial batches are gone: prepare the real one again
    h->ran_pipeline = false;
    return PJ_OK;
}

}  // extern "C"
