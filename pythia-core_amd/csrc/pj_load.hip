// pj_load.hip -- raw binary tables from files into HBM.
//
// The reference's binary scan source is MemMappedTable::load
// (operators/loaders/table.cpp:170-347): glob() the pattern, take the
// matches in glob's sorted order, skip directories, empty and non-regular
// files, and map each remaining file whole as packed tuples with no header
// (one LinkedTupleBuffer a file). Here the same files land back to back in
// one device buffer: reader threads pread() fixed-size chunks straight into
// pinned staging slots while earlier slots are copied to the device, so disk
// (or page cache) reads overlap host-to-device copies; a slot is refilled
// only after its previous copy has completed (an event a slot).
#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <glob.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "pj_internal.h"

namespace pj {
namespace {

constexpr uint32_t kLoadSlots = 8;
constexpr size_t kLoadChunk = 32u << 20;   // bytes a slot (a multiple of 4 KiB)

struct TableFile {
    std::string path;
    uint64_t bytes;
};

// The files MemMappedTable::doload would map, in its order (glob's sorted
// order; the permuted order of GLOB_NOSORT is not reproduced).
int list_files(const char* pattern, uint32_t width, std::vector<TableFile>& files, uint64_t* total)
{
    files.clear();
    *total = 0;
    glob_t g;
    const int rc = glob(pattern, GLOB_MARK, nullptr, &g);
    if (rc == GLOB_NOMATCH)
        return PJ_OK;   // no files: an empty table
    if (rc != 0)
        return fail(PJ_ERR_INVALID, "glob(\"%s\") failed (%d)", pattern, rc);
    int err = PJ_OK;
    for (size_t i = 0; i < g.gl_pathc && err == PJ_OK; ++i) {
        const std::string name = g.gl_pathv[i];
        if (name.empty() || name.back() == '/')
            continue;   // a directory
        struct stat st;
        if (stat(name.c_str(), &st) != 0) {
            err = fail(PJ_ERR_INVALID, "stat(\"%s\") failed", name.c_str());
            break;
        }
        if (!S_ISREG(st.st_mode) || st.st_size == 0)
            continue;
        if ((uint64_t)st.st_size % width) {
            err = fail(PJ_ERR_INVALID, "\"%s\" holds %llu bytes, not a whole number of %u-byte tuples", name.c_str(),
                       (unsigned long long)st.st_size, width);
            break;
        }
        files.push_back({name, (uint64_t)st.st_size});
        *total += (uint64_t)st.st_size;
    }
    globfree(&g);
    return err;
}

// Pinned staging slots, kept for the life of the process (allocating
// 256 MiB of pinned memory costs tens of milliseconds).
struct Staging {
    std::mutex mu;
    int device = -1;
    char* slot[kLoadSlots] = {};
    hipEvent_t done[kLoadSlots] = {};
    bool ready = false;

    int init()
    {
        int dev = 0;
        PJ_HIP(hipGetDevice(&dev));
        if (ready && dev == device)
            return PJ_OK;
        if (ready)
            return fail(PJ_ERR_STATE, "the loader's staging buffers belong to device %d, not %d", device, dev);
        for (uint32_t s = 0; s < kLoadSlots; ++s) {
            PJ_HIP(hipHostMalloc((void**)&slot[s], kLoadChunk, hipHostMallocDefault));
            PJ_HIP(hipEventCreateWithFlags(&done[s], hipEventDisableTiming));
        }
        device = dev;
        ready = true;
        return PJ_OK;
    }
};

Staging& staging()
{
    static Staging* s = new Staging();   // never destroyed: pinned memory outlives static destructors
    return *s;
}

struct Chunk {
    uint32_t file;
    uint64_t off, len, dst;
};

}  // namespace
}  // namespace pj

using namespace pj;

extern "C" {

int pj_binary_tuples(const char* pattern, uint32_t tuple_width, uint64_t* ntuples)
{
    if (!pattern || !ntuples || tuple_width == 0)
        return fail(PJ_ERR_INVALID, "null argument or zero tuple width");
    std::vector<TableFile> files;
    uint64_t total = 0;
    const int rc = list_files(pattern, tuple_width, files, &total);
    *ntuples = total / tuple_width;
    return rc;
}

int pj_load_binary(const char* pattern, uint32_t tuple_width, void* d_dst, uint64_t cap, uint64_t* ntuples,
                   void* stream)
{
    if (!pattern || !ntuples || tuple_width == 0)
        return fail(PJ_ERR_INVALID, "null argument or zero tuple width");
    std::vector<TableFile> files;
    uint64_t total = 0;
    int rc = list_files(pattern, tuple_width, files, &total);
    *ntuples = total / tuple_width;
    if (rc)
        return rc;
    if (*ntuples > cap)
        return fail(PJ_ERR_CAPACITY, "%llu tuples, the buffer holds %llu", (unsigned long long)*ntuples,
                    (unsigned long long)cap);
    if (total == 0)
        return PJ_OK;
    if (!d_dst)
        return fail(PJ_ERR_INVALID, "null device buffer");
    std::vector<Chunk> chunks;
    uint64_t dst = 0;
    for (uint32_t f = 0; f < files.size(); ++f)
        for (uint64_t off = 0; off < files[f].bytes; off += kLoadChunk) {
            const uint64_t len = std::min<uint64_t>(kLoadChunk, files[f].bytes - off);
            chunks.push_back({f, off, len, dst});
            dst += len;
        }
    Staging& S = staging();
    std::lock_guard<std::mutex> g(S.mu);
    if ((rc = S.init()))
        return rc;
    hipStream_t st = (hipStream_t)stream;
    std::vector<int> fds(files.size(), -1);
    for (uint32_t f = 0; f < files.size(); ++f) {
        fds[f] = open(files[f].path.c_str(), O_RDONLY);
        if (fds[f] < 0) {
            rc = fail(PJ_ERR_INVALID, "open(\"%s\") failed", files[f].path.c_str());
            break;
        }
    }
    // Reader r fills slot r with chunks r, r + kLoadSlots, ...; the copies of
    // all readers go to the caller's stream, so later work on it sees them.
    // Error messages are per thread: readers keep the first failure here and
    // the calling thread reports it.
    std::atomic<int> err{rc};
    std::mutex err_mu;
    std::string err_msg;
    auto bad = [&](int code, const std::string& m) {
        std::lock_guard<std::mutex> eg(err_mu);
        if (err.load() == PJ_OK) {
            err_msg = m;
            err = code;
        }
    };
    std::mutex copy_mu;   // one enqueue at a time keeps the copy and its event paired
    auto reader = [&](uint32_t r) {
        for (size_t c = r; c < chunks.size() && err.load() == PJ_OK; c += kLoadSlots) {
            const Chunk& k = chunks[c];
            if (hipEventSynchronize(S.done[r]) != hipSuccess)
                return bad(PJ_ERR_DEVICE, "waiting for a staging copy failed");
            uint64_t got = 0;
            while (got < k.len) {
                const ssize_t n = pread(fds[k.file], S.slot[r] + got, k.len - got, (off_t)(k.off + got));
                if (n <= 0)
                    return bad(PJ_ERR_INVALID, "reading \"" + files[k.file].path + "\" failed at byte " +
                                                   std::to_string(k.off + got));
                got += (uint64_t)n;
            }
            std::lock_guard<std::mutex> cg(copy_mu);
            if (hipMemcpyAsync((char*)d_dst + k.dst, S.slot[r], k.len, hipMemcpyHostToDevice, st) != hipSuccess ||
                hipEventRecord(S.done[r], st) != hipSuccess)
                return bad(PJ_ERR_DEVICE, "enqueueing a host-to-device copy failed");
        }
    };
    if (err.load() == PJ_OK) {
        std::vector<std::thread> th;
        const uint32_t nr = (uint32_t)std::min<size_t>(kLoadSlots, chunks.size());
        for (uint32_t r = 0; r < nr; ++r)
            th.emplace_back(reader, r);
        for (auto& t : th)
            t.join();
    }
    for (int fd : fds)
        if (fd >= 0)
            close(fd);
    if (hipStreamSynchronize(st) != hipSuccess && err.load() == PJ_OK)
        return fail(PJ_ERR_DEVICE, "host-to-device copy failed");
    if (err.load() != PJ_OK && !err_msg.empty())
        return fail(err.load(), "%s", err_msg.c_str());
    return err.load();
}

}  // extern "C"
