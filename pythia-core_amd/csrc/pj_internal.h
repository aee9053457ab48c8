// pj_internal.h -- shared declarations of libpjoin: tuple layouts, the
// canonical 16-byte device record, error plumbing and small device helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/pjoin.h"
#include "pj_datagen.h"

namespace pj {

// ---------------------------------------------------------------- errors
void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);

#define PJ_HIP(call)                                                              \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess)                                                     \
            return ::pj::fail(PJ_ERR_DEVICE, "%s failed: %s (%s:%d)", #call,       \
                              hipGetErrorString(e_), __FILE__, __LINE__);          \
    } while (0)

// ---------------------------------------------------------------- layouts
// Where the canonical record's words come from in a packed tuple of one
// side. A canonical record is four 32-bit words {key_lo, key_hi, pay0, pay1}:
// the join key normalised to 64 bits (an INTEGER key is sign-extended, so
// INTEGER/LONG keys compare as the reference's IntLongEqual does; two
// INTEGER key columns are packed lo/hi), and up to two words of the
// side's projected non-key columns.
struct SideLayout {
    uint32_t tw;   // tuple width in 32-bit words
    int32_t k0;    // word index of key_lo
    int32_t k1;    // word index of key_hi, or -1: sign-extend key_lo
    int32_t p0;    // word index of pay0, or -1: zero
    int32_t p1;    // word index of pay1, or -1: zero
    uint32_t ident;  // 1 when the tuple is exactly the record (tw 4, 0,1,2,3)
};

// Output tuple assembly: word w of an output tuple is word src[w] of the
// 8-word concatenation {build record, probe record}.
enum { SRC_BK0 = 0, SRC_BK1, SRC_BP0, SRC_BP1, SRC_PK0, SRC_PK1, SRC_PP0, SRC_PP1 };
struct OutFormat {
    uint32_t words;      // output tuple width in words (<= 8)
    uint8_t src[8];
};

struct Layout {
    SideLayout build, probe;
    OutFormat out;
    uint32_t out_width;   // bytes
};

// Validates a descriptor and derives the layouts (see pj_join.hip).
int derive_layout(const pj_join_desc* d, Layout* L);

// ---------------------------------------------------------------- device
constexpr uint32_t kJoinThreads = 1024;
constexpr uint32_t kTableCap = 8944;       // records per LDS table (~140 KiB)
constexpr uint32_t kTableBuckets = 4096;   // bucket directory entries
constexpr uint32_t kLdsBytes = 163840;     // gfx950 LDS per CU
constexpr uint32_t kScatterThreads = 512;

struct Item {           // one join work item
    uint32_t r_off, r_cnt;   // build records [r_off, r_off + r_cnt) of R_part
    uint32_t s_off, s_cnt;   // probe records of S_part
    uint32_t part;           // build partition id; records with other ids are skipped
    uint32_t unused;
};

}  // namespace pj
