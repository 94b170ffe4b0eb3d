// cjx_host.h - host-side state of a context and the launch helpers the C
// ABI (cjx_api.cpp) and the request server (cjx_server.cpp) share.  Not part
// of the public ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "../../include/cujson_amd.h"
#include "cjx_internal.h"

namespace cjx_host {

constexpr uint64_t kMaxLaunchBytes = 0x7FFFFFFFull;  // positions and indices stay in int32

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    // Grows (never shrinks); contents are not preserved.
    bool ensure(size_t bytes) {
        if (bytes <= cap) return true;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        size_t want = bytes + bytes / 8 + 4096;
        if (hipMalloc(&p, want) != hipSuccess) {
            p = nullptr;
            return false;
        }
        cap = want;
        return true;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

struct PinnedBuf {
    void *p = nullptr;
    size_t cap = 0;
    // flags: hipHostMallocCoherent for memory that kernels read or write
    // directly (visible to the host without a cache write-back)
    bool ensure(size_t bytes, unsigned flags = hipHostMallocDefault) {
        if (bytes <= cap) return true;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        if (hipHostMalloc(&p, bytes, flags) != hipSuccess) {
            p = nullptr;
            return false;
        }
        cap = bytes;
        return true;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};

}  // namespace cjx_host

struct cjx_context {
    using DevBuf = cjx_host::DevBuf;
    using PinnedBuf = cjx_host::PinnedBuf;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // Timing ring: event triples (tokenizer start, tokenizer end, phase-3 end)
    // for the last kTimingRing parses while timing is enabled.
    static constexpr int kTimingRing = 256;
    hipEvent_t tev[kTimingRing][3] = {};
    uint64_t timing_count = 0;
    bool timing = false;
    cjx::TokShape tok_shape;  // tokenizer launch shape (cjx_context_set_tokenizer_shape)
    uint32_t n_tiny = 0;      // small units of at most one round, first in the small list
    DevBuf in, structural, pair;
    DevBuf span;  // host units copied as one span of their source, before placement
    DevBuf units, uout, ls, tile_st, layout, chunk_unit, big_units, slow_units, unit_part,
        rec_count, recs;
    uint32_t *tile_unit = nullptr;    // inside `units`, after the unit descriptors
    uint32_t *small_units = nullptr;  // after the tile -> unit map
    uint64_t *unit_src = nullptr;     // after those: source offsets of relocated units
    // Pinned upload slots for unit tables: a call may return before its copy
    // has run, so the next calls use other slots; a slot is reused only after
    // the event recorded behind its copy has completed.
    static constexpr int kUploadSlots = 8;
    PinnedBuf up[kUploadSlots];
    hipEvent_t up_ev[kUploadSlots] = {};
    uint64_t up_count = 0;
    DevBuf oc_char, oc_idx, keys, sorted, tmp, spos, hist;
    DevBuf ck, gmin, gmin1, gmin2, gagg, gd0, lows, highs;
    DevBuf row_nl;         // newline flag per structural row (record tables)
    bool want_nl = false;  // the launch being enqueued writes row_nl
    PinnedBuf h_units, h_uout, h_ls, h_stage;
    // Chunked host input: a second stream copies the input in parts while the
    // tokenizer runs on the parts already in (created on first use).
    static constexpr int kCopyEvents = 64;
    hipStream_t copy_stream = nullptr;
    hipEvent_t copy_ev[kCopyEvents] = {};
    hipEvent_t idle_ev = nullptr;
    // Structural rows of a chunked host parse copied out while later parts
    // are still coming in: each part's row count (its last tile's status)
    // lands in h_part, and the rows go out on out_stream.
    hipStream_t out_stream = nullptr;
    hipEvent_t part_ev[kCopyEvents] = {};
    // Work of a launch that does not depend on the kernels beside it runs on
    // a second stream, forked from and joined back into the context stream
    // by events (created on first use): the strict UTF-8 check beside the
    // tokenizer, and in document batches the small units' waves beside the
    // tiles and the sort path beside the chunk matcher.  Off (everything on
    // one stream) when fork_ok is false.
    bool fork_ok = true;
    hipStream_t aux = nullptr;
    hipEvent_t fork_ev[2] = {}, join_ev[2] = {};
    PinnedBuf h_part;
    // Newline scan of device text: segment counts, bases and status; the
    // newline offsets and gathered line starts of the device chunker.
    DevBuf nl_seg, nl_off, nl_idx;
    DevBuf val_exact;  // typed values: numbers for the exact step, and their count
    PinnedBuf h_nl;
    // File loads into device memory: pinned staging buffers, refilled from the
    // file while earlier ones are copied (created on first use).
    static constexpr int kLoadBufs = 4;
    static constexpr size_t kLoadPart = 32ull << 20;
    PinnedBuf load_buf[kLoadBufs];
    hipEvent_t load_ev[kLoadBufs] = {};
    std::vector<uint64_t> batch_off;  // byte offsets of the documents of the batch in flight
    // One-unit device launches (cjx_parse_device): the unit table is written
    // by the reset kernel from this descriptor instead of being uploaded.
    bool one_unit_dev = false;
    cjx::UnitDesc one_unit{};
    // Shape of the launch in flight (device form).
    uint32_t n_units = 0;
    bool events_recorded = false;
    // Captured launch sequence of the last graph-mode call and what it baked in.
    hipGraphExec_t graph = nullptr;
    struct GraphKey {
        cjx::LaunchBufs b;
        const uint8_t *d_in;
        uint32_t *d_struct;
        int32_t *d_pair;
        uint32_t ntiles, n_units, n_small;
        uint64_t rows, max_rows, chunks;
        unsigned flags;
        bool concat;
        bool one_unit_dev;
        cjx::UnitDesc one_unit;
        cjx::TokShape tok_shape;
        uint32_t n_tiny;
    } graph_key{};
};


namespace cjx_host {

using cjx::LaunchBufs;
using cjx::UnitDesc;
using cjx::UnitOut;

struct Shape {
    uint32_t ntiles, n_units, n_small;
    uint64_t rows;      // total rows of all output regions
    uint64_t max_rows;  // largest region
    uint64_t chunks;    // bound on the matching chunks of all units
    uint64_t groups;    // bound on their 64-chunk groups (per unit, rounded up)
    bool concat;        // pair values numbered across units (JSON Lines), else per unit
};

int set_device(cjx_context *ctx);
uint64_t assign_tiles(std::vector<UnitDesc> &units, uint32_t &n_small);
uint64_t set_regions(std::vector<UnitDesc> &units, bool full);
void shape_rows(Shape &sh, const std::vector<UnitDesc> &units);
int ensure_work(cjx_context *ctx, const Shape &sh);
int upload_units(cjx_context *ctx, const std::vector<UnitDesc> &units, uint32_t ntiles,
                 uint32_t n_small, const uint64_t *unit_src = nullptr);
LaunchBufs bufs_of(cjx_context *ctx);
int enqueue_parse(cjx_context *ctx, const uint8_t *d_in, const Shape &sh, unsigned flags,
                  uint32_t *d_struct, int32_t *d_pair);
int read_status(cjx_context *ctx, uint32_t n_units);
int ensure_copy_stream(cjx_context *ctx);
int unit_code(const UnitOut &o);

}  // namespace cjx_host
