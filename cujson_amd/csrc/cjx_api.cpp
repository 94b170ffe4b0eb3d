// cjx_api.cpp - host driver and C ABI of the MI355X JSON parse.
//
// Orchestration replaces parse_standard_json (cujson/parse_standard_json.cu:1584-1687)
// and parse_json_lines (cujson/parse_json_lines.cu:1092-1264): one H2D copy,
// one tokenizer launch over every unit, the phase-3 kernels per unit, one
// device synchronization, and D2H copies straight into the reference's host
// layout.  No host round trip sits between the phases: every count the later
// kernels need is read on the device.
#include <hip/hip_runtime.h>

#include <errno.h>
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <system_error>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>
#include <functional>
#include <thread>
#include <vector>

#include "../../include/cujson_amd.h"
#include "cjx_chunk.h"
#include "cjx_host.h"
#include "cjx_internal.h"

using namespace cjx;
using cjx_chunk::ChunkList;
using cjx_chunk::bytes_rule;
using cjx_chunk::chunk_rule_ok;
using cjx_chunk::count_rule;
using cjx_chunk::emit_chunks;

static_assert(sizeof(RecordOut) == sizeof(cjx_record), "record layout");
static_assert(sizeof(ValueOut) == sizeof(cjx_value), "value layout");
static_assert(sizeof(cjx_doc_result) == 32, "document result layout");

namespace cjx_host {

int set_device(cjx_context *ctx) {
    return hipSetDevice(ctx->device) == hipSuccess ? CJX_OK : CJX_ERR_DEVICE;
}

// Pinned host buffers for results.  Allocating and pinning a large host
// buffer per call costs about as much as the copy into it, so released
// result buffers are kept for reuse: a request takes the smallest cached
// buffer of at least its size and at most about twice it; at most
// kPoolBytes stay cached.  The pool is process-wide and never torn down (the
// runtime may be gone at exit).
struct PinnedPool {
    std::mutex mu;
    std::unordered_map<void *, size_t> cap;  // every buffer handed out or cached
    std::multimap<size_t, void *> idle;      // cached: capacity -> buffer
    size_t cached = 0;
};
constexpr size_t kPoolBytes = 8ull << 30;

PinnedPool &pinned_pool() {
    static PinnedPool *p = new PinnedPool();
    return *p;
}

void *pinned_get(size_t bytes) {
    PinnedPool &P = pinned_pool();
    bytes = std::max<size_t>(bytes, 64);
    {
        std::lock_guard<std::mutex> g(P.mu);
        auto it = P.idle.lower_bound(bytes);
        if (it != P.idle.end() && it->first <= 2 * bytes + (1u << 20)) {
            void *p = it->second;
            P.cached -= it->first;
            P.idle.erase(it);
            return p;
        }
    }
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> g(P.mu);
    P.cap[p] = bytes;
    return p;
}

void pinned_put(void *p) {
    if (!p) return;
    PinnedPool &P = pinned_pool();
    {
        std::lock_guard<std::mutex> g(P.mu);
        auto it = P.cap.find(p);
        if (it != P.cap.end()) {
            const size_t c = it->second;
            if (P.cached + c <= kPoolBytes) {
                P.idle.emplace(c, p);
                P.cached += c;
                return;
            }
            P.cap.erase(it);
        }
    }
    (void)hipHostFree(p);
}

// Tokenizer work of the units: a unit of at most one wave's bytes is small
// (one wave, no tiles); the others are cut into tiles.  Returns the tile
// count and fills n_small.
uint64_t assign_tiles(std::vector<UnitDesc> &units, uint32_t &n_small) {
    uint64_t tiles = 0;
    n_small = 0;
    for (UnitDesc &u : units) {
        u.tile0 = (uint32_t)tiles;
        if (u.n <= (uint64_t)kWaveBytes) {
            u.ntiles = 0;
            n_small++;
        } else {
            u.ntiles = (uint32_t)((u.n + kTileBytes - 1) / kTileBytes);
            tiles += u.ntiles;
        }
    }
    return tiles;
}

// Units are given as (host offset, size); they are placed in the device
// input at 16-byte aligned offsets.  Fills descriptors; returns false on an
// invalid shape.  Output regions are set by set_regions.
bool plan_units(const std::vector<uint64_t> &sizes, const std::vector<uint64_t> &pos_base,
                std::vector<UnitDesc> &units, uint64_t &dev_bytes, uint32_t &ntiles,
                uint32_t &n_small) {
    units.resize(sizes.size());
    uint64_t off = 0;
    for (size_t i = 0; i < sizes.size(); i++) {
        UnitDesc &u = units[i];
        u.in_off = off;
        u.n = sizes[i];
        u.pos_base = pos_base[i];
        off += (sizes[i] + 15) & ~15ull;
    }
    dev_bytes = off + 64;
    const uint64_t tiles = assign_tiles(units, n_small);
    ntiles = (uint32_t)tiles;
    return tiles <= 0xFFFFFFF0ull && !units.empty();
}

// Output regions: unit u owns rows [out_off, out_off + out_cap) of every
// per-structural and per-bracket buffer.  `full` gives each unit its exact
// bound (n rows); otherwise a unit gets n/3 + 256 rows (typical JSON has
// fewer than 0.15 structurals per byte) and the parse is retried with full
// regions when one overflows.  Returns the total row count.
uint64_t set_regions(std::vector<UnitDesc> &units, bool full) {
    uint64_t off = 0;
    for (UnitDesc &u : units) {
        const uint64_t cap = full ? u.n : std::min<uint64_t>(u.n, u.n / 3 + 256);
        u.out_off = (uint32_t)off;
        u.out_cap = (uint32_t)cap;
        off += cap;
    }
    return off;
}


// Fills the row totals of a shape from the units' regions.
void shape_rows(Shape &sh, const std::vector<UnitDesc> &units) {
    sh.rows = sh.max_rows = sh.chunks = 0;
    for (const UnitDesc &u : units) {
        sh.rows += u.out_cap;
        sh.max_rows = std::max<uint64_t>(sh.max_rows, u.out_cap);
        sh.chunks += (u.out_cap + kMatchChunk - 1) / kMatchChunk + 1;
    }
    sh.groups = sh.chunks / 64 + units.size();
}

int ensure_work(cjx_context *ctx, const Shape &sh) {
    bool ok = ctx->units.ensure(sh.n_units * sizeof(UnitDesc)) &&
              ctx->uout.ensure(sh.n_units * sizeof(UnitOut)) &&
              ctx->ls.ensure(sizeof(LaunchState)) &&
              ctx->tile_st.ensure((size_t)sh.ntiles * 16) && ctx->oc_char.ensure(sh.rows + 16) &&
              ctx->oc_idx.ensure(sh.rows * 4 + 16) && ctx->keys.ensure(sh.rows * 2 + 16) &&
              ctx->sorted.ensure(sh.rows * 4 + 16) && ctx->spos.ensure(sh.rows * 4 + 16) &&
              ctx->hist.ensure(512 * 256 * 4) &&
              ctx->ck.ensure(sh.chunks * sizeof(ChunkStat) + 16) &&
              ctx->gmin.ensure(sh.chunks * 4 + 16) && ctx->gmin1.ensure(sh.groups * 4 + 16) &&
              ctx->gmin2.ensure(sh.groups * 4 + 16) &&
              ctx->gagg.ensure(sh.groups * sizeof(DepthAgg) + 16) &&
              ctx->gd0.ensure(sh.groups * 4 + 16) &&
              ctx->lows.ensure(sh.chunks * kMatchKeys * 4 + 16) &&
              ctx->highs.ensure(sh.chunks * kMatchKeys * 4 + 16) &&
              ctx->chunk_unit.ensure(sh.chunks * 4 + 16) &&
              ctx->big_units.ensure((size_t)sh.n_units * 4 + 16) &&
              ctx->slow_units.ensure((size_t)sh.n_units * 4 + 16) &&
              ctx->unit_part.ensure(((size_t)sh.n_units / 256 + 2) * 32) &&
              ctx->h_units.ensure(sh.n_units * sizeof(UnitDesc)) &&
              ctx->h_uout.ensure(sh.n_units * sizeof(UnitOut)) &&
              ctx->h_ls.ensure(sizeof(LaunchState));
    return ok ? CJX_OK : CJX_ERR_ALLOC;
}

// Uploads the unit descriptors and the tile -> unit map (one copy through a
// pinned upload slot) into ctx->units.
// The unit table, the tile -> unit map, the small units and, for units that
// are relocated, their source offsets (unit_src, one per unit), in one copy.
int upload_units(cjx_context *ctx, const std::vector<UnitDesc> &units, uint32_t ntiles,
                 uint32_t n_small, const uint64_t *unit_src) {
    const size_t nb = units.size() * sizeof(UnitDesc);
    const size_t tb0 = ((size_t)ntiles + n_small) * 4;
    const size_t sb = unit_src ? units.size() * 8 : 0;
    const size_t tb = sb ? (tb0 + 7) & ~(size_t)7 : tb0;  // source offsets 8-byte aligned
    if (!ctx->units.ensure(nb + tb + sb + 16)) return CJX_ERR_ALLOC;
    const int slot = (int)(ctx->up_count++ % cjx_context::kUploadSlots);
    if (!ctx->up_ev[slot] && hipEventCreate(&ctx->up_ev[slot]) != hipSuccess)
        return CJX_ERR_DEVICE;
    if (hipEventSynchronize(ctx->up_ev[slot]) != hipSuccess) return CJX_ERR_DEVICE;
    if (!ctx->up[slot].ensure(nb + tb + sb + 16)) return CJX_ERR_ALLOC;
    uint8_t *h = static_cast<uint8_t *>(ctx->up[slot].p);
    std::memcpy(h, units.data(), nb);
    uint32_t *tu = reinterpret_cast<uint32_t *>(h + nb);
    // then the small units: those of at most one tokenizer round first (the
    // one-round kernel takes them), each part in unit order
    uint32_t *su = tu + ntiles;
    uint32_t n_tiny = 0;
    for (size_t u = 0; u < units.size(); u++) {
        for (uint32_t t = 0; t < units[u].ntiles; t++) tu[units[u].tile0 + t] = (uint32_t)u;
        if (units[u].ntiles == 0 && units[u].n <= (uint64_t)kRoundBytes) {
            *su++ = (uint32_t)u;
            n_tiny++;
        }
    }
    for (size_t u = 0; u < units.size(); u++)
        if (units[u].ntiles == 0 && units[u].n > (uint64_t)kRoundBytes) *su++ = (uint32_t)u;
    (void)n_small;
    ctx->n_tiny = n_tiny;
    if (sb) std::memcpy(h + nb + tb, unit_src, sb);
    if (hipMemcpyAsync(ctx->units.p, h, nb + tb + sb, hipMemcpyHostToDevice, ctx->stream) !=
        hipSuccess)
        return CJX_ERR_DEVICE;
    if (hipEventRecord(ctx->up_ev[slot], ctx->stream) != hipSuccess) return CJX_ERR_DEVICE;
    ctx->tile_unit = reinterpret_cast<uint32_t *>(ctx->units.as<uint8_t>() + nb);
    ctx->small_units = ctx->tile_unit + ntiles;
    ctx->unit_src = sb ? reinterpret_cast<uint64_t *>(ctx->units.as<uint8_t>() + nb + tb) : nullptr;
    return CJX_OK;
}

LaunchBufs bufs_of(cjx_context *ctx) {
    LaunchBufs b;
    b.units = ctx->units.as<UnitDesc>();
    b.tile_unit = ctx->tile_unit;
    b.small_units = ctx->small_units;
    b.chunk_unit = ctx->chunk_unit.as<uint32_t>();
    b.big_units = ctx->big_units.as<uint32_t>();
    b.slow_units = ctx->slow_units.as<uint32_t>();
    b.unit_part = ctx->unit_part.as<uint32_t>();
    b.uout = ctx->uout.as<UnitOut>();
    b.ls = ctx->ls.as<LaunchState>();
    b.tile_st = ctx->tile_st.as<unsigned long long>();
    b.oc_char = ctx->oc_char.as<uint8_t>();
    b.oc_idx = ctx->oc_idx.as<uint32_t>();
    b.keys = ctx->keys.as<uint16_t>();
    b.sorted = ctx->sorted.as<uint32_t>();
    b.tmp = ctx->tmp.p ? ctx->tmp.as<uint32_t>() : nullptr;
    b.spos = ctx->spos.as<uint32_t>();
    b.hist = ctx->hist.as<uint32_t>();
    b.ck = ctx->ck.as<ChunkStat>();
    b.gmin = ctx->gmin.as<int32_t>();
    b.gmin1 = ctx->gmin1.as<int32_t>();
    b.gmin2 = ctx->gmin2.as<int32_t>();
    b.gagg = ctx->gagg.as<DepthAgg>();
    b.gd0 = ctx->gd0.as<int32_t>();
    b.lows = ctx->lows.as<uint32_t>();
    b.highs = ctx->highs.as<uint32_t>();
    b.row_nl = ctx->want_nl ? ctx->row_nl.as<uint8_t>() : nullptr;
    return b;
}

// The second stream and its events (created on first use); false when the
// context runs everything on its own stream.
bool ensure_aux(cjx_context *ctx) {
    if (!ctx->fork_ok) return false;
    if (ctx->aux) return true;
    for (int i = 0; i < 2; i++)
        if ((!ctx->fork_ev[i] &&
             hipEventCreateWithFlags(&ctx->fork_ev[i], hipEventDisableTiming) != hipSuccess) ||
            (!ctx->join_ev[i] &&
             hipEventCreateWithFlags(&ctx->join_ev[i], hipEventDisableTiming) != hipSuccess))
            return false;
    if (hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking) != hipSuccess) {
        ctx->aux = nullptr;
        return false;
    }
    return true;
}

// aux waits for everything enqueued on the context stream so far.
bool fork_aux(cjx_context *ctx, int k) {
    return hipEventRecord(ctx->fork_ev[k], ctx->stream) == hipSuccess &&
           hipStreamWaitEvent(ctx->aux, ctx->fork_ev[k], 0) == hipSuccess;
}

// The context stream waits for everything enqueued on aux so far.
bool join_aux(cjx_context *ctx, int k) {
    return hipEventRecord(ctx->join_ev[k], ctx->aux) == hipSuccess &&
           hipStreamWaitEvent(ctx->stream, ctx->join_ev[k], 0) == hipSuccess;
}

// Phase 3 of a launch (unit bases, matching, sort path) on the context
// stream, after the tokenizer.
int enqueue_back(cjx_context *ctx, const Shape &sh, unsigned flags, int32_t *d_pair) {
    const bool u16 = (flags & CJX_FLAG_DEPTH_U16) != 0;
    LaunchBufs b = bufs_of(ctx);
    hipStream_t s = ctx->stream;
    // Concatenated results number pairs across units (JSON Lines).
    if (launch_unit_bases(s, b, sh.n_units, sh.concat) != hipSuccess) return CJX_ERR_DEVICE;
    // Well-nested units are matched directly; the rest take the sort path.
    // Once the walks have told them apart, the chunk matcher and the sort
    // path touch different units and can run side by side.  That pays for
    // batches of independent documents, whose faulty and deep members give
    // the sort path real work; a single unit or JSON Lines chunks, whose sort
    // path finds nothing to do, keep one stream (the fork and join cost more
    // than the overlap: 10-20 us of launch gaps measured, and about 30 us per
    // 2 GB document inside a captured graph).
    if (launch_match_front(s, b, sh.n_units, (uint32_t)sh.chunks, d_pair) != hipSuccess)
        return CJX_ERR_DEVICE;
    const bool fork = sh.n_units > 1 && !sh.concat && ensure_aux(ctx);
    if (fork && !fork_aux(ctx, 1)) return CJX_ERR_DEVICE;
    hipStream_t ss = fork ? ctx->aux : s;
    if (sh.n_units == 1) {  // one unit: the multi-block sort path
        const uint32_t G = phase3_blocks(sh.max_rows);
        const uint32_t PG = pair_blocks(sh.max_rows);
        if (launch_phase3(ss, b, 0, G, PG, u16, d_pair) != hipSuccess) return CJX_ERR_DEVICE;
    } else if (launch_sort_units(ss, b, sh.n_units, u16, d_pair) != hipSuccess) {
        return CJX_ERR_DEVICE;
    }
    if (launch_match_local(s, b, (uint32_t)sh.chunks, d_pair) != hipSuccess) return CJX_ERR_DEVICE;
    if (fork && !join_aux(ctx, 1)) return CJX_ERR_DEVICE;
    if (ctx->timing) {
        hipEvent_t *tev = ctx->tev[ctx->timing_count % cjx_context::kTimingRing];
        (void)hipEventRecord(tev[2], s);
        ctx->timing_count++;
    }
    return CJX_OK;
}

// Enqueues the whole parse of `units` (already in ctx->units on the device)
// on the context stream.
int enqueue_parse(cjx_context *ctx, const uint8_t *d_in, const Shape &sh, unsigned flags,
                  uint32_t *d_struct, int32_t *d_pair) {
    const bool u16 = (flags & CJX_FLAG_DEPTH_U16) != 0;
    if (u16 && !ctx->tmp.ensure(sh.rows * 4 + 16)) return CJX_ERR_ALLOC;
    LaunchBufs b = bufs_of(ctx);
    hipStream_t s = ctx->stream;
    if ((ctx->one_unit_dev ? launch_reset_one(s, b, ctx->one_unit, sh.ntiles, sh.n_small)
                           : launch_reset(s, b, sh.ntiles, sh.n_units)) != hipSuccess)
        return CJX_ERR_DEVICE;
    hipEvent_t *tev = ctx->tev[ctx->timing_count % cjx_context::kTimingRing];
    if (ctx->timing) (void)hipEventRecord(tev[0], s);
    const bool strict = (flags & CJX_FLAG_UTF8_STRICT) != 0;
    const bool jsonl = (flags & CJX_FLAG_JSONL) != 0;
    // Beside the tile tokenizer on the second stream: the strict check, which
    // shares no output with it (it only sets the units' UTF-8 flags, which the
    // tokenizer leaves alone in strict mode; the tokenizer is bound by its
    // instruction issue, the check by HBM), and in document batches the small
    // units' waves (their units are not the tiles' units; measured 4-8 us
    // faster per 50,000-document batch than after the tiles).
    const bool two = sh.ntiles && sh.n_small && sh.n_units > 1 && !sh.concat;
    const bool fork = (strict || two) && ensure_aux(ctx);
    if (fork && !fork_aux(ctx, 0)) return CJX_ERR_DEVICE;
    if (strict && launch_utf8_strict(fork ? ctx->aux : s, b, d_in, sh.ntiles, sh.n_small) !=
                      hipSuccess)
        return CJX_ERR_DEVICE;
    if (fork && two) {
        if (launch_tokenize(ctx->aux, b, d_in, 0, sh.n_small, jsonl, !strict, d_struct,
                            ctx->tok_shape, ctx->n_tiny) != hipSuccess ||
            launch_tokenize(s, b, d_in, sh.ntiles, 0, jsonl, !strict, d_struct,
                            ctx->tok_shape) != hipSuccess)
            return CJX_ERR_DEVICE;
    } else if (launch_tokenize(s, b, d_in, sh.ntiles, sh.n_small, jsonl, !strict, d_struct,
                               ctx->tok_shape, ctx->n_tiny) != hipSuccess) {
        return CJX_ERR_DEVICE;
    }
    if (fork && !join_aux(ctx, 0)) return CJX_ERR_DEVICE;
    if (ctx->timing) (void)hipEventRecord(tev[1], s);
    return enqueue_back(ctx, sh, flags, d_pair);
}

// The second stream for host-to-device copies and its events (created on
// first use).
int ensure_copy_stream(cjx_context *ctx) {
    if (ctx->copy_stream) return CJX_OK;
    if (hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->idle_ev, hipEventDisableTiming) != hipSuccess)
        return CJX_ERR_DEVICE;
    for (hipEvent_t &e : ctx->copy_ev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return CJX_ERR_DEVICE;
    for (hipEvent_t &e : ctx->load_ev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return CJX_ERR_DEVICE;
    return CJX_OK;
}

// The output stream of the streamed copy out, its part events and the
// pinned part counts (created on first use: a context that never streams
// rows holds no third stream, and so takes no third hardware queue).
int ensure_out_stream(cjx_context *ctx) {
    if (ctx->out_stream) return CJX_OK;
    if (!ctx->h_part.ensure(cjx_context::kCopyEvents * sizeof(uint64_t))) return CJX_ERR_ALLOC;
    for (hipEvent_t &e : ctx->part_ev)
        if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            return CJX_ERR_DEVICE;
    if (hipStreamCreateWithFlags(&ctx->out_stream, hipStreamNonBlocking) != hipSuccess) {
        ctx->out_stream = nullptr;
        return CJX_ERR_DEVICE;
    }
    return CJX_OK;
}

// Structural rows streamed to a host result buffer during a chunked parse.
struct RowStream {
    int32_t *buf = nullptr;  // pinned, host layout: row i of the unit at buf[1 + i]
    int32_t *dev = nullptr;  // buf's address on the device (k_rows_out writes through it;
                             // the host reads after the output stream is synchronized)
    uint64_t cap_rows = 0;   // rows the buffer's layout holds (2 (cap_rows + 2) cells)
    uint64_t copied = 0;     // rows [0, copied) are queued on the output stream
    bool ok = true;          // false: the result is copied after the parse instead
};

// A single large unit whose bytes are still in host memory: the input is
// copied in parts of `part` bytes (whole tiles) on the copy stream, and the
// tokenizer runs as one launch per part, each waiting for its part's copy,
// so phases 1-2 overlap the transfer.  Tile tickets continue across the
// launches, which run in order on the context stream, so the tiles and the
// look-back are exactly those of one launch.  With `counts`, each part's
// row count (the inclusive count in its last tile's status) is copied to
// ctx->h_part after the part's launch and ctx->part_ev[k] marks it.
int enqueue_parse_chunked(cjx_context *ctx, const uint8_t *host, uint64_t n, const Shape &sh,
                          unsigned flags, uint32_t *d_struct, int32_t *d_pair, uint64_t part,
                          bool counts) {
    const bool u16 = (flags & CJX_FLAG_DEPTH_U16) != 0;
    if (u16 && !ctx->tmp.ensure(sh.rows * 4 + 16)) return CJX_ERR_ALLOC;
    if (int rc = ensure_copy_stream(ctx)) return rc;
    const uint64_t parts = (n + part - 1) / part;
    if (parts > (uint64_t)cjx_context::kCopyEvents || part % kTileBytes) return CJX_ERR_INPUT;
    LaunchBufs b = bufs_of(ctx);
    hipStream_t s = ctx->stream;
    uint8_t *d_in = ctx->in.as<uint8_t>();
    // the copies start once earlier work on the context stream (which may
    // still read the input buffer) is done
    if (hipEventRecord(ctx->idle_ev, s) != hipSuccess ||
        hipStreamWaitEvent(ctx->copy_stream, ctx->idle_ev, 0) != hipSuccess)
        return CJX_ERR_DEVICE;
    for (uint64_t k = 0; k < parts; k++) {
        const uint64_t off = k * part, len = std::min<uint64_t>(part, n - off);
        if (hipMemcpyAsync(d_in + off, host + off, len, hipMemcpyHostToDevice,
                           ctx->copy_stream) != hipSuccess ||
            hipEventRecord(ctx->copy_ev[k], ctx->copy_stream) != hipSuccess)
            return CJX_ERR_DEVICE;
    }
    if (launch_reset(s, b, sh.ntiles, sh.n_units) != hipSuccess) return CJX_ERR_DEVICE;
    hipEvent_t *tev = ctx->tev[ctx->timing_count % cjx_context::kTimingRing];
    const bool strict = (flags & CJX_FLAG_UTF8_STRICT) != 0;
    const bool jsonl = (flags & CJX_FLAG_JSONL) != 0;
    uint32_t launched = 0;
    for (uint64_t k = 0; k < parts; k++) {
        if (hipStreamWaitEvent(s, ctx->copy_ev[k], 0) != hipSuccess) return CJX_ERR_DEVICE;
        if (k == 0 && ctx->timing) (void)hipEventRecord(tev[0], s);
        const uint32_t upto =
            k + 1 == parts ? sh.ntiles : (uint32_t)std::min<uint64_t>(sh.ntiles, (k + 1) * part / kTileBytes);
        if (launch_tokenize_tiles(s, b, d_in, upto - launched, jsonl, !strict, d_struct,
                                  ctx->tok_shape) != hipSuccess)
            return CJX_ERR_DEVICE;
        if (counts && (hipMemcpyAsync(static_cast<uint64_t *>(ctx->h_part.p) + k,
                                      b.tile_st + 2 * (uint64_t)(upto - 1), 8,
                                      hipMemcpyDeviceToHost, s) != hipSuccess ||
                       hipEventRecord(ctx->part_ev[k], s) != hipSuccess))
            return CJX_ERR_DEVICE;
        launched = upto;
    }
    if (strict && launch_utf8_strict(s, b, d_in, sh.ntiles, sh.n_small) != hipSuccess)
        return CJX_ERR_DEVICE;
    if (ctx->timing) (void)hipEventRecord(tev[1], s);
    return enqueue_back(ctx, sh, flags, d_pair);
}

// enqueue_parse, or (CJX_FLAG_GRAPH) a replay of the same sequence captured
// as a HIP graph.  The graph is re-captured whenever the launch shape, the
// flags or any device pointer baked into its kernel arguments changes; the
// inputs it reads (unit table, input bytes) are written before the launch.
int enqueue_maybe_graph(cjx_context *ctx, const uint8_t *d_in, const Shape &sh, unsigned flags,
                        uint32_t *d_struct, int32_t *d_pair) {
    if (!(flags & CJX_FLAG_GRAPH)) return enqueue_parse(ctx, d_in, sh, flags, d_struct, d_pair);
    if ((flags & CJX_FLAG_DEPTH_U16) && !ctx->tmp.ensure(sh.rows * 4 + 16)) return CJX_ERR_ALLOC;
    cjx_context::GraphKey key;
    std::memset(&key, 0, sizeof(key));  // padding too: keys are compared bytewise
    key.b = bufs_of(ctx);
    key.d_in = d_in;
    key.d_struct = d_struct;
    key.d_pair = d_pair;
    key.ntiles = sh.ntiles;
    key.n_units = sh.n_units;
    key.n_small = sh.n_small;
    key.rows = sh.rows;
    key.max_rows = sh.max_rows;
    key.chunks = sh.chunks;
    key.flags = flags;
    key.concat = sh.concat;
    key.one_unit_dev = ctx->one_unit_dev;
    if (ctx->one_unit_dev) key.one_unit = ctx->one_unit;
    key.tok_shape = ctx->tok_shape;
    key.n_tiny = ctx->n_tiny;
    if (!ctx->graph || std::memcmp(&key, &ctx->graph_key, sizeof(key)) != 0) {
        if (ctx->graph) (void)hipGraphExecDestroy(ctx->graph);
        ctx->graph = nullptr;
        const bool timing = ctx->timing;
        ctx->timing = false;  // no event records inside the graph
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            ctx->timing = timing;
            return CJX_ERR_DEVICE;
        }
        const int rc = enqueue_parse(ctx, d_in, sh, flags, d_struct, d_pair);
        const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
        ctx->timing = timing;
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess || !g) return CJX_ERR_DEVICE;
        const hipError_t ei = hipGraphInstantiate(&ctx->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            ctx->graph = nullptr;
            return CJX_ERR_DEVICE;
        }
        std::memcpy(&ctx->graph_key, &key, sizeof(key));
    }
    hipEvent_t *tev = ctx->tev[ctx->timing_count % cjx_context::kTimingRing];
    if (ctx->timing) (void)hipEventRecord(tev[0], ctx->stream);
    if (hipGraphLaunch(ctx->graph, ctx->stream) != hipSuccess) return CJX_ERR_DEVICE;
    if (ctx->timing) {  // the whole graph is reported as tokenize, phase 3 as 0
        (void)hipEventRecord(tev[1], ctx->stream);
        (void)hipEventRecord(tev[2], ctx->stream);
        ctx->timing_count++;
    }
    return CJX_OK;
}

int read_status(cjx_context *ctx, uint32_t n_units) {
    hipStream_t s = ctx->stream;
    if (hipMemcpyAsync(ctx->h_uout.p, ctx->uout.p, n_units * sizeof(UnitOut),
                       hipMemcpyDeviceToHost, s) != hipSuccess)
        return CJX_ERR_DEVICE;
    if (hipMemcpyAsync(ctx->h_ls.p, ctx->ls.p, sizeof(LaunchState), hipMemcpyDeviceToHost, s) !=
        hipSuccess)
        return CJX_ERR_DEVICE;
    if (hipStreamSynchronize(s) != hipSuccess) return CJX_ERR_DEVICE;
    return CJX_OK;
}

int unit_code(const UnitOut &o) {
    if (o.utf8_err) return CJX_ERR_UTF8;
    if (o.struct_err) return CJX_ERR_STRUCTURE;
    return CJX_OK;
}

void zero_result(cjx_result *out) {
    std::memset(out, 0, sizeof(*out));
    out->error_chunk = -1;
}

// Parses units whose bytes are on the host: copies them in, parses, and
// assembles the reference's host layout.  Retries once with full capacity
// when the first capacity guess is exceeded.
// Copies units whose bytes are on the host into the device input (through a
// pinned staging buffer when there are many small ones) and parses them,
// retrying once with full output regions if the first guess overflows.
// Leaves the rows in ctx->structural / ctx->pair and the status in h_uout/h_ls.
// Part size for copying a single large unit in parts (0: copy it whole):
// units of at least 64 MiB go in at most 64 parts of at least 16 MiB, whole
// tiles each.
uint64_t chunked_part(const std::vector<UnitDesc> &units) {
    constexpr uint64_t kMinChunked = 64ull << 20, kMinPart = 16ull << 20;
    if (units.size() != 1 || units[0].n < kMinChunked) return 0;
    uint64_t part = std::max<uint64_t>(kMinPart, (units[0].n + 63) / 64);
    part = (part + kTileBytes - 1) / kTileBytes * kTileBytes;
    return part;
}

// While the parts of a chunked parse come in, copies each tokenized part's
// structural rows to the host result buffer (k_rows_out on the output
// stream), so the copy out overlaps the copy in.  The buffer is sized after the first part
// from its row density (25 % and 64 Ki rows of margin, at most the unit's
// row region); rows beyond it stop the streaming and the result is copied
// after the parse instead.  Tiles publish inclusive counts of the rows
// before their end, so a part's count is the row where the next part starts.
void stream_rows(cjx_context *ctx, RowStream &rs, uint64_t n, uint64_t part, uint32_t out_cap,
                 const uint32_t *d_struct) {
    const uint64_t parts = (n + part - 1) / part;
    const uint64_t *cnt = static_cast<const uint64_t *>(ctx->h_part.p);
    for (uint64_t k = 0; k < parts && rs.ok; k++) {
        if (hipEventSynchronize(ctx->part_ev[k]) != hipSuccess) {
            rs.ok = false;
            break;
        }
        const uint64_t c = std::min<uint64_t>(cnt[k] & kStCountMask, out_cap);
        if (k == 0) {
            const double rho = (double)c / (double)std::min(part, n);
            const uint64_t est = std::min<uint64_t>(
                (uint64_t)(rho * 1.25 * (double)n) + (64u << 10), out_cap);
            rs.buf = static_cast<int32_t *>(pinned_get(2 * (est + 2) * sizeof(int32_t)));
            rs.cap_rows = est;
            void *dv = nullptr;
            if (!rs.buf || hipHostGetDevicePointer(&dv, rs.buf, 0) != hipSuccess)
                rs.ok = false;
            rs.dev = static_cast<int32_t *>(dv);
        }
        if (!rs.ok || c > rs.cap_rows) {
            rs.ok = false;
            break;
        }
        // a kernel writes the rows, so the copy out does not queue behind the
        // input's DMA copies
        if (c > rs.copied &&
            (hipStreamWaitEvent(ctx->out_stream, ctx->part_ev[k], 0) != hipSuccess ||
             launch_rows_out(ctx->out_stream, d_struct + rs.copied,
                             reinterpret_cast<uint32_t *>(rs.dev + 1 + rs.copied),
                             c - rs.copied) != hipSuccess)) {
            rs.ok = false;
            break;
        }
        rs.copied = std::max(rs.copied, c);
    }
}

// Whether the units' bytes lie densely in the source: the span from the
// first unit's start to the last unit's end holds at most an eighth more
// than the units themselves, plus 1 MiB.  Then one copy of the span and a
// placement on the device replace the host's gathering of the units into a
// staging buffer.  The copy reads exactly [lo, lo + bytes) of the caller's
// memory and lands at device offset skew = lo's address mod 16, so that
// device and host addresses agree mod 16.
bool dense_span(const uint8_t *data, const std::vector<uint64_t> &offs,
                const std::vector<uint64_t> &sizes, uint64_t total, const uint8_t *&lo,
                uint64_t &bytes, uint64_t &skew) {
    uint64_t a = UINT64_MAX, b = 0;
    for (size_t i = 0; i < offs.size(); i++) {
        a = std::min(a, offs[i]);
        b = std::max(b, offs[i] + sizes[i]);
    }
    lo = data + a;
    bytes = b - a;
    skew = reinterpret_cast<uintptr_t>(lo) & 15u;
    return bytes <= total + total / 8 + (1u << 20);
}

int run_host_units(cjx_context *ctx, const uint8_t *data, const std::vector<uint64_t> &offs,
                   const std::vector<uint64_t> &sizes, const std::vector<uint64_t> &pos_base,
                   unsigned flags, bool concat, std::vector<UnitDesc> &units, float *ms,
                   RowStream *rs = nullptr) {
    int rc = set_device(ctx);
    if (rc) return rc;
    uint64_t dev_bytes = 0;
    uint32_t ntiles = 0, n_small = 0;
    if (!plan_units(sizes, pos_base, units, dev_bytes, ntiles, n_small)) return CJX_ERR_INPUT;
    uint64_t total = 0;
    for (uint64_t z : sizes) total += z;
    if (total == 0 || total > kMaxLaunchBytes) return CJX_ERR_INPUT;
    if (!ctx->in.ensure(dev_bytes)) return CJX_ERR_ALLOC;
    hipStream_t s = ctx->stream;
    // One large unit: copy in parts overlapped with the tokenizer (first
    // attempt only; a re-run with exact regions finds the input in place).
    const uint64_t part = chunked_part(units);
    const uint8_t *span_lo = nullptr;
    uint64_t span_bytes = 0, span_skew = 0;
    std::vector<uint64_t> unit_src;  // units' offsets in the copied span
    if (part) {
    } else if (units.size() > 16 &&
               dense_span(data, offs, sizes, total, span_lo, span_bytes, span_skew)) {
        // One copy of the source span, then k_place_units moves every unit
        // to its aligned offset (the JSON Lines chunks of a file lie back to
        // back).
        if (!ctx->span.ensure(span_skew + span_bytes + 32)) return CJX_ERR_ALLOC;
        if (hipMemcpyAsync(ctx->span.as<uint8_t>() + span_skew, span_lo, span_bytes,
                           hipMemcpyHostToDevice, s) != hipSuccess)
            return CJX_ERR_DEVICE;
        span_bytes += span_skew;  // the device span: [0, skew) unused, then the bytes
        unit_src.resize(units.size());
        for (size_t i = 0; i < units.size(); i++)
            unit_src[i] = span_skew + (uint64_t)((data + offs[i]) - span_lo);
    } else if (units.size() > 16) {  // one pinned staging copy and one H2D instead of one per unit
        if (!ctx->h_stage.ensure(dev_bytes)) return CJX_ERR_ALLOC;
        uint8_t *st = static_cast<uint8_t *>(ctx->h_stage.p);
        if (hipStreamSynchronize(s) != hipSuccess) return CJX_ERR_DEVICE;  // staging reuse
        for (size_t i = 0; i < units.size(); i++)
            std::memcpy(st + units[i].in_off, data + offs[i], sizes[i]);
        if (hipMemcpyAsync(ctx->in.p, st, dev_bytes, hipMemcpyHostToDevice, s) != hipSuccess)
            return CJX_ERR_DEVICE;
    } else {
        for (size_t i = 0; i < units.size(); i++)
            if (hipMemcpyAsync(ctx->in.as<uint8_t>() + units[i].in_off, data + offs[i], sizes[i],
                               hipMemcpyHostToDevice, s) != hipSuccess)
                return CJX_ERR_DEVICE;
    }

    Shape sh{ntiles, (uint32_t)units.size(), n_small, 0, 0, 0, 0, concat};
    const bool records = (flags & CJX_FLAG_JSONL) && (flags & CJX_FLAG_RECORDS);
    for (int attempt = 0; attempt < 2; attempt++) {
        set_regions(units, attempt == 1);
        shape_rows(sh, units);
        if ((rc = ensure_work(ctx, sh))) return rc;
        if (!ctx->structural.ensure(sh.rows * 4 + 16) || !ctx->pair.ensure(sh.rows * 4 + 16) ||
            (records && !ctx->row_nl.ensure(sh.rows + 16)))
            return CJX_ERR_ALLOC;
        ctx->want_nl = records;  // the tokenizer also writes the newline flags
        const bool place = !unit_src.empty() && attempt == 0;
        if ((rc = upload_units(ctx, units, ntiles, n_small, place ? unit_src.data() : nullptr)))
            return rc;
        (void)hipEventRecord(ctx->ev0, s);
        if (place && launch_place_units(s, ctx->span.as<uint8_t>(), span_bytes, ctx->unit_src,
                                        ctx->units.as<UnitDesc>(), (uint32_t)units.size(),
                                        ctx->in.as<uint8_t>(), dev_bytes) != hipSuccess)
            return CJX_ERR_DEVICE;
        if (part && attempt == 0) {  // device_ms then includes waiting for the copies
            if (rs && ensure_out_stream(ctx) != CJX_OK) rs->ok = false;
            const bool counts = rs && rs->ok;
            rc = enqueue_parse_chunked(ctx, data + offs[0], sizes[0], sh, flags,
                                       ctx->structural.as<uint32_t>(), ctx->pair.as<int32_t>(),
                                       part, counts);
            if (!rc && counts)
                stream_rows(ctx, *rs, sizes[0], part, units[0].out_cap,
                            ctx->structural.as<uint32_t>());
        } else {
            if (rs) rs->ok = false;  // a re-run rewrites the rows
            rc = enqueue_parse(ctx, ctx->in.as<uint8_t>(), sh, flags,
                               ctx->structural.as<uint32_t>(), ctx->pair.as<int32_t>());
        }
        ctx->want_nl = false;
        if (rc) return rc;
        (void)hipEventRecord(ctx->ev1, s);
        if ((rc = read_status(ctx, sh.n_units))) return rc;
        const LaunchState *ls = static_cast<const LaunchState *>(ctx->h_ls.p);
        if (!ls->overflow) break;
        if (attempt == 1) return CJX_ERR_CAPACITY;
    }
    *ms = 0.f;
    (void)hipEventElapsedTime(ms, ctx->ev0, ctx->ev1);
    return CJX_OK;
}

// The per-record table of the JSON Lines units just parsed (their rows are
// still in ctx->structural): count and number on the device, read the count,
// fill the table on the device, copy it out.
int collect_records(cjx_context *ctx, uint32_t n_units, cjx_result *out) {
    hipStream_t s = ctx->stream;
    ctx->want_nl = true;  // the parse just run wrote the newline flags
    LaunchBufs b = bufs_of(ctx);
    ctx->want_nl = false;
    if (!ctx->rec_count.ensure((size_t)n_units * 4 + 16)) return CJX_ERR_ALLOC;
    if (launch_records(s, b, n_units, ctx->structural.as<uint32_t>(),
                       ctx->rec_count.as<uint32_t>(), nullptr, 0, false) != hipSuccess)
        return CJX_ERR_DEVICE;
    int rc = read_status(ctx, n_units);
    if (rc) return rc;
    const uint64_t R = static_cast<const LaunchState *>(ctx->h_ls.p)->n_records;
    out->n_records = (int64_t)R;
    if (R == 0) return CJX_OK;
    if (!ctx->recs.ensure(R * sizeof(RecordOut) + 16)) return CJX_ERR_ALLOC;
    cjx_record *h = nullptr;
    if (!(h = static_cast<cjx_record *>(pinned_get(R * sizeof(cjx_record)))))
        return CJX_ERR_ALLOC;
    if (launch_records(s, b, n_units, ctx->structural.as<uint32_t>(),
                       ctx->rec_count.as<uint32_t>(), ctx->recs.as<RecordOut>(), R, true) !=
            hipSuccess ||
        hipMemcpyAsync(h, ctx->recs.p, R * sizeof(cjx_record), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        pinned_put(h);
        return CJX_ERR_DEVICE;
    }
    out->records = h;
    return CJX_OK;
}

int parse_host_units(cjx_context *ctx, const uint8_t *data, const std::vector<uint64_t> &offs,
                     const std::vector<uint64_t> &sizes, const std::vector<uint64_t> &pos_base,
                     unsigned flags, cjx_result *out) {
    std::vector<UnitDesc> units;
    float ms = 0.f;
    RowStream rs;
    int rc = run_host_units(ctx, data, offs, sizes, pos_base, flags, sizes.size() > 1, units, &ms,
                            &rs);
    auto drop_stream = [&] {  // rows still in flight to rs.buf land before it is reused
        if (!rs.buf) return;
        (void)hipStreamSynchronize(ctx->out_stream);
        pinned_put(rs.buf);
        rs.buf = nullptr;
    };
    if (rc) {
        drop_stream();
        return rc;
    }
    hipStream_t s = ctx->stream;
    const UnitOut *uo = static_cast<const UnitOut *>(ctx->h_uout.p);
    const LaunchState *ls = static_cast<const LaunchState *>(ctx->h_ls.p);
    const uint64_t N = ls->n_total;
    out->n_structural = (int64_t)N;
    out->n_brackets = (int64_t)ls->m_total;
    out->chunk_count = (int32_t)units.size();
    out->device_ms = ms;
    out->error = CJX_OK;
    out->error_chunk = -1;
    for (size_t i = 0; i < units.size(); i++) {
        int c = unit_code(uo[i]);
        if (c != CJX_OK) {
            out->error = c;
            out->error_chunk = (int32_t)i;
            break;
        }
    }
    // Host layout, parse_standard_json.cu:1659-1680 / parse_json_lines.cu:1070-1089.
    const size_t cells = 2 * (N + 2);
    // every structural row already queued for rs.buf, which holds the layout
    const bool streamed = rs.buf && rs.ok && units.size() == 1 && rs.copied == N &&
                          N <= rs.cap_rows;
    int32_t *buf = nullptr;
    if (streamed) {
        buf = rs.buf;
        rs.buf = nullptr;
    } else {
        drop_stream();
        if (!(buf = static_cast<int32_t *>(pinned_get(cells * sizeof(int32_t)))))
            return CJX_ERR_ALLOC;
    }
    bool ok = true;
    if (streamed) {  // the pair values remain
        if (N)
            ok = hipMemcpyAsync(buf + N + 2, ctx->pair.p, N * 4, hipMemcpyDeviceToHost, s) ==
                 hipSuccess;
        if (hipStreamSynchronize(ctx->out_stream) != hipSuccess) ok = false;
    } else if (units.size() == 1) {  // one region starting at row 0: copy it as is
        if (N)
            ok = hipMemcpyAsync(buf + 1, ctx->structural.p, N * 4, hipMemcpyDeviceToHost, s) ==
                     hipSuccess &&
                 hipMemcpyAsync(buf + N + 2, ctx->pair.p, N * 4, hipMemcpyDeviceToHost, s) ==
                     hipSuccess;
    } else {  // pack the regions on the device, then one copy
        ok = ctx->layout.ensure(cells * 4) &&
             launch_assemble(s, bufs_of(ctx), (uint32_t)units.size(),
                             ctx->structural.as<uint32_t>(), ctx->pair.as<int32_t>(),
                             ctx->layout.as<int32_t>(), nullptr, (uint32_t)N, true) == hipSuccess &&
             hipMemcpyAsync(buf, ctx->layout.p, cells * 4, hipMemcpyDeviceToHost, s) == hipSuccess;
    }
    if (!ok || hipStreamSynchronize(s) != hipSuccess) {
        pinned_put(buf);
        return ok ? CJX_ERR_DEVICE : CJX_ERR_ALLOC;
    }
    buf[0] = 0;
    buf[N + 1] = (int32_t)(N + 1);
    buf[2 * N + 2] = -1;
    buf[2 * N + 3] = -1;
    out->buf = buf;
    if ((flags & CJX_FLAG_RECORDS) && (flags & CJX_FLAG_JSONL)) {
        const int rr = collect_records(ctx, (uint32_t)units.size(), out);
        if (rr) {
            out->buf = nullptr;
            pinned_put(buf);
            return rr;
        }
    }
    out->structural = buf;
    out->pair_pos = buf + N + 1;
    out->total_result_size = (int32_t)(N + 2);
    out->file_size = (int32_t)(N + 2);
    out->result_sizes = (int32_t *)std::malloc(units.size() * sizeof(int32_t));
    out->result_sizes_prefix = (int32_t *)std::malloc(units.size() * sizeof(int32_t));
    uint32_t prefix = 0;
    for (size_t i = 0; i < units.size(); i++) {
        const uint32_t nu = uo[i].n_end - uo[i].n_base;
        prefix += nu;
        out->result_sizes[i] = (int32_t)nu;
        out->result_sizes_prefix[i] = (int32_t)prefix;
    }
    return out->error;
}

// Batch of independent documents: positions and pair values are local to
// each document; rows are packed densely on the device and copied once.
int parse_host_batch(cjx_context *ctx, const uint8_t *data, const std::vector<uint64_t> &offs,
                     const std::vector<uint64_t> &sizes, unsigned flags, cjx_batch_result *out) {
    std::vector<UnitDesc> units;
    float ms = 0.f;
    std::vector<uint64_t> zero(sizes.size(), 0);
    int rc = run_host_units(ctx, data, offs, sizes, zero, flags, false, units, &ms);
    if (rc) return rc;
    hipStream_t s = ctx->stream;
    const UnitOut *uo = static_cast<const UnitOut *>(ctx->h_uout.p);
    const LaunchState *ls = static_cast<const LaunchState *>(ctx->h_ls.p);
    const uint64_t N = ls->n_total;
    int32_t *rows = nullptr;
    if (!(rows = static_cast<int32_t *>(pinned_get((2 * N + 2) * sizeof(int32_t)))))
        return CJX_ERR_ALLOC;
    bool ok = true;
    if (N) {
        ok = ctx->layout.ensure(2 * N * 4 + 16) &&
             launch_assemble(s, bufs_of(ctx), (uint32_t)units.size(),
                             ctx->structural.as<uint32_t>(), ctx->pair.as<int32_t>(),
                             ctx->layout.as<int32_t>(), ctx->layout.as<int32_t>() + N,
                             (uint32_t)N, false) == hipSuccess &&
             hipMemcpyAsync(rows, ctx->layout.p, 2 * N * 4, hipMemcpyDeviceToHost, s) ==
                 hipSuccess;
    }
    if (!ok || hipStreamSynchronize(s) != hipSuccess) {
        pinned_put(rows);
        return ok ? CJX_ERR_DEVICE : CJX_ERR_ALLOC;
    }
    out->structural = rows;
    out->pair_pos = rows + N;
    out->n_structural = (int64_t)N;
    out->n_docs = (int64_t)units.size();
    out->device_ms = ms;
    out->docs = static_cast<cjx_doc_result *>(std::calloc(units.size(), sizeof(cjx_doc_result)));
    int first_err = CJX_OK;
    for (size_t i = 0; i < units.size(); i++) {
        cjx_doc_result &d = out->docs[i];
        d.byte_base = (int64_t)offs[i];
        d.struct_begin = uo[i].row_base;
        d.struct_end = (int64_t)uo[i].row_base + (uo[i].n_end - uo[i].n_base);
        d.error = unit_code(uo[i]);
        if (d.error && !first_err) first_err = d.error;
    }
    out->error = first_err;
    return CJX_OK;
}

int parse_units_device(cjx_context *ctx, const uint8_t *d_input, const size_t *doc_off,
                       const size_t *doc_size, size_t n_docs, unsigned flags,
                       int32_t *d_structural, int32_t *d_pair, size_t cap,
                       cjx_record *d_records, size_t max_records);

}  // namespace cjx_host

using namespace cjx_host;

extern "C" {

const char *cjx_version(void) { return "cujson_amd 0.1 (gfx950)"; }

const char *cjx_strerror(int code) {
    switch (code) {
        case CJX_OK: return "ok";
        case CJX_ERR_INPUT: return "invalid input";
        case CJX_ERR_UTF8: return "invalid UTF-8";
        case CJX_ERR_STRUCTURE: return "invalid JSON structure";
        case CJX_ERR_ALLOC: return "allocation failed";
        case CJX_ERR_DEVICE: return "device error";
        case CJX_ERR_CAPACITY: return "output capacity exceeded";
        case CJX_ERR_TOO_LARGE: return "input larger than the entry point's maximum";
        default: return "unknown error";
    }
}

int cjx_context_create(int device, cjx_context **out) {
    if (!out) return CJX_ERR_INPUT;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return CJX_ERR_DEVICE;
    if (device < 0 || device >= ndev) return CJX_ERR_INPUT;
    cjx_context *ctx = new cjx_context();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
        delete ctx;
        return CJX_ERR_DEVICE;
    }
    *out = ctx;
    return CJX_OK;
}

void cjx_context_destroy(cjx_context *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    DevBuf *bufs[] = {&ctx->in,      &ctx->structural, &ctx->pair,   &ctx->units,
                      &ctx->uout,    &ctx->ls,         &ctx->tile_st, &ctx->layout,
                      &ctx->oc_char, &ctx->oc_idx,     &ctx->keys,   &ctx->sorted,
                      &ctx->tmp,     &ctx->spos,       &ctx->hist,
                      &ctx->ck,      &ctx->gmin,       &ctx->gmin1,  &ctx->gmin2,
                      &ctx->gagg,    &ctx->gd0,        &ctx->lows,   &ctx->highs,
                      &ctx->chunk_unit, &ctx->big_units, &ctx->slow_units, &ctx->unit_part,
                      &ctx->rec_count,
                      &ctx->recs,       &ctx->row_nl,   &ctx->nl_seg, &ctx->nl_off,
                      &ctx->nl_idx,     &ctx->val_exact, &ctx->span};
    for (DevBuf *b : bufs) b->release();
    ctx->h_units.release();
    ctx->h_uout.release();
    ctx->h_ls.release();
    ctx->h_stage.release();
    ctx->h_nl.release();
    for (int i = 0; i < cjx_context::kUploadSlots; i++) {
        if (ctx->up_ev[i]) {
            (void)hipEventSynchronize(ctx->up_ev[i]);
            (void)hipEventDestroy(ctx->up_ev[i]);
        }
        ctx->up[i].release();
    }
    for (auto &tr : ctx->tev)
        for (hipEvent_t e : tr)
            if (e) (void)hipEventDestroy(e);
    if (ctx->copy_stream) {
        (void)hipStreamSynchronize(ctx->copy_stream);
        (void)hipStreamDestroy(ctx->copy_stream);
    }
    if (ctx->out_stream) {
        (void)hipStreamSynchronize(ctx->out_stream);
        (void)hipStreamDestroy(ctx->out_stream);
    }
    if (ctx->aux) {
        (void)hipStreamSynchronize(ctx->aux);
        (void)hipStreamDestroy(ctx->aux);
    }
    for (hipEvent_t e : ctx->fork_ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->join_ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->copy_ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->part_ev)
        if (e) (void)hipEventDestroy(e);
    ctx->h_part.release();
    for (int i = 0; i < cjx_context::kLoadBufs; i++) {
        if (ctx->load_ev[i]) (void)hipEventDestroy(ctx->load_ev[i]);
        ctx->load_buf[i].release();
    }
    if (ctx->idle_ev) (void)hipEventDestroy(ctx->idle_ev);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->graph) (void)hipGraphExecDestroy(ctx->graph);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

void *cjx_context_stream(cjx_context *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int cjx_context_set_timing(cjx_context *ctx, int enable) {
    if (!ctx) return CJX_ERR_INPUT;
    if (enable && !ctx->tev[0][0]) {
        if (set_device(ctx)) return CJX_ERR_DEVICE;
        for (auto &tr : ctx->tev)
            for (hipEvent_t &e : tr)
                if (hipEventCreate(&e) != hipSuccess) return CJX_ERR_DEVICE;
    }
    ctx->timing = enable != 0;
    ctx->timing_count = 0;
    return CJX_OK;
}

int cjx_context_set_tokenizer_shape(cjx_context *ctx, int tile_waves, int small_waves) {
    if (!ctx || tile_waves < 0 || small_waves < 0 ||
        !cjx::tok_shape_valid((uint32_t)tile_waves, (uint32_t)small_waves))
        return CJX_ERR_INPUT;
    ctx->tok_shape.tile_waves = (uint32_t)tile_waves;
    ctx->tok_shape.small_waves = (uint32_t)small_waves;
    return CJX_OK;
}

int cjx_context_tokenizer_shape(cjx_context *ctx, int *tile_waves, int *small_waves) {
    if (!ctx || !tile_waves || !small_waves) return CJX_ERR_INPUT;
    *tile_waves = (int)ctx->tok_shape.tile_waves;
    *small_waves = (int)ctx->tok_shape.small_waves;
    return CJX_OK;
}

int cjx_context_timings(cjx_context *ctx, cjx_timings *out, int max) {
    if (!ctx || !out || max <= 0 || !ctx->timing) return -1;
    const uint64_t have = std::min<uint64_t>(ctx->timing_count, cjx_context::kTimingRing);
    const uint64_t k = std::min<uint64_t>(have, (uint64_t)max);
    for (uint64_t i = 0; i < k; i++) {  // oldest first, ending with the latest parse
        hipEvent_t *tr = ctx->tev[(ctx->timing_count - k + i) % cjx_context::kTimingRing];
        if (hipEventSynchronize(tr[2]) != hipSuccess) return -1;
        cjx_timings &t = out[i];
        (void)hipEventElapsedTime(&t.tokenize_ms, tr[0], tr[1]);
        (void)hipEventElapsedTime(&t.phase3_ms, tr[1], tr[2]);
        (void)hipEventElapsedTime(&t.total_ms, tr[0], tr[2]);
    }
    return (int)k;
}

int cjx_load(const char *path, cjx_input *out) {
    if (!out) return CJX_ERR_INPUT;
    out->data = nullptr;
    out->size = 0;
    FILE *f = path ? std::fopen(path, "rb") : nullptr;
    if (!f) return CJX_ERR_INPUT;  // load_file.cu:10-13
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (sz < 0) {
        std::fclose(f);
        return CJX_ERR_INPUT;
    }
    uint8_t *buf = nullptr;
    if (hipHostMalloc((void **)&buf, (size_t)sz + 1, hipHostMallocDefault) != hipSuccess) {
        std::fclose(f);
        return CJX_ERR_ALLOC;  // load_file.cu:23-27
    }
    size_t got = sz ? std::fread(buf, 1, (size_t)sz, f) : 0;
    std::fclose(f);
    if (got != (size_t)sz) {
        (void)hipHostFree(buf);
        return CJX_ERR_INPUT;
    }
    out->data = buf;
    out->size = (size_t)sz;
    return CJX_OK;
}

void cjx_free_input(cjx_input *in) {
    if (!in) return;
    if (in->data) (void)hipHostFree(in->data);
    in->data = nullptr;
    in->size = 0;
}

}  // extern "C"

namespace {

// Newline count and last byte of device text; the segment counts and bases
// stay in ctx->nl_seg for launch_nl_write.  Synchronous.
int nl_scan_total(cjx_context *ctx, const uint8_t *d_in, uint64_t n, uint64_t **seg_base,
                  uint64_t *total, uint8_t *last) {
    const uint64_t g = cjx::nl_segments(n);
    const size_t cnt_bytes = (size_t)((g * 4 + 15) & ~15ull);
    if (!ctx->nl_seg.ensure(cnt_bytes + (size_t)(g + 1) * 8 + 16) || !ctx->h_nl.ensure(64))
        return CJX_ERR_ALLOC;
    uint8_t *base = ctx->nl_seg.as<uint8_t>();
    uint32_t *cnt = reinterpret_cast<uint32_t *>(base);
    *seg_base = reinterpret_cast<uint64_t *>(base + cnt_bytes);
    uint64_t *stat = *seg_base + g + 1;
    uint64_t *h = static_cast<uint64_t *>(ctx->h_nl.p);
    if (cjx::launch_nl_count(ctx->stream, d_in, n, cnt, *seg_base, stat) != hipSuccess ||
        hipMemcpyAsync(h, stat, 16, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return CJX_ERR_DEVICE;
    *total = h[0];
    *last = (uint8_t)h[1];
    return CJX_OK;
}

bool device_text_ok(const uint8_t *d_in, size_t n) {
    return (d_in || !n) && (reinterpret_cast<uintptr_t>(d_in) & 15u) == 0;
}

}  // namespace

extern "C" {

int cjx_scan_lines_device(cjx_context *ctx, const uint8_t *d_input, size_t n,
                          uint64_t *d_newlines, size_t cap, size_t *n_newlines) {
    if (!n_newlines) return CJX_ERR_INPUT;
    *n_newlines = 0;
    if (!ctx || !device_text_ok(d_input, n)) return CJX_ERR_INPUT;
    if (int rc = set_device(ctx)) return rc;
    uint64_t *seg_base = nullptr, total = 0;
    uint8_t last = 0;
    if (int rc = nl_scan_total(ctx, d_input, n, &seg_base, &total, &last)) return rc;
    *n_newlines = (size_t)total;
    if (total > cap) return CJX_ERR_CAPACITY;
    if (total && !d_newlines) return CJX_ERR_INPUT;
    if (total && (cjx::launch_nl_write(ctx->stream, d_input, n, seg_base, d_newlines, cap) !=
                      hipSuccess ||
                  hipStreamSynchronize(ctx->stream) != hipSuccess))
        return CJX_ERR_DEVICE;
    return CJX_OK;
}

int cjx_chunk_lines_device(cjx_context *ctx, const uint8_t *d_input, size_t n, int mode,
                           size_t param, size_t *chunk_off, size_t *chunk_size, size_t cap,
                           size_t *n_chunks) {
    if (!n_chunks) return CJX_ERR_INPUT;
    *n_chunks = 0;
    if (!ctx || !device_text_ok(d_input, n)) return CJX_ERR_INPUT;
    if (!chunk_rule_ok(mode, param)) return CJX_ERR_INPUT;
    if (int rc = set_device(ctx)) return rc;
    uint64_t *seg_base = nullptr, total = 0;
    uint8_t last = 0;
    if (int rc = nl_scan_total(ctx, d_input, n, &seg_base, &total, &last)) return rc;
    // line starts: 0, the byte after each '\n', and n after an unterminated last line
    const uint64_t lines = total + (n && last != '\n' ? 1 : 0);
    hipStream_t s = ctx->stream;
    if (!ctx->nl_off.ensure((size_t)(total ? total : 1) * 8)) return CJX_ERR_ALLOC;
    uint64_t *d_nl = ctx->nl_off.as<uint64_t>();
    if (total && cjx::launch_nl_write(s, d_input, n, seg_base, d_nl, total) != hipSuccess)
        return CJX_ERR_DEVICE;
    ChunkList out;
    if (mode == 0) {
        // the count rule needs param + 1 line starts: gathered on the device
        std::vector<uint64_t> li;
        count_rule(lines, param, li);
        const size_t k = li.size(), bytes = k * 8;
        if (!ctx->nl_idx.ensure(2 * bytes) || !ctx->h_nl.ensure(bytes)) return CJX_ERR_ALLOC;
        uint64_t *d_idx = ctx->nl_idx.as<uint64_t>(), *d_at = d_idx + k;
        uint64_t *h = static_cast<uint64_t *>(ctx->h_nl.p);
        std::memcpy(h, li.data(), bytes);
        if (hipMemcpyAsync(d_idx, h, bytes, hipMemcpyHostToDevice, s) != hipSuccess ||
            cjx::launch_nl_gather(s, d_nl, total, n, d_idx, k, d_at) != hipSuccess ||
            hipMemcpyAsync(h, d_at, bytes, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return CJX_ERR_DEVICE;
        for (size_t i = 0; i < param; i++) out.emplace_back(h[i], h[i + 1] - h[i]);
    } else {
        // the size rule walks the line starts greedily: on the host, from the
        // device's newline offsets (line j >= 1 starts after newline j - 1)
        if (!ctx->h_nl.ensure((size_t)(total ? total : 1) * 8)) return CJX_ERR_ALLOC;
        const uint64_t *h = static_cast<const uint64_t *>(ctx->h_nl.p);
        if ((total && hipMemcpyAsync(ctx->h_nl.p, d_nl, total * 8, hipMemcpyDeviceToHost, s) !=
                          hipSuccess) ||
            hipStreamSynchronize(s) != hipSuccess)
            return CJX_ERR_DEVICE;
        auto at = [&](size_t j) -> uint64_t { return j == 0 ? 0 : (j <= total ? h[j - 1] + 1 : n); };
        bytes_rule(at, (size_t)lines + 1, n, param, out);
    }
    return emit_chunks(out, chunk_off, chunk_size, cap, n_chunks);
}

int cjx_values_device(cjx_context *ctx, const uint8_t *d_input, size_t n,
                      const int32_t *d_structural, size_t n_structural, cjx_value *d_values,
                      uint8_t *d_strings, size_t *n_exact) {
    if (n_exact) *n_exact = 0;
    if (!ctx || !d_input || n == 0 || n > 0xFFFFFFFFull || !d_values ||
        (n_structural && !d_structural) || n_structural >= n + 1)
        return CJX_ERR_INPUT;
    if (int rc = set_device(ctx)) return rc;
    // the list of numbers for the exact step, then its count
    if (!ctx->val_exact.ensure((n_structural + 1) * 4 + 16)) return CJX_ERR_ALLOC;
    uint32_t *list = ctx->val_exact.as<uint32_t>();
    uint32_t *count = list + n_structural + 1;
    hipStream_t s = ctx->stream;
    if (launch_values(s, d_input, n, d_structural, n_structural,
                      reinterpret_cast<ValueOut *>(d_values), d_strings, list, count) != hipSuccess)
        return CJX_ERR_DEVICE;
    if (n_exact) {
        if (!ctx->h_nl.ensure(64)) return CJX_ERR_ALLOC;
        if (hipMemcpyAsync(ctx->h_nl.p, count, 4, hipMemcpyDeviceToHost, s) != hipSuccess)
            return CJX_ERR_DEVICE;
    }
    if (hipStreamSynchronize(s) != hipSuccess) return CJX_ERR_DEVICE;
    if (n_exact) *n_exact = *static_cast<const uint32_t *>(ctx->h_nl.p);
    return CJX_OK;
}

int cjx_values(cjx_context *ctx, const uint8_t *data, size_t n, const cjx_result *r,
               cjx_value *values, uint8_t *strings) {
    if (!ctx || !data || n == 0 || n > 0xFFFFFFFFull || !r || !r->structural || !values ||
        r->n_structural < 0)
        return CJX_ERR_INPUT;
    if (int rc = set_device(ctx)) return rc;
    const size_t N = (size_t)r->n_structural;
    // device copies: the text, its structural row, the table and the strings
    if (!ctx->in.ensure(n + 16) || !ctx->structural.ensure(N * 4 + 16) ||
        !ctx->layout.ensure((N + 1) * sizeof(cjx_value) + (strings ? n : 0) + 16))
        return CJX_ERR_ALLOC;
    hipStream_t s = ctx->stream;
    cjx_value *d_values = ctx->layout.as<cjx_value>();
    uint8_t *d_strings = strings ? reinterpret_cast<uint8_t *>(d_values + N + 1) : nullptr;
    if (hipMemcpyAsync(ctx->in.p, data, n, hipMemcpyHostToDevice, s) != hipSuccess ||
        (N && hipMemcpyAsync(ctx->structural.p, r->structural + 1, N * 4, hipMemcpyHostToDevice,
                             s) != hipSuccess))
        return CJX_ERR_DEVICE;
    int rc = cjx_values_device(ctx, ctx->in.as<uint8_t>(), n, ctx->structural.as<int32_t>(), N,
                               d_values, d_strings, nullptr);
    if (rc) return rc;
    if (hipMemcpyAsync(values, d_values, (N + 1) * sizeof(cjx_value), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        (strings &&
         hipMemcpyAsync(strings, d_strings, n, hipMemcpyDeviceToHost, s) != hipSuccess) ||
        hipStreamSynchronize(s) != hipSuccess)
        return CJX_ERR_DEVICE;
    return CJX_OK;
}

int cjx_load_device(cjx_context *ctx, const char *path, uint8_t *d_buf, size_t cap, size_t *n) {
    if (!n) return CJX_ERR_INPUT;
    *n = 0;
    if (!ctx || !path) return CJX_ERR_INPUT;
    const int fd = ::open(path, O_RDONLY);
    if (fd < 0) return CJX_ERR_INPUT;  // load_file.cu:10-13
    struct stat st;
    if (::fstat(fd, &st) != 0 || st.st_size < 0) {
        ::close(fd);
        return CJX_ERR_INPUT;
    }
    const uint64_t size = (uint64_t)st.st_size;
    *n = (size_t)size;
    int rc = CJX_OK;
    if (size > cap) rc = CJX_ERR_CAPACITY;
    else if (size && !d_buf) rc = CJX_ERR_INPUT;
    else rc = set_device(ctx);
    if (!rc) rc = ensure_copy_stream(ctx);
    for (int k = 0; !rc && k < cjx_context::kLoadBufs && (uint64_t)k * cjx_context::kLoadPart < size; k++)
        if (!ctx->load_buf[k].ensure(cjx_context::kLoadPart)) rc = CJX_ERR_ALLOC;
    // the copies start once earlier work on the context stream (which may
    // still read the buffer) is done
    if (!rc && (hipEventRecord(ctx->idle_ev, ctx->stream) != hipSuccess ||
                hipStreamWaitEvent(ctx->copy_stream, ctx->idle_ev, 0) != hipSuccess))
        rc = CJX_ERR_DEVICE;
    // Part i is read into staging buffer i % kLoadBufs once that buffer's
    // previous copy has finished, so the file reads overlap the copies.
    uint64_t part = 0;
    for (uint64_t off = 0; !rc && off < size; off += cjx_context::kLoadPart, part++) {
        const int k = (int)(part % cjx_context::kLoadBufs);
        const uint64_t len = std::min<uint64_t>(cjx_context::kLoadPart, size - off);
        if (part >= (uint64_t)cjx_context::kLoadBufs &&
            hipEventSynchronize(ctx->load_ev[k]) != hipSuccess) {
            rc = CJX_ERR_DEVICE;
            break;
        }
        uint8_t *buf = static_cast<uint8_t *>(ctx->load_buf[k].p);
        for (uint64_t got = 0; got < len;) {
            const ssize_t r = ::pread(fd, buf + got, (size_t)(len - got), (off_t)(off + got));
            if (r < 0 && errno == EINTR) continue;
            if (r <= 0) {
                rc = CJX_ERR_INPUT;  // a short file: read error (load_file.cu:29-34)
                break;
            }
            got += (uint64_t)r;
        }
        if (!rc && (hipMemcpyAsync(d_buf + off, buf, len, hipMemcpyHostToDevice,
                                   ctx->copy_stream) != hipSuccess ||
                    hipEventRecord(ctx->load_ev[k], ctx->copy_stream) != hipSuccess))
            rc = CJX_ERR_DEVICE;
    }
    ::close(fd);
    // every queued copy finishes before the staging buffers can be reused
    if (ctx->copy_stream && hipStreamSynchronize(ctx->copy_stream) != hipSuccess && !rc)
        rc = CJX_ERR_DEVICE;
    return rc;
}

int cjx_parse_standard(cjx_context *ctx, const uint8_t *data, size_t n, unsigned flags,
                       cjx_result *out) {
    if (!out) return CJX_ERR_INPUT;
    zero_result(out);
    if (!ctx || !data || n == 0) return out->error = CJX_ERR_INPUT;  // PSJ:1597-1600
    std::vector<uint64_t> offs{0}, sizes{n}, pos{0};
    int rc = parse_host_units(ctx, data, offs, sizes, pos, flags & ~CJX_FLAG_JSONL, out);
    out->error = rc;
    return rc;
}

int cjx_parse_lines(cjx_context *ctx, const uint8_t *data, size_t n, const size_t *chunk_off,
                    const size_t *chunk_size, size_t n_chunks, unsigned flags,
                    cjx_result *out) {
    if (!out) return CJX_ERR_INPUT;
    zero_result(out);
    // Argument checks of parse_json_lines.cu:1103-1150.
    if (!ctx || n == 0 || !data || n_chunks == 0 || !chunk_off || !chunk_size)
        return out->error = CJX_ERR_INPUT;
    std::vector<uint64_t> offs(n_chunks), sizes(n_chunks), pos(n_chunks);
    uint64_t last_chunk_index = 0;
    for (size_t i = 0; i < n_chunks; i++) {
        if (chunk_size[i] == 0 || chunk_size[i] > n || chunk_off[i] > n ||
            chunk_off[i] + chunk_size[i] > n) {
            out->error = CJX_ERR_INPUT;
            out->error_chunk = (int32_t)i;
            return CJX_ERR_INPUT;
        }
        offs[i] = chunk_off[i];
        sizes[i] = chunk_size[i];
        pos[i] = last_chunk_index;  // lastChunkIndex, parse_json_lines.cu:1223
        last_chunk_index += chunk_size[i];
    }
    int rc = parse_host_units(ctx, data, offs, sizes, pos, flags | CJX_FLAG_JSONL, out);
    out->error = rc;
    return rc;
}

void cjx_free_result(cjx_result *r) {
    if (!r) return;
    pinned_put(r->buf);
    pinned_put(r->records);
    std::free(r->result_sizes);
    std::free(r->result_sizes_prefix);
    std::memset(r, 0, sizeof(*r));
    r->error_chunk = -1;
}

int cjx_parse_device(cjx_context *ctx, const uint8_t *d_input, size_t n, unsigned flags,
                     int32_t *d_structural, int32_t *d_pair, size_t cap) {
    if (!ctx || !d_input || n == 0 || !d_structural || !d_pair) return CJX_ERR_INPUT;
    if (n > kMaxLaunchBytes || (reinterpret_cast<uintptr_t>(d_input) & 15u)) return CJX_ERR_INPUT;
    int rc = set_device(ctx);
    if (rc) return rc;
    std::vector<UnitDesc> units;
    uint64_t dev_bytes = 0;
    uint32_t ntiles = 0, n_small = 0;
    if (!plan_units({n}, {0}, units, dev_bytes, ntiles, n_small)) return CJX_ERR_INPUT;
    units[0].out_off = 0;
    units[0].out_cap = (uint32_t)std::min<uint64_t>(cap, n);
    Shape sh{ntiles, 1, n_small, 0, 0, 0, 0, false};
    shape_rows(sh, units);
    if ((rc = ensure_work(ctx, sh))) return rc;
    // The reset kernel writes the one-unit table from its arguments: no
    // upload, so no copy-engine round trip before the launch.
    if (!ctx->units.ensure(sizeof(UnitDesc) + ((size_t)ntiles + 1) * 4 + 16)) return CJX_ERR_ALLOC;
    ctx->tile_unit = reinterpret_cast<uint32_t *>(ctx->units.as<uint8_t>() + sizeof(UnitDesc));
    ctx->small_units = ctx->tile_unit + ntiles;
    ctx->unit_src = nullptr;
    ctx->one_unit = units[0];
    ctx->one_unit_dev = true;
    ctx->n_tiny = (n_small && n <= (uint64_t)kRoundBytes) ? 1u : 0u;
    ctx->n_units = 1;
    (void)hipEventRecord(ctx->ev0, ctx->stream);
    rc = enqueue_maybe_graph(ctx, d_input, sh, flags & ~CJX_FLAG_JSONL,
                       reinterpret_cast<uint32_t *>(d_structural), d_pair);
    ctx->one_unit_dev = false;
    (void)hipEventRecord(ctx->ev1, ctx->stream);
    ctx->events_recorded = true;
    return rc;
}

int cjx_device_wait(cjx_context *ctx, cjx_device_status *st) {
    if (!ctx) return CJX_ERR_INPUT;
    int rc = read_status(ctx, ctx->n_units ? ctx->n_units : 1);
    if (rc) return rc;
    const UnitOut *uo = static_cast<const UnitOut *>(ctx->h_uout.p);
    const LaunchState *ls = static_cast<const LaunchState *>(ctx->h_ls.p);
    int code = unit_code(uo[0]);
    if (ls->overflow) code = CJX_ERR_CAPACITY;
    if (st) {
        st->n_structural = ls->n_total;
        st->n_brackets = ls->m_total;
        st->error = code;
        st->capacity_exceeded = ls->overflow ? 1 : 0;
    }
    return code;
}

int cjx_parse_batch(cjx_context *ctx, const uint8_t *data, const size_t *doc_off,
                    const size_t *doc_size, size_t n_docs, unsigned flags,
                    cjx_batch_result *out) {
    if (!out) return CJX_ERR_INPUT;
    std::memset(out, 0, sizeof(*out));
    if (!ctx || !data || !doc_off || !doc_size || n_docs == 0) return out->error = CJX_ERR_INPUT;
    std::vector<uint64_t> offs(n_docs), sizes(n_docs);
    for (size_t i = 0; i < n_docs; i++) {
        if (doc_size[i] == 0) return out->error = CJX_ERR_INPUT;  // PSJ:1597-1600 per document
        offs[i] = doc_off[i];
        sizes[i] = doc_size[i];
    }
    const int rc = parse_host_batch(ctx, data, offs, sizes, flags & ~CJX_FLAG_JSONL, out);
    if (rc) out->error = rc;
    return rc;
}

void cjx_free_batch_result(cjx_batch_result *r) {
    if (!r) return;
    pinned_put(r->structural);
    std::free(r->docs);
    std::memset(r, 0, sizeof(*r));
}

int cjx_parse_batch_device(cjx_context *ctx, const uint8_t *d_input, const size_t *doc_off,
                           const size_t *doc_size, size_t n_docs, unsigned flags,
                           int32_t *d_structural, int32_t *d_pair, size_t cap) {
    return parse_units_device(ctx, d_input, doc_off, doc_size, n_docs,
                              flags & ~(CJX_FLAG_JSONL | CJX_FLAG_RECORDS), d_structural, d_pair,
                              cap, nullptr, 0);
}

int cjx_parse_lines_device(cjx_context *ctx, const uint8_t *d_input, const size_t *chunk_off,
                           const size_t *chunk_size, size_t n_chunks, unsigned flags,
                           int32_t *d_structural, int32_t *d_pair, size_t cap,
                           cjx_record *d_records, size_t max_records) {
    return parse_units_device(ctx, d_input, chunk_off, chunk_size, n_chunks,
                              flags | CJX_FLAG_JSONL, d_structural, d_pair, cap, d_records,
                              max_records);
}

}  // extern "C"

namespace cjx_host {

// Units at offsets of a device buffer.  Independent documents (no JSONL
// flag): positions and pair values local to each unit.  JSON Lines chunks:
// positions count from the first chunk (lastChunkIndex) and pair values are
// numbered across chunks, as in the host layout.  When a unit does not start
// at a 16-byte aligned offset, all units are first copied into a packed,
// aligned layout in the context's input buffer (k_place_units), which the
// tokenizer then reads.
int parse_units_device(cjx_context *ctx, const uint8_t *d_input, const size_t *doc_off,
                       const size_t *doc_size, size_t n_docs, unsigned flags,
                       int32_t *d_structural, int32_t *d_pair, size_t cap,
                       cjx_record *d_records, size_t max_records) {
    if (!ctx || !d_input || !doc_off || !doc_size || n_docs == 0 || !d_structural || !d_pair)
        return CJX_ERR_INPUT;
    const bool jsonl = (flags & CJX_FLAG_JSONL) != 0;
    int rc = set_device(ctx);
    if (rc) return rc;
    std::vector<UnitDesc> units(n_docs);
    uint64_t total = 0, full = 0, placed = 0, src_end = 0;
    // Addresses, not offsets, decide: relocation reads from d_input rounded
    // down to 16 bytes (offsets shifted by the difference).
    const uint64_t shift = reinterpret_cast<uintptr_t>(d_input) & 15u;
    bool relocate = false;
    for (size_t i = 0; i < n_docs; i++) {
        if (doc_size[i] == 0) return CJX_ERR_INPUT;
        if ((doc_off[i] + shift) & 15) relocate = true;
    }
    std::vector<uint64_t> src;
    for (size_t i = 0; i < n_docs; i++) {
        UnitDesc &u = units[i];
        u.in_off = relocate ? placed : doc_off[i];
        placed += (doc_size[i] + 15) & ~15ull;
        if (relocate) src.push_back(doc_off[i] + shift);
        src_end = std::max<uint64_t>(src_end, (doc_off[i] + shift + doc_size[i] + 15) & ~15ull);
        u.n = doc_size[i];
        u.pos_base = jsonl ? total : 0;
        total += doc_size[i];
        full += doc_size[i];
    }
    uint32_t n_small = 0;
    const uint64_t tiles = assign_tiles(units, n_small);
    if (tiles > 0xFFFFFFF0ull || total > kMaxLaunchBytes) return CJX_ERR_INPUT;
    // Regions: each document's exact bound when the caller's rows hold them
    // all, else the n/3 + 256 guess (an overflow reports CJX_ERR_CAPACITY).
    const bool exact = full <= cap;
    if (set_regions(units, exact) > cap) return CJX_ERR_CAPACITY;
    Shape sh{(uint32_t)tiles, (uint32_t)n_docs, n_small, 0, 0, 0, 0, jsonl};
    shape_rows(sh, units);
    if ((rc = ensure_work(ctx, sh))) return rc;
    if (relocate && !ctx->in.ensure(placed + 16)) return CJX_ERR_ALLOC;
    if ((rc = upload_units(ctx, units, (uint32_t)tiles, n_small, relocate ? src.data() : nullptr)))
        return rc;
    ctx->n_units = (uint32_t)n_docs;
    ctx->batch_off.assign(doc_off, doc_off + n_docs);
    (void)hipEventRecord(ctx->ev0, ctx->stream);
    const uint8_t *d_text = d_input;
    if (relocate) {
        if (launch_place_units(ctx->stream, d_input - shift, src_end, ctx->unit_src,
                               ctx->units.as<UnitDesc>(), (uint32_t)n_docs,
                               ctx->in.as<uint8_t>(), placed) != hipSuccess)
            return CJX_ERR_DEVICE;
        d_text = ctx->in.as<uint8_t>();
    }
    const bool records = jsonl && (flags & CJX_FLAG_RECORDS) && d_records;
    if (records && (!ctx->row_nl.ensure(sh.rows + 16) || !ctx->rec_count.ensure(n_docs * 4 + 16)))
        return CJX_ERR_ALLOC;
    ctx->want_nl = records;  // the tokenizer also writes the newline flags
    rc = enqueue_maybe_graph(ctx, d_text, sh, flags & ~CJX_FLAG_RECORDS,
                             reinterpret_cast<uint32_t *>(d_structural), d_pair);
    if (!rc && records) {
        const LaunchBufs b = bufs_of(ctx);
        if (launch_records(ctx->stream, b, (uint32_t)n_docs,
                           reinterpret_cast<uint32_t *>(d_structural),
                           ctx->rec_count.as<uint32_t>(), nullptr, 0, false) != hipSuccess ||
            launch_records(ctx->stream, b, (uint32_t)n_docs,
                           reinterpret_cast<uint32_t *>(d_structural),
                           ctx->rec_count.as<uint32_t>(), reinterpret_cast<RecordOut *>(d_records),
                           max_records, true) != hipSuccess)
            rc = CJX_ERR_DEVICE;
        // the timing ring's end event of this launch moves behind the table
        if (!rc && ctx->timing && ctx->timing_count)
            (void)hipEventRecord(
                ctx->tev[(ctx->timing_count - 1) % cjx_context::kTimingRing][2], ctx->stream);
    }
    ctx->want_nl = false;
    (void)hipEventRecord(ctx->ev1, ctx->stream);
    ctx->events_recorded = true;
    return rc;
}

}  // namespace cjx_host

extern "C" {

int cjx_batch_wait(cjx_context *ctx, cjx_doc_result *docs, size_t n_docs, cjx_device_status *st) {
    if (!ctx || ctx->n_units == 0) return CJX_ERR_INPUT;
    int rc = read_status(ctx, ctx->n_units);
    if (rc) return rc;
    const UnitOut *uo = static_cast<const UnitOut *>(ctx->h_uout.p);
    const LaunchState *ls = static_cast<const LaunchState *>(ctx->h_ls.p);
    int code = CJX_OK;
    for (uint32_t i = 0; i < ctx->n_units; i++) {
        const int c = unit_code(uo[i]);
        if (c && !code) code = c;
        if (docs && i < n_docs) {
            docs[i].byte_base = i < ctx->batch_off.size() ? (int64_t)ctx->batch_off[i] : 0;
            docs[i].struct_begin = uo[i].n_base;
            docs[i].struct_end = uo[i].n_end;
            docs[i].error = c;
        }
    }
    if (ls->overflow) code = CJX_ERR_CAPACITY;
    if (st) {
        st->n_structural = ls->n_total;
        st->n_brackets = ls->m_total;
        st->error = code;
        st->capacity_exceeded = (ls->overflow || ls->rec_overflow) ? 1 : 0;
        st->n_records = ls->n_records;
    }
    return (ls->overflow || ls->rec_overflow) ? CJX_ERR_CAPACITY : CJX_OK;
}

}  // extern "C"
