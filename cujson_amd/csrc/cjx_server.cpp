// cjx_server.cpp - request server with continuous batching (no counterpart
// in the reference, which parses one document per synchronous call).
//
// Callers submit documents from any thread; a worker thread keeps a set of
// slots (one context, i.e. one HIP stream and its resident buffers, each).
// Whenever a slot is idle, the worker takes the pending requests in arrival
// order that fit the slot's share of the device-memory budget (and share
// the first one's flags), up to max_batch of them, and runs them as one
// batch (every document its own parse unit) on that slot.  Batches on
// different slots overlap on the device; requests that arrive meanwhile wait
// only for a free slot, never for the other batches to finish.  A finished
// batch is laid out on the device as one reference-layout buffer per request
// (the cuJSONResult layout of parse_standard_json) and copied out once into a
// pinned region; each completion points into that region, which returns to
// the server's pool when the last of its completions is released.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/cujson_amd.h"
#include "cjx_host.h"
#include "cjx_internal.h"

using namespace cjx;
using namespace cjx_host;

namespace {

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
    uint64_t ticket;
    const uint8_t *data;
    size_t n;
    unsigned flags;
    double t_submit, t_launch;
    size_t cost;
};

// Device bytes a request holds while in flight: its input and, per guessed
// output row (n/3 + 256), the structural and pair rows, bracket arrays, sort
// scratch, matching tables and the laid-out copy (about 40 bytes).
size_t request_cost(size_t n) { return n + (n / 3 + 256) * 40 + 4096; }

constexpr size_t kStageSmall = 64 << 10;  // smaller inputs go through the staging copy

// A pinned host region holding the laid-out results of one batch.
struct Region {
    void *p = nullptr;
    size_t cap = 0;
    size_t refs = 0;
};

struct Slot {
    cjx_context *ctx = nullptr;
    enum State { kIdle, kKernels, kCopyOut } state = kIdle;
    std::vector<Request> reqs;
    std::vector<UnitDesc> units;
    Shape sh{};
    unsigned flags = 0;
    bool full_regions = false;
    size_t cost = 0;
    uint64_t N = 0;
    Region *region = nullptr;
    hipEvent_t ev = nullptr;
    PinnedBuf stage;
};

}  // namespace

struct cjx_server {
    int device = 0;
    size_t slot_budget = 0, max_batch = 0, in_use = 0;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request> pending;
    std::deque<cjx_completion> done;
    std::vector<Region *> regions, free_regions;  // all regions / the unused ones
    uint64_t next_ticket = 1;
    bool stop = false;
    std::thread worker;
    uint64_t batches = 0, docs = 0, bytes = 0, max_inflight = 0, relaunches = 0;
};

namespace {

// Smallest free region of at least `bytes`, else a new one (sizes are powers
// of two from 1 MiB).  Called with the server lock held.
Region *acquire_region(cjx_server *srv, size_t bytes) {
    Region *best = nullptr;
    size_t bi = 0;
    for (size_t i = 0; i < srv->free_regions.size(); i++) {
        Region *r = srv->free_regions[i];
        if (r->cap >= bytes && (!best || r->cap < best->cap)) {
            best = r;
            bi = i;
        }
    }
    if (best) {
        srv->free_regions.erase(srv->free_regions.begin() + (ptrdiff_t)bi);
        return best;
    }
    size_t cap = 1 << 20;
    while (cap < bytes) cap <<= 1;
    Region *r = new Region();
    if (hipHostMalloc(&r->p, cap, hipHostMallocDefault) != hipSuccess) {
        delete r;
        return nullptr;
    }
    r->cap = cap;
    srv->regions.push_back(r);
    return r;
}

// Enqueues the batch of slot s (inputs, kernels, status copies) and records
// its event.
int launch_batch(Slot &s) {
    cjx_context *ctx = s.ctx;
    hipStream_t st = ctx->stream;
    const size_t k = s.reqs.size();
    std::vector<UnitDesc> &units = s.units;
    units.assign(k, UnitDesc{});
    uint64_t off = 0;
    for (size_t i = 0; i < k; i++) {
        units[i].in_off = off;
        units[i].n = s.reqs[i].n;
        units[i].pos_base = 0;
        off += (s.reqs[i].n + 15) & ~15ull;
    }
    uint32_t n_small = 0;
    const uint64_t tiles = assign_tiles(units, n_small);
    set_regions(units, s.full_regions);
    s.sh = Shape{(uint32_t)tiles, (uint32_t)k, n_small, 0, 0, 0, 0, false};
    shape_rows(s.sh, units);
    int rc;
    if ((rc = ensure_work(ctx, s.sh))) return rc;
    if (!ctx->in.ensure(off + 64) || !ctx->structural.ensure(s.sh.rows * 4 + 16) ||
        !ctx->pair.ensure(s.sh.rows * 4 + 16) || !s.stage.ensure(off + 64))
        return CJX_ERR_ALLOC;
    // Small inputs are gathered in the slot's pinned staging buffer and copied
    // in runs of consecutive staged documents; large ones are copied straight
    // from the caller's memory.
    uint8_t *stage = static_cast<uint8_t *>(s.stage.p);
    uint8_t *d_in = ctx->in.as<uint8_t>();
    uint64_t lo = 0, hi = 0;  // pending staged run
    for (size_t i = 0; i <= k; i++) {
        const bool staged = i < k && s.reqs[i].n < kStageSmall;
        if (!staged && hi > lo) {
            if (hipMemcpyAsync(d_in + lo, stage + lo, hi - lo, hipMemcpyHostToDevice, st) !=
                hipSuccess)
                return CJX_ERR_DEVICE;
            lo = hi = 0;
        }
        if (i == k) break;
        const Request &r = s.reqs[i];
        if (staged) {
            std::memcpy(stage + units[i].in_off, r.data, r.n);
            if (hi == lo) lo = units[i].in_off;
            hi = units[i].in_off + r.n;
        } else if (hipMemcpyAsync(d_in + units[i].in_off, r.data, r.n, hipMemcpyHostToDevice,
                                  st) != hipSuccess) {
            return CJX_ERR_DEVICE;
        }
    }
    if ((rc = upload_units(ctx, units, (uint32_t)tiles, n_small))) return rc;
    if ((rc = enqueue_parse(ctx, d_in, s.sh, s.flags, ctx->structural.as<uint32_t>(),
                            ctx->pair.as<int32_t>())))
        return rc;
    if (hipMemcpyAsync(ctx->h_uout.p, ctx->uout.p, k * sizeof(UnitOut), hipMemcpyDeviceToHost,
                       st) != hipSuccess ||
        hipMemcpyAsync(ctx->h_ls.p, ctx->ls.p, sizeof(LaunchState), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipEventRecord(s.ev, st) != hipSuccess)
        return CJX_ERR_DEVICE;
    const double t = now_ms();
    for (Request &r : s.reqs) r.t_launch = t;
    s.state = Slot::kKernels;
    return CJX_OK;
}

// Kernels done: relaunch with exact regions on overflow, else lay the
// results out per request and copy them into a pinned region.
int after_kernels(cjx_server *srv, Slot &s) {
    const LaunchState *ls = static_cast<const LaunchState *>(s.ctx->h_ls.p);
    if (ls->overflow) {
        s.full_regions = true;
        srv->relaunches++;
        return launch_batch(s);
    }
    s.N = ls->n_total;
    const size_t k = s.reqs.size();
    const size_t cells = 2 * s.N + 4 * k;
    {
        std::lock_guard<std::mutex> g(srv->mu);
        s.region = acquire_region(srv, cells * 4);
    }
    if (!s.region || !s.ctx->layout.ensure(cells * 4 + 16)) return CJX_ERR_ALLOC;
    hipStream_t st = s.ctx->stream;
    if (launch_assemble_docs(st, bufs_of(s.ctx), (uint32_t)k, s.ctx->structural.as<uint32_t>(),
                             s.ctx->pair.as<int32_t>(), s.ctx->layout.as<int32_t>(),
                             (uint32_t)s.N) != hipSuccess ||
        hipMemcpyAsync(s.region->p, s.ctx->layout.p, cells * 4, hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipEventRecord(s.ev, st) != hipSuccess)
        return CJX_ERR_DEVICE;
    s.state = Slot::kCopyOut;
    return CJX_OK;
}

// The region holds every request's buffer: hand them out.
void finish_batch(cjx_server *srv, Slot &s) {
    const UnitOut *uo = static_cast<const UnitOut *>(s.ctx->h_uout.p);
    int32_t *base = static_cast<int32_t *>(s.region->p);
    const double t = now_ms();
    std::lock_guard<std::mutex> g(srv->mu);
    s.region->refs = s.reqs.size();
    for (size_t i = 0; i < s.reqs.size(); i++) {
        cjx_completion c;
        std::memset(&c, 0, sizeof(c));
        c.ticket = s.reqs[i].ticket;
        c.error = unit_code(uo[i]);
        c.n_structural = (int64_t)(uo[i].n_end - uo[i].n_base);
        c.buf = base + 2 * (uint64_t)uo[i].row_base + 4 * i;
        c.owner = s.region;
        c.latency_ms = t - s.reqs[i].t_submit;
        c.queue_ms = s.reqs[i].t_launch - s.reqs[i].t_submit;
        srv->done.push_back(c);
    }
    srv->in_use -= s.cost;
    srv->docs += s.reqs.size();
    s.reqs.clear();
    s.region = nullptr;
    s.state = Slot::kIdle;
    srv->cv_done.notify_all();
}

void fail_batch(cjx_server *srv, Slot &s, int rc) {
    std::lock_guard<std::mutex> g(srv->mu);
    for (const Request &r : s.reqs) {
        cjx_completion c;
        std::memset(&c, 0, sizeof(c));
        c.ticket = r.ticket;
        c.error = rc;
        srv->done.push_back(c);
    }
    if (s.region) srv->free_regions.push_back(s.region);
    s.region = nullptr;
    srv->in_use -= s.cost;
    s.reqs.clear();
    s.state = Slot::kIdle;
    srv->cv_done.notify_all();
}

// Moves pending requests into the idle slot s: the oldest one, then the
// following ones with the same flags, in arrival order, while they fit the
// slot's budget share and max_batch (requests with other flags keep their
// places).  Returns false when nothing is pending.  Called with the server
// lock held.
bool admit(cjx_server *srv, Slot &s) {
    if (srv->pending.empty()) return false;
    const unsigned flags = srv->pending.front().flags;
    size_t cost = 0;
    std::deque<Request> &q = srv->pending;
    size_t keep = 0;  // requests left in place are compacted to q[0, keep)
    size_t i = 0;
    for (; i < q.size(); i++) {
        const Request &r = q[i];
        if (r.flags == flags) {
            // a request larger than the slot share still runs, alone
            if (s.reqs.size() == srv->max_batch ||
                (!s.reqs.empty() && cost + r.cost > srv->slot_budget))
                break;
            cost += r.cost;
            s.reqs.push_back(r);
        } else {
            q[keep++] = r;
        }
    }
    for (; i < q.size(); i++) q[keep++] = q[i];
    q.resize(keep);
    srv->in_use += cost;
    s.cost = cost;
    s.flags = flags;
    srv->batches++;
    srv->max_inflight = std::max<uint64_t>(srv->max_inflight, srv->in_use);
    for (const Request &r : s.reqs) srv->bytes += r.n;
    return true;
}

void worker_main(cjx_server *srv) {
    (void)hipSetDevice(srv->device);
    for (;;) {
        bool progressed = false;
        for (Slot &s : srv->slots) {
            if (s.state == Slot::kIdle) continue;
            const hipError_t q = hipEventQuery(s.ev);
            if (q == hipErrorNotReady) continue;
            progressed = true;
            if (q != hipSuccess) {
                fail_batch(srv, s, CJX_ERR_DEVICE);
                continue;
            }
            if (s.state == Slot::kKernels) {
                const int rc = after_kernels(srv, s);
                if (rc) fail_batch(srv, s, rc);
            } else {
                finish_batch(srv, s);
            }
        }
        for (Slot &s : srv->slots) {
            if (s.state != Slot::kIdle) continue;
            {
                std::lock_guard<std::mutex> g(srv->mu);
                if (!admit(srv, s)) break;
            }
            s.full_regions = false;
            const int rc = launch_batch(s);
            if (rc) fail_batch(srv, s, rc);
            progressed = true;
        }
        std::unique_lock<std::mutex> g(srv->mu);
        bool busy = false;
        for (const Slot &s : srv->slots) busy |= s.state != Slot::kIdle;
        if (srv->stop && !busy && srv->pending.empty()) break;
        if (!progressed) {
            if (busy) {  // poll the slot events
                g.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(20));
            } else {
                srv->cv_work.wait_for(g, std::chrono::milliseconds(5));
            }
        }
    }
}

}  // namespace

extern "C" {

int cjx_server_create(int device, size_t device_budget, int n_slots, size_t max_batch,
                      cjx_server **out) {
    if (!out) return CJX_ERR_INPUT;
    *out = nullptr;
    if (n_slots <= 0 || n_slots > 16 || device_budget == 0 || max_batch == 0 ||
        max_batch > (1u << 20))
        return CJX_ERR_INPUT;
    cjx_server *srv = new cjx_server();
    srv->device = device;
    srv->slot_budget = device_budget / (size_t)n_slots;
    srv->max_batch = max_batch;
    srv->slots.resize((size_t)n_slots);
    for (Slot &s : srv->slots) {
        int rc = cjx_context_create(device, &s.ctx);
        if (!rc && hipEventCreateWithFlags(&s.ev, hipEventDisableTiming) != hipSuccess)
            rc = CJX_ERR_DEVICE;
        if (rc) {
            for (Slot &t : srv->slots) {
                if (t.ev) (void)hipEventDestroy(t.ev);
                if (t.ctx) cjx_context_destroy(t.ctx);
            }
            delete srv;
            return rc;
        }
    }
    srv->worker = std::thread(worker_main, srv);
    *out = srv;
    return CJX_OK;
}

int cjx_server_submit(cjx_server *srv, const uint8_t *data, size_t n, unsigned flags,
                      uint64_t *ticket) {
    if (!srv || !data || n == 0 || n > kMaxLaunchBytes) return CJX_ERR_INPUT;
    Request r;
    r.data = data;
    r.n = n;
    r.flags = flags & (CJX_FLAG_UTF8_STRICT | CJX_FLAG_DEPTH_U16);
    r.t_submit = now_ms();
    r.t_launch = 0;
    r.cost = request_cost(n);
    {
        std::lock_guard<std::mutex> g(srv->mu);
        if (srv->stop) return CJX_ERR_INPUT;
        r.ticket = srv->next_ticket++;
        srv->pending.push_back(r);
    }
    if (ticket) *ticket = r.ticket;
    srv->cv_work.notify_one();
    return CJX_OK;
}

int cjx_server_wait(cjx_server *srv, cjx_completion *out, size_t cap, size_t *n_out,
                    int timeout_ms) {
    if (!srv || !out || !n_out) return CJX_ERR_INPUT;
    std::unique_lock<std::mutex> g(srv->mu);
    if (timeout_ms > 0 && srv->done.empty())
        srv->cv_done.wait_for(g, std::chrono::milliseconds(timeout_ms),
                              [&] { return !srv->done.empty(); });
    size_t k = 0;
    while (k < cap && !srv->done.empty()) {
        out[k++] = srv->done.front();
        srv->done.pop_front();
    }
    *n_out = k;
    return CJX_OK;
}

void cjx_server_release(cjx_server *srv, cjx_completion *c) {
    if (!srv || !c || !c->owner) return;
    Region *r = static_cast<Region *>(c->owner);
    std::lock_guard<std::mutex> g(srv->mu);
    if (r->refs && --r->refs == 0) srv->free_regions.push_back(r);
    c->owner = nullptr;
    c->buf = nullptr;
}

int cjx_server_get_stats(cjx_server *srv, cjx_server_stats *st) {
    if (!srv || !st) return CJX_ERR_INPUT;
    std::lock_guard<std::mutex> g(srv->mu);
    st->batches = srv->batches;
    st->documents = srv->docs;
    st->bytes = srv->bytes;
    st->max_inflight_bytes = srv->max_inflight;
    st->relaunches = srv->relaunches;
    st->pending = srv->pending.size();
    st->region_bytes = 0;
    for (const Region *r : srv->regions) st->region_bytes += r->cap;
    return CJX_OK;
}

void cjx_server_destroy(cjx_server *srv) {
    if (!srv) return;
    {
        std::lock_guard<std::mutex> g(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    srv->worker.join();
    for (Slot &s : srv->slots) {
        s.stage.release();
        if (s.ev) (void)hipEventDestroy(s.ev);
        if (s.ctx) cjx_context_destroy(s.ctx);
    }
    for (Region *r : srv->regions) {
        (void)hipHostFree(r->p);
        delete r;
    }
    delete srv;
}

}  // extern "C"
