// cjx_server.cpp - request server with continuous batching (no counterpart
// in the reference, which parses one document per synchronous call).
//
// Callers submit documents from any thread.  The server keeps a set of
// slots, each one context (one HIP stream and its resident buffers) driven
// by its own host thread.  Whenever a slot is idle, it takes the oldest
// pending request and the following ones with the same flags, in arrival
// order, while they fit the slot's share of the device-memory budget,
// max_batch and a byte cap, and runs them as one batch (every document its
// own parse unit).  Half of the slots are express slots: they take only
// requests below 4 MiB, in batches of at most 16 MiB, so small requests never
// wait behind the transfers of very large documents.  Batches on different
// slots overlap on the device; requests that arrive meanwhile wait only for
// a free slot, never for the other batches to finish.
//
// A finished batch is laid out as one reference-layout buffer per request
// (the cuJSONResult layout of parse_standard_json) in a pinned region; each
// completion points into that region, which returns to the server's pool
// when the last of its completions is released.  Batches up to 16 MiB never
// touch the copy engines: their inputs are gathered in pinned memory and
// pulled in by a copy kernel, and their laid-out results are pushed into the
// region by another, so they do not queue behind the large copies of other
// slots' batches.  Status words always come back through a copy kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
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
constexpr size_t kZeroCopyBytes = 16 << 20;     // batch inputs / results moved by kernels
constexpr size_t kExpressDocBytes = 4 << 20;    // requests below this may take express slots
constexpr size_t kExpressBatchBytes = 16 << 20; // input bytes of an express batch
constexpr size_t kBatchBytes = 64 << 20;        // input bytes of any other batch (or one request)

// A pinned host region holding the laid-out results of one batch.
struct Region {
    void *p = nullptr;   // host address
    void *dp = nullptr;  // the same memory as kernels address it
    size_t cap = 0;
    size_t refs = 0;
};

// Device address of pinned host memory.
void *device_view(void *h) {
    void *d = nullptr;
    return hipHostGetDevicePointer(&d, h, 0) == hipSuccess ? d : nullptr;
}

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
    bool express = false;
    PinnedBuf stage, status;  // coherent: kernels read / write them directly
    const UnitOut *uo = nullptr;       // in `status`
    const LaunchState *ls = nullptr;   // in `status`, after the k unit records
};

}  // namespace

struct cjx_server {
    int device = 0;
    size_t slot_budget = 0, max_batch = 0, in_use = 0;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_any, cv_express, cv_done;  // general / express slots, callers
    std::deque<Request> pending;
    std::deque<cjx_completion> done;
    std::vector<Region *> regions, free_regions;  // all regions / the unused ones
    uint64_t next_ticket = 1;
    bool stop = false;
    std::vector<std::thread> workers;  // one per slot
    uint64_t batches = 0, docs = 0, bytes = 0, max_inflight = 0, relaunches = 0;
    size_t pending_large = 0;  // pending requests of at least kExpressDocBytes
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
    if (hipHostMalloc(&r->p, cap, hipHostMallocCoherent) != hipSuccess) {
        delete r;
        return nullptr;
    }
    r->dp = device_view(r->p);
    if (!r->dp) {
        (void)hipHostFree(r->p);
        delete r;
        return nullptr;
    }
    r->cap = cap;
    srv->regions.push_back(r);
    return r;
}

// Where the calling slot thread last met a failed HIP call (source line and
// error), printed for failed batches when CJX_SERVER_TRACE is set.
thread_local int t_fail_line = 0;
thread_local hipError_t t_fail_err = hipSuccess;
int dev_fail(int line, hipError_t e = hipErrorUnknown) {
    t_fail_line = line;
    t_fail_err = e != hipErrorUnknown ? e : hipGetLastError();
    return CJX_ERR_DEVICE;
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
    const size_t status_bytes = k * sizeof(UnitOut) + sizeof(LaunchState);
    const bool zero_copy = off <= kZeroCopyBytes;
    uint64_t stage_end = 0;  // staged documents keep their device offsets
    for (size_t i = 0; i < k; i++)
        if (zero_copy || s.reqs[i].n < kStageSmall)
            stage_end = units[i].in_off + s.reqs[i].n;
    if (!ctx->in.ensure(off + 64) || !ctx->structural.ensure(s.sh.rows * 4 + 16) ||
        !ctx->pair.ensure(s.sh.rows * 4 + 16) ||
        !s.stage.ensure(stage_end + 64, hipHostMallocCoherent) ||
        !s.status.ensure(status_bytes + 64, hipHostMallocCoherent))
        return CJX_ERR_ALLOC;
    void *stage_dev = device_view(s.stage.p), *status_dev = device_view(s.status.p);
    if (!stage_dev || !status_dev) return dev_fail(__LINE__);
    // A small batch is gathered whole in the slot's pinned staging buffer
    // and pulled in by a copy kernel.  Otherwise small inputs are gathered and
    // copied in runs of consecutive staged documents, and large ones are
    // copied straight from the caller's memory.
    uint8_t *stage = static_cast<uint8_t *>(s.stage.p);
    uint8_t *d_in = ctx->in.as<uint8_t>();
    uint64_t lo = 0, hi = 0;  // pending staged run
    for (size_t i = 0; i <= k && zero_copy; i++) {
        if (i == k) {
            if (launch_copy(st, stage_dev, d_in, off, nullptr, nullptr, 0) != hipSuccess)
                return dev_fail(__LINE__);
            break;
        }
        std::memcpy(stage + units[i].in_off, s.reqs[i].data, s.reqs[i].n);
    }
    for (size_t i = 0; i <= k && !zero_copy; i++) {
        const bool staged = i < k && s.reqs[i].n < kStageSmall;
        if (!staged && hi > lo) {
            if (hipMemcpyAsync(d_in + lo, stage + lo, hi - lo, hipMemcpyHostToDevice, st) !=
                hipSuccess)
                return dev_fail(__LINE__);
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
            return dev_fail(__LINE__);
        }
    }
    if ((rc = upload_units(ctx, units, (uint32_t)tiles, n_small))) return rc;
    if ((rc = enqueue_parse(ctx, d_in, s.sh, s.flags, ctx->structural.as<uint32_t>(),
                            ctx->pair.as<int32_t>())))
        return rc == CJX_ERR_DEVICE ? dev_fail(__LINE__) : rc;
    uint8_t *sd = static_cast<uint8_t *>(status_dev);
    if (launch_copy(st, ctx->uout.p, sd, k * sizeof(UnitOut), ctx->ls.p,
                    sd + k * sizeof(UnitOut), sizeof(LaunchState)) != hipSuccess ||
        hipEventRecord(s.ev, st) != hipSuccess)
        return dev_fail(__LINE__);
    s.uo = static_cast<const UnitOut *>(s.status.p);
    s.ls = reinterpret_cast<const LaunchState *>(static_cast<uint8_t *>(s.status.p) +
                                                 k * sizeof(UnitOut));
    const double t = now_ms();
    for (Request &r : s.reqs) r.t_launch = t;
    s.state = Slot::kKernels;
    return CJX_OK;
}

// Kernels done: relaunch with exact regions on overflow, else lay the
// results out per request and copy them into a pinned region.
int after_kernels(cjx_server *srv, Slot &s) {
    if (s.ls->overflow) {
        s.full_regions = true;
        {
            std::lock_guard<std::mutex> g(srv->mu);
            srv->relaunches++;
        }
        return launch_batch(s);
    }
    s.N = s.ls->n_total;
    const size_t k = s.reqs.size();
    const size_t cells = 2 * s.N + 4 * k;
    {
        std::lock_guard<std::mutex> g(srv->mu);
        s.region = acquire_region(srv, cells * 4);
    }
    if (!s.region) return CJX_ERR_ALLOC;
    hipStream_t st = s.ctx->stream;
    const LaunchBufs b = bufs_of(s.ctx);
    const uint32_t *d_st = s.ctx->structural.as<uint32_t>();
    const int32_t *d_pr = s.ctx->pair.as<int32_t>();
    if (!s.ctx->layout.ensure(cells * 4 + 16)) return CJX_ERR_ALLOC;
    if (launch_assemble_docs(st, b, (uint32_t)k, d_st, d_pr, s.ctx->layout.as<int32_t>(),
                             (uint32_t)s.N) != hipSuccess)
        return dev_fail(__LINE__);
    // small results: a copy kernel into the region; large ones: the copy engine
    const hipError_t e =
        cells * 4 <= kZeroCopyBytes
            ? launch_copy(st, s.ctx->layout.p, s.region->dp, cells * 4, nullptr, nullptr, 0)
            : hipMemcpyAsync(s.region->p, s.ctx->layout.p, cells * 4, hipMemcpyDeviceToHost, st);
    if (e != hipSuccess || hipEventRecord(s.ev, st) != hipSuccess) return dev_fail(__LINE__);
    s.state = Slot::kCopyOut;
    return CJX_OK;
}

// The region holds every request's buffer: hand them out.
void finish_batch(cjx_server *srv, Slot &s) {
    const UnitOut *uo = s.uo;
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
    static std::atomic<int> traced{0};
    static const bool trace = std::getenv("CJX_SERVER_TRACE") != nullptr;
    if (trace && traced.fetch_add(1) < 20)
        std::fprintf(stderr, "[cjx server] batch of %zu documents failed: code %d, line %d, %s\n",
                     s.reqs.size(), rc, t_fail_line, hipGetErrorString(t_fail_err));
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

// Moves pending requests into the idle slot s: the oldest eligible one
// (express slots: below kExpressDocBytes), then the following eligible ones
// with the same flags, in arrival order, while they fit the slot's budget
// share, max_batch and the batch byte cap.  Requests passed over keep their
// places.  Returns false when none is eligible.  Called with the server lock
// held.
bool admit(cjx_server *srv, Slot &s) {
    std::deque<Request> &q = srv->pending;
    const size_t byte_cap = s.express ? kExpressBatchBytes : kBatchBytes;
    // general slots serve large requests first: nothing else can take them
    const bool large_only = !s.express && srv->pending_large > 0;
    auto eligible = [&](const Request &r) {
        return s.express ? r.n < kExpressDocBytes : (!large_only || r.n >= kExpressDocBytes);
    };
    size_t first = 0;
    while (first < q.size() && !eligible(q[first])) first++;
    if (first == q.size()) return false;
    const unsigned flags = q[first].flags;
    size_t cost = 0, bytes = 0;
    std::vector<char> taken;
    size_t i = first;
    for (; i < q.size(); i++) {
        const Request &r = q[i];
        const bool take = r.flags == flags && eligible(r);
        if (take) {
            // the first request always runs, alone if it exceeds the caps
            if (!s.reqs.empty() && (s.reqs.size() == srv->max_batch ||
                                    cost + r.cost > srv->slot_budget || bytes + r.n > byte_cap))
                break;
            cost += r.cost;
            bytes += r.n;
            s.reqs.push_back(r);
        }
        taken.push_back(take);
    }
    // drop the taken requests from q[first, i): the passed-over ones move up
    // to end at i, then the front is erased
    size_t w = i;
    for (size_t j = i; j-- > first;)
        if (!taken[j - first]) q[--w] = q[j];
    for (size_t j = first; j-- > 0;) q[--w] = q[j];
    q.erase(q.begin(), q.begin() + (ptrdiff_t)w);
    for (const Request &r : s.reqs) srv->pending_large -= r.n >= kExpressDocBytes;
    srv->in_use += cost;
    s.cost = cost;
    s.flags = flags;
    srv->batches++;
    srv->max_inflight = std::max<uint64_t>(srv->max_inflight, srv->in_use);
    for (const Request &r : s.reqs) srv->bytes += r.n;
    return true;
}

// Grows the buffers of slot s for its largest batches (a batch at the byte
// cap of max_batch equal documents, and one document of max_request bytes on
// general slots) and adds a result region for the largest of them, so that
// serving does not allocate (device frees synchronize the whole device).
int reserve_slot(cjx_server *srv, Slot &s, size_t max_request) {
    cjx_context *ctx = s.ctx;
    const size_t cap = s.express ? kExpressBatchBytes : std::max(kBatchBytes, max_request);
    const size_t batch = s.express ? kExpressBatchBytes : kBatchBytes;
    uint64_t max_cells = 0;
    for (int pass = 0; pass < 2; pass++) {
        const size_t k = pass == 0 ? 1 : std::max<size_t>(1, std::min(srv->max_batch, batch / 256));
        const size_t each = pass == 0 ? cap : batch / k;
        std::vector<UnitDesc> units(k, UnitDesc{});
        uint64_t off = 0;
        for (size_t i = 0; i < k; i++) {
            units[i].in_off = off;
            units[i].n = each;
            off += (each + 15) & ~15ull;
        }
        uint32_t n_small = 0;
        const uint64_t tiles = assign_tiles(units, n_small);
        set_regions(units, false);
        Shape sh{(uint32_t)tiles, (uint32_t)k, n_small, 0, 0, 0, 0, false};
        shape_rows(sh, units);
        int rc;
        if ((rc = ensure_work(ctx, sh))) return rc;
        const uint64_t cells = 2 * sh.rows + 4 * k;
        const size_t table = k * sizeof(UnitDesc) + ((size_t)tiles + n_small) * 4 + 16;
        bool ok = ctx->in.ensure(off + 64) && ctx->structural.ensure(sh.rows * 4 + 16) &&
                  ctx->pair.ensure(sh.rows * 4 + 16) && ctx->tmp.ensure(sh.rows * 4 + 16) &&
                  ctx->layout.ensure(cells * 4 + 16) && ctx->units.ensure(table) &&
                  s.stage.ensure(batch + 64, hipHostMallocCoherent) &&
                  s.status.ensure(k * sizeof(UnitOut) + sizeof(LaunchState) + 64,
                                  hipHostMallocCoherent);
        for (int u = 0; u < cjx_context::kUploadSlots; u++) ok = ok && ctx->up[u].ensure(table);
        if (!ok) return CJX_ERR_ALLOC;
        max_cells = std::max(max_cells, cells);
    }
    std::lock_guard<std::mutex> g(srv->mu);
    Region *r = acquire_region(srv, max_cells * 4);
    if (!r) return CJX_ERR_ALLOC;
    srv->free_regions.push_back(r);
    return CJX_OK;
}

// Waits for the slot's event (polling keeps the wake-up latency low).
hipError_t wait_event(Slot &s) {
    for (;;) {
        const hipError_t q = hipEventQuery(s.ev);
        if (q != hipErrorNotReady) return q;
        std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

// Replaces a slot context's stream by one created with a CU mask of every
// CU of the device: no CU is left out, and the runtime gives such a stream a
// hardware queue of its own, so the slot never waits behind another stream's
// copies on a shared queue.  Where that cannot be had, the slot keeps the
// context's ordinary stream.  A slot keeps all its work on that one stream
// (no second stream beside it), so it holds exactly one queue.
int own_queue_stream(int device, cjx_context *ctx) {
    ctx->fork_ok = false;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess || prop.multiProcessorCount <= 0)
        return CJX_OK;
    const uint32_t cus = (uint32_t)prop.multiProcessorCount;
    std::vector<uint32_t> mask((cus + 31) / 32, 0xFFFFFFFFu);
    if (cus % 32) mask.back() = (1u << (cus % 32)) - 1u;
    hipStream_t st = nullptr;
    if (hipExtStreamCreateWithCUMask(&st, (uint32_t)mask.size(), mask.data()) != hipSuccess) {
        (void)hipGetLastError();  // clear it: the ordinary stream stays
        return CJX_OK;
    }
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
    ctx->stream = st;
    return CJX_OK;
}

// One host thread per slot: admit a batch, launch it, wait, lay it out,
// hand the results out, repeat.  The threads share only the request queue,
// the region pool and the statistics (under the server lock).
void slot_main(cjx_server *srv, Slot *sp) {
    Slot &s = *sp;
    (void)hipSetDevice(srv->device);
    std::condition_variable &cv = s.express ? srv->cv_express : srv->cv_any;
    for (;;) {
        {
            std::unique_lock<std::mutex> g(srv->mu);
            bool got = false;
            while (!(got = admit(srv, s))) {
                // general slots leave once nothing is pending; express slots
                // once nothing they may take is pending
                if (srv->stop) break;
                cv.wait_for(g, std::chrono::milliseconds(2));
            }
            if (!got) break;
        }
        s.full_regions = false;
        int rc = launch_batch(s);
        while (!rc) {
            const hipError_t we = wait_event(s);
            if (we != hipSuccess) {
                rc = dev_fail(__LINE__, we);
                break;
            }
            if (s.state == Slot::kKernels) {
                rc = after_kernels(srv, s);  // lays out, or re-runs with exact regions
                continue;
            }
            finish_batch(srv, s);
            break;
        }
        if (rc) fail_batch(srv, s, rc);
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
    for (int i = 0; i < n_slots; i++) srv->slots[(size_t)i].express = i >= n_slots - n_slots / 2;
    for (Slot &s : srv->slots) {
        int rc = cjx_context_create(device, &s.ctx);
        if (!rc && hipEventCreateWithFlags(&s.ev, hipEventDisableTiming) != hipSuccess)
            rc = CJX_ERR_DEVICE;
        if (!rc) rc = own_queue_stream(device, s.ctx);
        if (rc) {
            for (Slot &t : srv->slots) {
                if (t.ev) (void)hipEventDestroy(t.ev);
                if (t.ctx) cjx_context_destroy(t.ctx);
            }
            delete srv;
            return rc;
        }
    }
    for (Slot &s : srv->slots) srv->workers.emplace_back(slot_main, srv, &s);
    *out = srv;
    return CJX_OK;
}

int cjx_server_reserve(cjx_server *srv, size_t max_request_bytes) {
    if (!srv || max_request_bytes > kMaxLaunchBytes) return CJX_ERR_INPUT;
    {
        std::lock_guard<std::mutex> g(srv->mu);
        if (srv->next_ticket != 1) return CJX_ERR_INPUT;  // only before the first request
    }
    if (hipSetDevice(srv->device) != hipSuccess) return CJX_ERR_DEVICE;
    for (Slot &s : srv->slots) {
        const int rc = reserve_slot(srv, s, max_request_bytes);
        if (rc) return rc;
    }
    return hipDeviceSynchronize() == hipSuccess ? CJX_OK : CJX_ERR_DEVICE;
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
        srv->pending_large += n >= kExpressDocBytes;
    }
    if (ticket) *ticket = r.ticket;
    srv->cv_any.notify_one();
    if (n < kExpressDocBytes) srv->cv_express.notify_one();
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
    srv->cv_any.notify_all();
    srv->cv_express.notify_all();
    for (std::thread &t : srv->workers) t.join();
    for (Slot &s : srv->slots) {
        s.stage.release();
        s.status.release();
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
