// cjx_loadgen.cpp - open-loop load driver for the request server (a bench
// tool, built as libcjx_loadgen.so and called in-process by bench.py).
//
// Workload: Poisson arrivals at a fixed rate for a fixed time; each request
// is a standard-JSON document drawn from a pool, its size class chosen from
// a mixture (default 80% 1-16 KiB, 19% 64 KiB-4 MiB, 1% 64-256 MiB, sizes
// uniform within a class).  Pool documents come from the synthetic
// generator (cjx_gen_record below 8 KiB, cjx_gen_standard above) and live in
// pinned host memory, so payloads start in host memory and every request
// crosses the host link both ways.
//
// Latency is measured end to end from the scheduled arrival time to the
// moment the collector thread holds the completion (so a late submission
// counts against the server, not for it).  Every completion is checked
// against the result of its pool document parsed alone by the plain path
// (cjx_parse_standard): error code and structural count always, a checksum
// of the whole buffer for small documents and of sampled cells for the rest.
// The plain path is then timed on the same request sequence, one document
// per call, for comparison.  Before serving, the host-link ceiling of this
// traffic (inputs in, results out, several streams, no parse) is measured
// with the copy engines: it bounds what any server can sustain here.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/cujson_amd.h"

extern "C" {
size_t cjx_gen_standard(uint64_t seed, uint64_t n_objects, uint64_t target_bytes, int deep_depth,
                        uint64_t chunk_bytes, int boundary_every, uint8_t *buf, size_t cap);
size_t cjx_gen_record(uint64_t seed, uint64_t target, int extra_depth, uint8_t *buf, size_t cap);
}

namespace {

using clk = std::chrono::steady_clock;

double ms_since(clk::time_point t0) {
    return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
}

struct Config {
    double rate = 20000, seconds = 60, plain_seconds = 10;
    size_t budget = 64ull << 30, max_batch = 4096;
    int slots = 8, pool_small = 4096, pool_mid = 128, pool_large = 8;
    uint64_t seed = 20240605;
    double p_small = 0.80, p_mid = 0.19;  // the rest is large
    size_t small_lo = 1 << 10, small_hi = 16 << 10, mid_lo = 64 << 10, mid_hi = 4 << 20,
           large_lo = 64 << 20, large_hi = 256 << 20;
};

struct Doc {
    uint8_t *p = nullptr;
    size_t n = 0;
    int cls = 0;
    // the plain path's result for this document alone
    int error = 0;
    int64_t n_structural = 0;
    uint64_t sum = 0;
};

// Checksum of the result cells: all of them for small results, else 3
// windows of 1024 cells (start, middle, end).
uint64_t cell_sum(const int32_t *buf, int64_t n_structural) {
    const uint64_t cells = 2 * (uint64_t)(n_structural + 2);
    auto mix = [](uint64_t h, uint64_t i, int32_t v) {
        return h + ((uint64_t)(uint32_t)v + 0x9E3779B97F4A7C15ull * (i + 1)) * 0xBF58476D1CE4E5B9ull;
    };
    uint64_t h = 0;
    if (cells <= 3 * 4096) {
        for (uint64_t i = 0; i < cells; i++) h = mix(h, i, buf[i]);
        return h;
    }
    const uint64_t starts[3] = {0, cells / 2 - 512, cells - 1024};
    for (uint64_t s : starts)
        for (uint64_t i = s; i < s + 1024; i++) h = mix(h, i, buf[i]);
    return h;
}

double pct(std::vector<double> &v, double q) {
    if (v.empty()) return 0;
    const size_t k = std::min(v.size() - 1, (size_t)std::floor(q * (double)(v.size() - 1) + 0.5));
    std::nth_element(v.begin(), v.begin() + (ptrdiff_t)k, v.end());
    return v[k];
}

bool parse_args(const char *s, Config &c) {
    // "key=value key=value ..."
    std::string a = s ? s : "";
    size_t i = 0;
    while (i < a.size()) {
        while (i < a.size() && a[i] == ' ') i++;
        size_t j = a.find(' ', i);
        if (j == std::string::npos) j = a.size();
        const std::string kv = a.substr(i, j - i);
        i = j;
        if (kv.empty()) continue;
        const size_t e = kv.find('=');
        if (e == std::string::npos) return false;
        const std::string k = kv.substr(0, e);
        const double v = std::atof(kv.c_str() + e + 1);
        if (k == "rate") c.rate = v;
        else if (k == "seconds") c.seconds = v;
        else if (k == "plain_seconds") c.plain_seconds = v;
        else if (k == "budget") c.budget = (size_t)v;
        else if (k == "max_batch") c.max_batch = (size_t)v;
        else if (k == "slots") c.slots = (int)v;
        else if (k == "pool_small") c.pool_small = (int)v;
        else if (k == "pool_mid") c.pool_mid = (int)v;
        else if (k == "pool_large") c.pool_large = (int)v;
        else if (k == "seed") c.seed = (uint64_t)v;
        else if (k == "p_small") c.p_small = v;
        else if (k == "p_mid") c.p_mid = v;
        else if (k == "large_lo") c.large_lo = (size_t)v;
        else if (k == "large_hi") c.large_hi = (size_t)v;
        else return false;
    }
    return c.rate > 0 && c.seconds > 0 && c.slots > 0 && c.pool_small > 0 && c.pool_mid > 0 &&
           c.pool_large >= 0;
}

// Generates the pool on several threads into one pinned allocation.
bool make_pool(const Config &c, std::vector<Doc> &pool, std::vector<int> *by_class,
               void **pinned) {
    std::mt19937_64 rng(c.seed);
    struct Spec { size_t target; int cls; uint64_t seed; };
    std::vector<Spec> specs;
    auto add = [&](int count, size_t lo, size_t hi, int cls) {
        std::uniform_int_distribution<size_t> d(lo, hi);
        for (int i = 0; i < count; i++) specs.push_back({d(rng), cls, rng()});
    };
    add(c.pool_small, c.small_lo, c.small_hi, 0);
    add(c.pool_mid, c.mid_lo, c.mid_hi, 1);
    add(c.pool_large, c.large_lo, c.large_hi, 2);
    // room for the generators' overshoot: records may exceed their target,
    // standard documents stop within one object of it
    auto cap_of = [](const Spec &s) {
        const size_t c = s.target < 8192 ? 2 * s.target + 8192 : s.target + (64 << 10);
        return (c + 255) & ~(size_t)255;
    };
    std::vector<size_t> off(specs.size());
    size_t total = 0;
    for (size_t i = 0; i < specs.size(); i++) {
        off[i] = total;
        total += cap_of(specs[i]);
    }
    uint8_t *base = nullptr;
    if (hipHostMalloc((void **)&base, total, hipHostMallocDefault) != hipSuccess) return false;
    *pinned = base;
    pool.assign(specs.size(), Doc{});
    std::atomic<size_t> next{0};
    auto work = [&]() {
        for (size_t i; (i = next.fetch_add(1)) < specs.size();) {
            const Spec &s = specs[i];
            uint8_t *dst = base + off[i];
            const size_t cap = cap_of(s);
            size_t n;
            if (s.target < 8192) {
                n = cjx_gen_record(s.seed, s.target, (int)(s.seed % 6), dst, cap);
                if (n) n -= 1;  // drop the record's '\n'
            } else {
                n = cjx_gen_standard(s.seed, 0, s.target, 0, 0, 0, dst, cap);
            }
            pool[i].p = dst;
            pool[i].n = n;
            pool[i].cls = s.cls;
        }
    };
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++) th.emplace_back(work);
    for (auto &t : th) t.join();
    for (size_t i = 0; i < pool.size(); i++) {
        if (pool[i].n == 0) return false;
        by_class[pool[i].cls].push_back((int)i);
    }
    return true;
}

// Host-link ceiling of the serving traffic: 4 streams, each repeatedly
// copying a 128 MiB input in and `out_ratio` of it out through the copy
// engines (no parse), as the server does for one request.  Returns input
// GB/s, or 0 on failure.
double link_ceiling(const uint8_t *src, size_t src_bytes, double out_ratio) {
    const int S = 4, reps = 6;
    const size_t in_b = std::min<size_t>(src_bytes, 128u << 20) & ~(size_t)15;
    const size_t out_b = (size_t)((double)in_b * out_ratio) & ~(size_t)15;
    std::vector<void *> din(S, nullptr), dout(S, nullptr), hout(S, nullptr);
    std::vector<hipStream_t> st(S, nullptr);
    double gbs = 0;
    bool ok = in_b > 0 && out_b > 0;
    for (int i = 0; i < S && ok; i++)
        ok = hipMalloc(&din[i], in_b) == hipSuccess && hipMalloc(&dout[i], out_b) == hipSuccess &&
             hipHostMalloc(&hout[i], out_b, hipHostMallocCoherent) == hipSuccess &&
             hipStreamCreate(&st[i]) == hipSuccess;
    if (ok) {
        auto issue = [&](int i) {
            (void)hipMemcpyAsync(din[i], src, in_b, hipMemcpyHostToDevice, st[i]);
            (void)hipMemcpyAsync(hout[i], dout[i], out_b, hipMemcpyDeviceToHost, st[i]);
        };
        for (int i = 0; i < S; i++) issue(i);
        ok = hipDeviceSynchronize() == hipSuccess;
        const auto t0 = clk::now();
        for (int r = 0; r < reps; r++)
            for (int i = 0; i < S; i++) issue(i);
        ok = ok && hipDeviceSynchronize() == hipSuccess;
        const double ms = ms_since(t0);
        if (ok && ms > 0) gbs = (double)reps * S * (double)in_b / (ms * 1e6);
    }
    for (int i = 0; i < S; i++) {
        if (st[i]) (void)hipStreamDestroy(st[i]);
        if (din[i]) (void)hipFree(din[i]);
        if (dout[i]) (void)hipFree(dout[i]);
        if (hout[i]) (void)hipHostFree(hout[i]);
    }
    return gbs;
}

}  // namespace

extern "C" {

// Runs the serving workload; writes one JSON object into out.  Returns 0 on
// success.  args: "key=value ..." (rate, seconds, plain_seconds, budget,
// max_batch, slots, pool_small, pool_mid, pool_large, seed, p_small, p_mid,
// large_lo, large_hi).
int cjx_loadgen_run(const char *args, char *out, size_t cap) {
    Config c;
    if (!out || cap < 64) return 1;
    out[0] = 0;
    if (!parse_args(args, c)) {
        std::snprintf(out, cap, "{\"error\": \"bad arguments\"}");
        return 1;
    }
    if (c.pool_large == 0) c.p_mid = 1.0 - c.p_small;
    std::vector<Doc> pool;
    std::vector<int> by_class[3];
    void *pinned = nullptr;
    const auto t_gen = clk::now();
    if (!make_pool(c, pool, by_class, &pinned)) {
        if (pinned) (void)hipHostFree(pinned);
        std::snprintf(out, cap, "{\"error\": \"pool generation failed\"}");
        return 1;
    }
    const double gen_ms = ms_since(t_gen);
    int rc = 0;
    // Reference results: every pool document parsed alone by the plain path.
    cjx_context *ctx = nullptr;
    if ((rc = cjx_context_create(0, &ctx))) {
        (void)hipHostFree(pinned);
        std::snprintf(out, cap, "{\"error\": \"context: %d\"}", rc);
        return 1;
    }
    for (Doc &d : pool) {
        cjx_result r;
        std::memset(&r, 0, sizeof(r));
        const int e = cjx_parse_standard(ctx, d.p, d.n, 0, &r);
        d.error = e;
        d.n_structural = r.buf ? r.n_structural : 0;
        d.sum = r.buf ? cell_sum(r.buf, r.n_structural) : 0;
        cjx_free_result(&r);
    }
    // The server's slot streams should have the hardware queues to
    // themselves: no other context lives during the timed run.
    cjx_context_destroy(ctx);
    ctx = nullptr;
    // Arrival schedule.
    std::mt19937_64 rng(c.seed ^ 0x5DEECE66Dull);
    std::exponential_distribution<double> gap(c.rate / 1000.0);  // in ms
    std::uniform_real_distribution<double> u01(0.0, 1.0);
    std::vector<double> t_arr;
    std::vector<int> doc_of;
    for (double t = gap(rng); t < c.seconds * 1000.0; t += gap(rng)) {
        const double x = u01(rng);
        const int cls = x < c.p_small ? 0 : (x < c.p_small + c.p_mid || by_class[2].empty() ? 1 : 2);
        const std::vector<int> &v = by_class[cls];
        t_arr.push_back(t);
        doc_of.push_back(v[(size_t)(u01(rng) * (double)v.size()) % v.size()]);
    }
    const size_t n_req = t_arr.size();
    // result bytes per input byte over the schedule, and the link ceiling
    double in_sum = 0, out_sum = 0;
    for (size_t i = 0; i < n_req; i++) {
        const Doc &d = pool[(size_t)doc_of[i]];
        in_sum += (double)d.n;
        out_sum += 8.0 * (double)(d.n_structural + 2);
    }
    const double out_ratio = in_sum > 0 ? out_sum / in_sum : 0.0;
    const Doc *biggest = &pool[0];
    for (const Doc &d : pool)
        if (d.n > biggest->n) biggest = &d;
    const double ceiling = link_ceiling(biggest->p, biggest->n, out_ratio);

    cjx_server *srv = nullptr;
    if ((rc = cjx_server_create(0, c.budget, c.slots, c.max_batch, &srv))) {
        cjx_context_destroy(ctx);
        (void)hipHostFree(pinned);
        std::snprintf(out, cap, "{\"error\": \"server: %d\"}", rc);
        return 1;
    }
    // Resident buffers for the largest request, then a warm-up with every
    // pool document once.
    if ((rc = cjx_server_reserve(srv, c.large_hi))) {
        cjx_server_destroy(srv);
        cjx_context_destroy(ctx);
        (void)hipHostFree(pinned);
        std::snprintf(out, cap, "{\"error\": \"reserve: %d\"}", rc);
        return 1;
    }
    {
        for (const Doc &d : pool) cjx_server_submit(srv, d.p, d.n, 0, nullptr);
        size_t got = 0;
        std::vector<cjx_completion> cb(256);
        while (got < pool.size()) {
            size_t k = 0;
            cjx_server_wait(srv, cb.data(), cb.size(), &k, 1000);
            for (size_t i = 0; i < k; i++) cjx_server_release(srv, &cb[i]);
            got += k;
        }
    }
    cjx_server_stats st0;
    cjx_server_get_stats(srv, &st0);

    // Open loop: one submitting thread on the schedule, one collector.
    std::vector<double> lat(n_req, -1.0), srv_lat(n_req, -1.0), queue(n_req, -1.0);
    std::mutex tmu;
    std::unordered_map<uint64_t, size_t> req_of;
    req_of.reserve(n_req * 2);
    std::atomic<size_t> mismatches{0}, submitted{0};
    // mismatches by kind: another error code (and the last such code), another
    // structural count, another cell sum
    std::atomic<size_t> mis_code{0}, mis_count{0}, mis_sum{0};
    std::atomic<int> last_code{0};
    std::atomic<bool> submit_failed{false};
    const auto t0 = clk::now();
    double last_done = 0;
    std::thread submitter([&]() {
        for (size_t i = 0; i < n_req; i++) {
            const auto due = t0 + std::chrono::duration_cast<clk::duration>(
                                      std::chrono::duration<double, std::milli>(t_arr[i]));
            std::this_thread::sleep_until(due);
            const Doc &d = pool[(size_t)doc_of[i]];
            uint64_t t = 0;
            {
                std::lock_guard<std::mutex> g(tmu);  // the collector looks tickets up
                if (cjx_server_submit(srv, d.p, d.n, 0, &t) != CJX_OK) {
                    submit_failed = true;
                    break;
                }
                req_of[t] = i;
            }
            submitted++;
        }
    });
    size_t done = 0;
    std::vector<cjx_completion> cb(1024);
    const double hard_stop_ms = c.seconds * 1000.0 + 30000.0;  // drain limit
    double next_report = 10000.0;
    while (done < n_req && !submit_failed) {
        size_t k = 0;
        cjx_server_wait(srv, cb.data(), cb.size(), &k, 50);
        const double now = ms_since(t0);
        for (size_t j = 0; j < k; j++) {
            cjx_completion &cc = cb[j];
            size_t i;
            {
                std::lock_guard<std::mutex> g(tmu);
                i = req_of[cc.ticket];
            }
            const Doc &d = pool[(size_t)doc_of[i]];
            bool ok = cc.error == d.error;
            if (!ok) {
                mis_code++;
                last_code = cc.error;
            } else if (d.error == CJX_OK) {
                if (cc.n_structural != d.n_structural || !cc.buf) {
                    ok = false;
                    mis_count++;
                } else if (cell_sum(cc.buf, cc.n_structural) != d.sum) {
                    ok = false;
                    mis_sum++;
                }
            }
            if (!ok) mismatches++;
            lat[i] = now - t_arr[i];
            srv_lat[i] = cc.latency_ms;
            queue[i] = cc.queue_ms;
            cjx_server_release(srv, &cc);
        }
        done += k;
        if (k) last_done = now;
        if (now >= next_report) {  // progress line every 10 s
            std::fprintf(stderr, "[serve] %.0f s: submitted %zu, completed %zu\n", now / 1000.0,
                         submitted.load(), done);
            next_report += 10000.0;
        }
        if (now > hard_stop_ms) break;
    }
    submitter.join();
    cjx_server_stats st1;
    cjx_server_get_stats(srv, &st1);
    cjx_server_destroy(srv);

    // Statistics of the timed requests.
    double bytes = 0;
    std::vector<double> all, by[3], sl, qs;
    for (size_t i = 0; i < n_req; i++) {
        if (lat[i] < 0) continue;
        const Doc &d = pool[(size_t)doc_of[i]];
        bytes += (double)d.n;
        all.push_back(lat[i]);
        by[d.cls].push_back(lat[i]);
        sl.push_back(srv_lat[i]);
        qs.push_back(queue[i]);
    }
    const double span_ms = last_done;
    const double offered_bytes = [&] {
        double b = 0;
        for (size_t i = 0; i < n_req; i++) b += (double)pool[(size_t)doc_of[i]].n;
        return b;
    }();

    // Plain path on the same sequence, one document per synchronous call.
    size_t plain_n = 0;
    double plain_bytes = 0;
    std::vector<double> plain_lat;
    if (cjx_context_create(0, &ctx) != CJX_OK) ctx = nullptr;
    const auto tp = clk::now();
    for (size_t i = 0; i < n_req && ms_since(tp) < c.plain_seconds * 1000.0; i++) {
        const Doc &d = pool[(size_t)doc_of[i]];
        const auto ts = clk::now();
        cjx_result r;
        std::memset(&r, 0, sizeof(r));
        if (!ctx) break;
        cjx_parse_standard(ctx, d.p, d.n, 0, &r);
        cjx_free_result(&r);
        plain_lat.push_back(ms_since(ts));
        plain_bytes += (double)d.n;
        plain_n++;
    }
    const double plain_ms = ms_since(tp);
    cjx_context_destroy(ctx);
    (void)hipHostFree(pinned);

    const size_t n_done = all.size();
    std::snprintf(
        out, cap,
        "{\"requests\": %zu, \"completed\": %zu, \"mismatches\": %zu, "
        "\"mismatch_kinds\": {\"error_code\": %zu, \"last_error_code\": %d, "
        "\"structural_count\": %zu, \"cell_sum\": %zu}, \"submit_failed\": %s, "
        "\"rate_req_s\": %.1f, \"seconds\": %.1f, \"offered_GBps\": %.3f, "
        "\"sustained_GBps\": %.3f, \"completed_req_s\": %.1f, "
        "\"p50_ms\": %.4f, \"p99_ms\": %.4f, \"p999_ms\": %.4f, \"max_ms\": %.4f, "
        "\"p50_small_ms\": %.4f, \"p99_small_ms\": %.4f, \"p50_mid_ms\": %.4f, "
        "\"p99_mid_ms\": %.4f, \"p50_large_ms\": %.4f, \"p99_large_ms\": %.4f, "
        "\"server_p50_ms\": %.4f, \"server_p99_ms\": %.4f, \"queue_p50_ms\": %.4f, "
        "\"queue_p99_ms\": %.4f, \"batches\": %llu, \"docs_per_batch\": %.2f, "
        "\"relaunches\": %llu, \"max_inflight_GB\": %.3f, \"region_GB\": %.3f, "
        "\"plain_requests\": %zu, \"plain_GBps\": %.3f, \"plain_req_s\": %.1f, "
        "\"plain_p50_ms\": %.4f, \"plain_p99_ms\": %.4f, \"pool_docs\": %zu, "
        "\"pool_generate_s\": %.2f, \"budget_GB\": %.1f, \"slots\": %d, \"max_batch\": %zu, "
        "\"result_bytes_per_input_byte\": %.4f, \"link_ceiling_GBps\": %.3f}",
        n_req, n_done, mismatches.load(), mis_code.load(), last_code.load(), mis_count.load(),
        mis_sum.load(), submit_failed ? "true" : "false", c.rate, c.seconds,
        offered_bytes / (c.seconds * 1e9), span_ms > 0 ? bytes / (span_ms * 1e6) : 0.0,
        span_ms > 0 ? (double)n_done / (span_ms / 1000.0) : 0.0, pct(all, 0.5), pct(all, 0.99),
        pct(all, 0.999), all.empty() ? 0.0 : *std::max_element(all.begin(), all.end()),
        pct(by[0], 0.5), pct(by[0], 0.99), pct(by[1], 0.5), pct(by[1], 0.99), pct(by[2], 0.5),
        pct(by[2], 0.99), pct(sl, 0.5), pct(sl, 0.99), pct(qs, 0.5), pct(qs, 0.99),
        (unsigned long long)(st1.batches - st0.batches),
        st1.batches > st0.batches ? (double)(st1.documents - st0.documents) /
                                        (double)(st1.batches - st0.batches)
                                  : 0.0,
        (unsigned long long)(st1.relaunches - st0.relaunches), st1.max_inflight_bytes / 1e9,
        st1.region_bytes / 1e9, plain_n, plain_ms > 0 ? plain_bytes / (plain_ms * 1e6) : 0.0,
        plain_ms > 0 ? (double)plain_n / (plain_ms / 1000.0) : 0.0, pct(plain_lat, 0.5),
        pct(plain_lat, 0.99), pool.size(), gen_ms / 1000.0, c.budget / 1e9, c.slots, c.max_batch,
        out_ratio, ceiling);
    return (mismatches.load() || n_done != n_req || submit_failed) ? 2 : 0;
}

}  // extern "C"
