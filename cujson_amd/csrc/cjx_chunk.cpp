// cjx_chunk.cpp - the host chunker: cjx_chunk_lines (the line split of
// loadJSONLines_chunkCount / _chunkSizeBytes, cujson/load_file.cu:38-254)
// over a threaded newline scan of host text.  Host code only, no HIP.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <system_error>
#include <thread>
#include <vector>

#include "../../include/cujson_amd.h"
#include "cjx_chunk.h"

namespace cjx_chunk {

int emit_chunks(const ChunkList &out, size_t *chunk_off, size_t *chunk_size, size_t cap,
                size_t *n_chunks) {
    *n_chunks = out.size();
    for (size_t i = 0; i < out.size() && i < cap; i++) {
        if (chunk_off) chunk_off[i] = (size_t)out[i].first;
        if (chunk_size) chunk_size[i] = (size_t)out[i].second;
    }
    return out.size() <= cap ? CJX_OK : CJX_ERR_CAPACITY;
}

namespace {

// Test hook: when positive, the threaded line scan behaves as if only this
// many of its threads could be created (cjx_debug_fail_threads).
std::atomic<int> g_thread_limit{0};

}  // namespace

// From 64 MiB the text is cut into segments of at least 16 MiB scanned by up
// to 16 threads, each collecting its own offsets; the lists are joined in
// order.  If a thread cannot be created, the threads already started are
// joined and the whole text is scanned on the calling thread.  Returns false
// when memory runs out (nothing escapes into the C ABI).
bool host_line_starts(const uint8_t *data, size_t n, std::vector<uint64_t> &offs) {
    auto scan = [data](size_t a, size_t b, std::vector<uint64_t> &v) {
        const uint8_t *p = data + a, *end = data + b;
        while (p < end) {
            const void *hit = std::memchr(p, '\n', (size_t)(end - p));
            if (!hit) break;
            p = static_cast<const uint8_t *>(hit) + 1;
            v.push_back((uint64_t)(p - data));
        }
    };
    const size_t hw = std::max<size_t>(1, std::thread::hardware_concurrency());
    const size_t T = n >= (64u << 20) ? std::min<size_t>({hw, 16, n / (16u << 20)}) : 1;
    try {
        if (T > 1) {
            std::vector<std::vector<uint64_t>> part(T);
            std::vector<char> ok(T, 0);
            std::vector<std::thread> th;
            th.reserve(T);
            bool started = true;
            try {
                for (size_t t = 0; t < T; t++) {
                    const int lim = g_thread_limit.load();
                    if (lim > 0 && t >= (size_t)lim)
                        throw std::system_error(std::make_error_code(std::errc::resource_unavailable_try_again));
                    th.emplace_back([&, t] {
                        try {
                            scan(n * t / T, n * (t + 1) / T, part[t]);
                            ok[t] = 1;
                        } catch (...) {
                        }
                    });
                }
            } catch (const std::system_error &) {
                started = false;
            }
            for (std::thread &x : th) x.join();
            bool all = started;
            for (size_t t = 0; t < T && all; t++) all = ok[t] != 0;
            if (all) {
                size_t total = offs.size();
                for (const auto &v : part) total += v.size();
                offs.reserve(total);
                for (const auto &v : part) offs.insert(offs.end(), v.begin(), v.end());
                return true;
            }
            if (started) return false;  // a thread ran out of memory
        }
        scan(0, n, offs);  // one thread (small texts, or threads unavailable)
        return true;
    } catch (const std::bad_alloc &) {
        return false;
    }
}

}  // namespace cjx_chunk

using namespace cjx_chunk;

extern "C" {

int cjx_chunk_lines(const uint8_t *data, size_t n, int mode, size_t param, size_t *chunk_off,
                    size_t *chunk_size, size_t cap, size_t *n_chunks) {
    if (!n_chunks) return CJX_ERR_INPUT;
    *n_chunks = 0;
    if (!data && n) return CJX_ERR_INPUT;
    if (!chunk_rule_ok(mode, param)) return CJX_ERR_INPUT;
    // Line starts (load_file.cu:72-81): 0, every byte after '\n', and n if the
    // last line is unterminated.
    try {
        std::vector<uint64_t> offs;
        offs.push_back(0);
        if (!host_line_starts(data, n, offs)) return CJX_ERR_ALLOC;
        if (offs.back() < n) offs.push_back(n);
        ChunkList out;
        if (mode == 0) {
            std::vector<uint64_t> li;
            count_rule(offs.size() - 1, param, li);
            for (size_t i = 0; i < param; i++)
                out.emplace_back(offs[li[i]], offs[li[i + 1]] - offs[li[i]]);
        } else {
            bytes_rule([&](size_t j) { return offs[j]; }, offs.size(), n, param, out);
        }
        return emit_chunks(out, chunk_off, chunk_size, cap, n_chunks);
    } catch (const std::bad_alloc &) {
        return CJX_ERR_ALLOC;
    }
}

// Test hook (not in the public header): the threaded line scan of
// cjx_chunk_lines acts as if only `k` threads could be created (k <= 0:
// no limit), so tests can force its single-thread fallback.
void cjx_debug_fail_threads(int k) { g_thread_limit.store(k); }

}  // extern "C"


