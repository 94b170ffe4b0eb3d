// cujson/cjx_bridge.h - glue between the reference-shaped C++ API and the C
// ABI of include/cujson_amd.h: one context per host thread on device 0 (the
// reference uses the implicit device 0 of the CUDA runtime), conversion of a
// cjx_result into a cuJSONResult, and the reference's error behaviour.
#ifndef CUJSON_AMD_BRIDGE_H
#define CUJSON_AMD_BRIDGE_H

#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "../cujson_amd.h"
#include "cujson_types.h"

namespace cujson_amd_detail {

// Lazily created, per host thread; destroyed at thread exit.
struct ThreadContext {
    cjx_context *ctx = nullptr;
    int rc = CJX_OK;
    ThreadContext() { rc = cjx_context_create(0, &ctx); }
    ~ThreadContext() {
        if (ctx) cjx_context_destroy(ctx);
    }
};

inline cjx_context *context() {
    thread_local ThreadContext t;
    if (!t.ctx) std::cerr << "cujson_amd: no HIP device (" << cjx_strerror(t.rc) << ")\n";
    return t.ctx;
}

// The reference prints and calls exit(0) on invalid UTF-8 or brackets
// (parse_standard_json.cu:1626, :1567-1570; parse_json_lines.cu:1172, :1058).
// Define CUJSON_AMD_NO_EXIT to get an empty result instead; the code is then
// available from cujson_amd_last_error().
inline int &last_error() {
    thread_local int e = CJX_OK;
    return e;
}

inline cuJSONResult to_result(cjx_result &r, int rc) {
    last_error() = rc;
    cuJSONResult out{};
    if (rc == CJX_ERR_STRUCTURE || rc == CJX_ERR_UTF8) {
#ifndef CUJSON_AMD_NO_EXIT
        if (rc == CJX_ERR_STRUCTURE) std::printf("error found!");
        std::fflush(stdout);
        std::exit(0);
#else
        cjx_free_result(&r);
        return out;
#endif
    }
    if (rc != CJX_OK) {
        std::cerr << "\033[1;31m Error: " << cjx_strerror(rc) << ". \033[0m\n";
        cjx_free_result(&r);
        return out;
    }
    out.chunkCount = r.chunk_count;
    out.structural = r.structural;
    out.pair_pos = r.pair_pos;
    out.totalResultSize = r.total_result_size;
    out.fileSize = r.file_size;
    out.resultSizes.assign(r.result_sizes, r.result_sizes + r.chunk_count);
    out.resultSizesPrefix.assign(r.result_sizes_prefix, r.result_sizes_prefix + r.chunk_count);
    // The buffer now belongs to the cuJSONResult (release with cujson_free_result).
    std::free(r.result_sizes);
    std::free(r.result_sizes_prefix);
    return out;
}

inline cjx_result from_result(cuJSONResult &res) {
    cjx_result r{};
    r.buf = res.structural;
    r.error_chunk = -1;
    return r;
}

}  // namespace cujson_amd_detail

// Status of the last parse on this thread (CJX_* codes).
inline int cujson_amd_last_error() { return cujson_amd_detail::last_error(); }

// Releases a result's pinned buffer.  Replaces the cudaFreeHost(structural)
// of the reference's cuJSONIterator::freeJson (query_iterator_standard_json.cpp:143-146).
inline void cujson_free_result(cuJSONResult &res) {
    cjx_result r = cujson_amd_detail::from_result(res);
    cjx_free_result(&r);
    res.structural = res.pair_pos = nullptr;
}

// loadJSON (cujson/load_file.h:9, load_file.cu:7-36).
inline void load_error(int rc) {
    if (rc == CJX_ERR_INPUT) std::cerr << "Failed to open file\n";  // load_file.cu:10-13
    else std::cerr << "Failed to allocate pinned host memory (" << cjx_strerror(rc) << ")\n";
}

inline cuJSONInput loadJSON(const std::string &filePath) {
    cjx_input in{};
    const int rc = cjx_load(filePath.c_str(), &in);
    if (rc != CJX_OK) {
        load_error(rc);
        return cuJSONInput{nullptr, 0};
    }
    return cuJSONInput{in.data, in.size};
}

namespace cujson_amd_detail {

inline cuJSONLinesInput load_lines(const std::string &filePath, int mode, size_t param) {
    cuJSONLinesInput out{};
    cjx_input in{};
    const int rc = cjx_load(filePath.c_str(), &in);
    if (rc != CJX_OK) {
        load_error(rc);
        return out;
    }
    out.data = in.data;
    out.size = in.size;
    size_t n = 0;
    cjx_chunk_lines(in.data, in.size, mode, param, nullptr, nullptr, 0, &n);
    std::vector<size_t> off(n), sz(n);
    if (n && cjx_chunk_lines(in.data, in.size, mode, param, off.data(), sz.data(), n, &n) != CJX_OK)
        return out;
    out.chunkCount = n;
    for (size_t i = 0; i < n; i++) {
        out.chunks.push_back(in.data + off[i]);
        out.chunksSize.push_back(sz[i]);
    }
    return out;
}

}  // namespace cujson_amd_detail

// loadJSONLines_* (cujson/load_file.h:10-12, load_file.cu:38-254).
inline cuJSONLinesInput loadJSONLines_chunkCount(const std::string &filePath, size_t chunkCount) {
    return cujson_amd_detail::load_lines(filePath, 0, chunkCount);
}
inline cuJSONLinesInput loadJSONLines_chunkSizeBytes(const std::string &filePath,
                                                     size_t chunkSizeBytes) {
    return cujson_amd_detail::load_lines(filePath, 1, chunkSizeBytes);
}
inline cuJSONLinesInput loadJSONLines_chunkSizeMegaBytes(const std::string &filePath,
                                                         size_t chunkSizeMegaBytes) {
    return cujson_amd_detail::load_lines(filePath, 1, chunkSizeMegaBytes * 1024 * 1024);
}

#endif  // CUJSON_AMD_BRIDGE_H
