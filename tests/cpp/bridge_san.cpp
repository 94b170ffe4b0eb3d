// Drives include/cujson/cjx_bridge.h (the glue between the reference-shaped
// C++ API and the C ABI) with results in malloc'd memory: built with
// AddressSanitizer and UndefinedBehaviorSanitizer against the host-only
// stand-ins of tests/cpp/abi_host_stubs.cpp and the library's own chunker.
//
//   bridge_san <result file> <json lines file> <chunk count>
//
// The result file holds the 2(N+2) int32 cells of a standard-JSON parse.
// Prints one "ok <what>" line per check and exits non-zero at the first
// failing one.
#define CUJSON_AMD_NO_EXIT
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <vector>

#include "cujson/cujson.h"
#include "cujson/cujsonlines.h"

#define CHECK(cond, what)                                     \
    do {                                                      \
        if (!(cond)) {                                        \
            std::printf("fail %s (line %d)\n", what, __LINE__); \
            return 1;                                         \
        }                                                     \
        std::printf("ok %s\n", what);                         \
    } while (0)

// A cjx_result as the library fills it, with its buffers from malloc.
static cjx_result make_result(const std::vector<int32_t> &cells, int chunks) {
    const int64_t N = (int64_t)cells.size() / 2 - 2;
    cjx_result r{};
    r.buf = static_cast<int32_t *>(std::malloc(cells.size() * 4));
    std::memcpy(r.buf, cells.data(), cells.size() * 4);
    r.structural = r.buf;
    r.pair_pos = r.buf + N + 1;
    r.n_structural = N;
    r.chunk_count = chunks;
    r.result_sizes = static_cast<int32_t *>(std::malloc(4 * (size_t)chunks));
    r.result_sizes_prefix = static_cast<int32_t *>(std::malloc(4 * (size_t)chunks));
    int64_t acc = 0;
    for (int i = 0; i < chunks; i++) {
        const int64_t part = N / chunks + (i < N % chunks ? 1 : 0);
        acc += part;
        r.result_sizes[i] = (int32_t)part;
        r.result_sizes_prefix[i] = (int32_t)acc;
    }
    r.total_result_size = r.file_size = (int32_t)(N + 2);
    r.error_chunk = -1;
    return r;
}

int main(int argc, char **argv) {
    if (argc != 4) return 2;
    std::ifstream f(argv[1], std::ios::binary | std::ios::ate);
    if (!f) return 3;
    std::vector<int32_t> cells((size_t)f.tellg() / 4);
    f.seekg(0);
    f.read(reinterpret_cast<char *>(cells.data()), (std::streamsize)(cells.size() * 4));
    const int N = (int)cells.size() / 2 - 2;

    // success: the buffer changes hands, the size arrays are copied and freed
    for (int chunks : {1, 3}) {
        cjx_result r = make_result(cells, chunks);
        int32_t *buf = r.buf;
        cuJSONResult out = cujson_amd_detail::to_result(r, CJX_OK);
        CHECK(out.structural == buf && out.pair_pos == buf + N + 1, "to_result buffers");
        CHECK(out.totalResultSize == N + 2 && out.fileSize == N + 2, "to_result sizes");
        CHECK(out.chunkCount == chunks && (int)out.resultSizes.size() == chunks &&
                  out.resultSizesPrefix.back() == N,
              "to_result chunk sizes");
        CHECK(std::memcmp(out.structural, cells.data(), cells.size() * 4) == 0, "cells kept");
        CHECK(cujson_amd_last_error() == CJX_OK, "last error ok");
        cujson_free_result(out);
        CHECK(out.structural == nullptr && out.pair_pos == nullptr, "cujson_free_result");
    }
    // error codes: the result is released and the code kept
    for (int rc : {CJX_ERR_INPUT, CJX_ERR_UTF8, CJX_ERR_STRUCTURE, CJX_ERR_CAPACITY}) {
        cjx_result r = make_result(cells, 2);
        cuJSONResult out = cujson_amd_detail::to_result(r, rc);
        CHECK(out.structural == nullptr && out.chunkCount == 0 && out.resultSizes.empty(),
              "to_result error empty");
        CHECK(cujson_amd_last_error() == rc, "last error code");
        CHECK(r.buf == nullptr, "error result released");
    }
    // the parse wrappers without a device: empty results, inputs released
    cuJSONInput in = loadJSON(argv[2]);
    CHECK(in.data != nullptr && in.size > 0, "loadJSON");
    cuJSONResult none = parse_standard_json(in);
    CHECK(none.structural == nullptr, "parse_standard_json without a device");
    CHECK(parse_standard_json(cuJSONInput{nullptr, 0}).structural == nullptr, "empty input");
    CHECK(loadJSON("/nonexistent/file.json").data == nullptr, "missing file");

    // loadJSONLines_*: chunks cut by the library's chunker over the loaded text
    const size_t count = std::strtoull(argv[3], nullptr, 10);
    cuJSONLinesInput li = loadJSONLines_chunkCount(argv[2], count);
    CHECK(li.data != nullptr && li.chunkCount == count && li.chunks.size() == count,
          "loadJSONLines_chunkCount");
    size_t total = 0;
    for (size_t i = 0; i < li.chunkCount; i++) {
        if (li.chunks[i] != li.data + total) return 4;
        total += li.chunksSize[i];
    }
    CHECK(total == li.size, "chunks tile the text");
    for (size_t i = 0; i < li.chunkCount; i++) std::printf("chunk %zu %zu\n",
                                                          (size_t)(li.chunks[i] - li.data),
                                                          li.chunksSize[i]);
    cuJSONResult lr = parse_json_lines(li);
    CHECK(lr.structural == nullptr, "parse_json_lines without a device");
    cuJSONLinesInput lb = loadJSONLines_chunkSizeBytes(argv[2], 64 << 10);
    CHECK(lb.data != nullptr && lb.chunkCount >= 1, "loadJSONLines_chunkSizeBytes");
    size_t tb = 0;
    for (size_t i = 0; i < lb.chunkCount; i++) tb += lb.chunksSize[i];
    CHECK(tb == lb.size, "byte chunks tile the text");
    cjx_input lbin{lb.data, lb.size};
    cjx_free_input(&lbin);
    return 0;
}
