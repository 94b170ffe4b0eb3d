// Runs the host chunker (cujson_amd/csrc/cjx_chunk.cpp) on a file; built with
// AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_sanitizers.py,
// which compares the chunks with the oracle's chunk rules.
//
//   chunk_san <file> <mode> <param> [thread limit]
//
// Prints "<rc> <count>", then "<offset> <size>" per chunk.  A thread limit
// above 0 makes the threaded newline scan fail past that many threads
// (its single-thread fallback).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

#include "cujson_amd.h"

extern "C" void cjx_debug_fail_threads(int k);

int main(int argc, char **argv) {
    if (argc != 4 && argc != 5) return 2;
    std::ifstream f(argv[1], std::ios::binary | std::ios::ate);
    if (!f) return 3;
    const size_t n = (size_t)f.tellg();
    std::vector<uint8_t> text(n);  // exactly n bytes: a read past the end is caught
    f.seekg(0);
    f.read(reinterpret_cast<char *>(text.data()), (std::streamsize)n);
    if (argc == 5) cjx_debug_fail_threads(std::atoi(argv[4]));
    const int mode = std::atoi(argv[2]);
    const size_t param = std::strtoull(argv[3], nullptr, 10);
    size_t count = 0;
    int rc = cjx_chunk_lines(text.data(), n, mode, param, nullptr, nullptr, 0, &count);
    if (rc != CJX_OK && rc != CJX_ERR_CAPACITY) {
        std::printf("%d 0\n", rc);
        return 0;
    }
    std::vector<size_t> off(count), size(count);
    rc = cjx_chunk_lines(text.data(), n, mode, param, off.data(), size.data(), count, &count);
    std::printf("%d %zu\n", rc, count);
    for (size_t i = 0; i < count; i++) std::printf("%zu %zu\n", off[i], size[i]);
    // a short capacity stores what fits and reports the count
    if (count > 1) {
        std::vector<size_t> o2(count - 1), s2(count - 1);
        size_t c2 = 0;
        const int r2 = cjx_chunk_lines(text.data(), n, mode, param, o2.data(), s2.data(),
                                       count - 1, &c2);
        if (r2 != CJX_ERR_CAPACITY || c2 != count || o2[count - 2] != off[count - 2]) return 5;
    }
    return 0;
}
