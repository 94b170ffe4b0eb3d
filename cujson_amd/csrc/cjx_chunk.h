// cjx_chunk.h - the loader's line chunking rules (cujson/load_file.cu) over
// line starts, shared by the host chunker (cjx_chunk.cpp) and the device
// chunker (cjx_chunk_lines_device in cjx_api.cpp).  Host code only, no HIP:
// it also builds with a plain C++ compiler (the sanitizer test does).
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <utility>
#include <vector>

namespace cjx_chunk {

using ChunkList = std::vector<std::pair<uint64_t, uint64_t>>;

// Chunk boundaries under the count rule (load_file.cu:83-103), as line
// indices li[0..count]: chunk i takes lines / count lines, one more while
// i < lines % count.
inline void count_rule(uint64_t lines, uint64_t count, std::vector<uint64_t> &li) {
    li.assign(count + 1, 0);
    const uint64_t per = lines / count, extra = lines % count;
    for (uint64_t i = 0; i < count; i++) li[i + 1] = li[i] + per + (i < extra ? 1 : 0);
}

// Greedy split under the size rule (load_file.cu:153-176) over the line
// starts at(0..len) (at(0) = 0, ascending, the last one n): a chunk grows
// line by line; the first line start more than `param` bytes past the chunk
// start closes the chunk at the line start before it, which starts the next
// chunk, and the line start that closed it is not tested again.  A chunk may
// be empty (a line longer than param).  Binary searches give the same cuts
// as the reference's line-by-line walk.
template <class LineStart>
void bytes_rule(LineStart at, size_t len, uint64_t n, uint64_t param, ChunkList &out) {
    uint64_t cstart = 0;
    size_t i = 1;
    while (i < len) {
        const uint64_t lim = param > UINT64_MAX - cstart ? UINT64_MAX : cstart + param;
        size_t lo = i, hi = len;  // first j in [i, len) with at(j) > lim, else len
        while (lo < hi) {
            const size_t mid = lo + (hi - lo) / 2;
            if (at(mid) > lim) hi = mid; else lo = mid + 1;
        }
        const size_t j = lo;
        if (j >= len) break;
        out.emplace_back(cstart, at(j - 1) - cstart);
        cstart = at(j - 1);
        i = j + 1;
    }
    if (cstart < n) out.emplace_back(cstart, n - cstart);
}

// Mode 0 with 1 .. 2^32 chunks, or mode 1.
inline bool chunk_rule_ok(int mode, size_t param) {
    return (mode == 0 && param != 0 && param <= (size_t(1) << 32)) || mode == 1;
}

// Chunk list into the caller's arrays (up to cap entries); the count is
// always stored.  CJX_ERR_CAPACITY when cap is short.
int emit_chunks(const ChunkList &out, size_t *chunk_off, size_t *chunk_size, size_t cap,
                size_t *n_chunks);

// Appends the offset after every '\n' of data[0..n) to offs, in order
// (threaded from 64 MiB).  False when memory runs out.
bool host_line_starts(const uint8_t *data, size_t n, std::vector<uint64_t> &offs);

}  // namespace cjx_chunk
