// cjx_utf8.hip - strict (RFC 3629) UTF-8 validation, the CJX_FLAG_UTF8_STRICT
// mode.  The reference only has its weaker per-word check (UTF8Validation,
// parse_standard_json.cu:264-327), which the tokenizer computes by default;
// strict mode replaces it with this kernel.
//
// Validity of a byte string equals these local conditions at every position
// p (with the three bytes before p, zero before the unit start, and three
// zero bytes after the end):
//   - p is a continuation byte (10xxxxxx) exactly when a lead expects one:
//     b[p-1] >= C0, or b[p-2] >= E0, or b[p-3] >= F0;
//   - b[p] is none of C0, C1, F5..FF;
//   - second-byte ranges: after E0 at least A0 (no overlong), after ED at most
//     9F (no surrogate), after F0 at least 90 (no overlong), after F4 at most
//     8F (at most U+10FFFF).
// The three positions past the end carry a zero byte, so a sequence cut by
// the end fails the first condition.
#include "cjx_device.h"

namespace cjx {

namespace {

constexpr int kU8Threads = 256;
constexpr uint32_t kU8ThreadBytes = kTileBytes / kU8Threads;  // 256 B per thread

__device__ __forceinline__ uint32_t bad_at(uint32_t x, uint32_t a, uint32_t b, uint32_t c) {
    const uint32_t is_cont = (x & 0xC0u) == 0x80u;
    const uint32_t must = (a >= 0xC0u) | (b >= 0xE0u) | (c >= 0xF0u);
    uint32_t bad = is_cont ^ must;
    bad |= (x == 0xC0u) | (x == 0xC1u) | (x >= 0xF5u);
    bad |= ((a == 0xE0u) & (x < 0xA0u)) | ((a == 0xEDu) & (x > 0x9Fu)) |
           ((a == 0xF0u) & (x < 0x90u)) | ((a == 0xF4u) & (x > 0x8Fu));
    return bad;
}

// Four bytes at o (a multiple of 4); bytes at or past n read as zero.  Whole
// words below n are loaded directly; the input is readable up to n rounded
// up to 16.
__device__ __forceinline__ uint32_t word_at(const uint8_t *base, uint64_t o, uint64_t n) {
    if (o >= n) return 0;
    const uint32_t v = *reinterpret_cast<const uint32_t *>(base + o);
    const uint64_t rem = n - o;
    return rem >= 4 ? v : (v & ((1u << (8 * rem)) - 1u));
}

}  // namespace

// One block per tokenizer tile; thread t checks bytes [256 t, 256 t + 256) of
// the tile, four at a time, skipping words that hold no byte >= 80 and follow
// a word without one.
// Blocks past the tiles take one small unit each (at most 16 KiB).
__global__ __launch_bounds__(kU8Threads) void k_utf8_strict(const uint8_t *__restrict__ in,
                                                            const UnitDesc *__restrict__ units,
                                                            const uint32_t *__restrict__ tile_unit,
                                                            const uint32_t *__restrict__ small_units,
                                                            uint32_t ntiles,
                                                            UnitOut *__restrict__ uo) {
    const uint32_t blk = blockIdx.x;
    const bool small = blk >= ntiles;
    const uint32_t u = small ? small_units[blk - ntiles] : tile_unit[blk];
    const UnitDesc U = units[u];
    const uint8_t *base = in + U.in_off;
    const uint64_t n = U.n;
    const uint64_t s = (small ? 0ull : (uint64_t)(blk - U.tile0) * kTileBytes) +
                       threadIdx.x * kU8ThreadBytes;
    if (s >= n) return;
    uint32_t prev = s >= 4 ? word_at(base, s - 4, n) : 0u;
    uint32_t bad = 0;
    const uint64_t end = s + kU8ThreadBytes < n ? s + kU8ThreadBytes : n;
    for (uint64_t o = s; o < end; o += 4) {
        const uint32_t cur = word_at(base, o, n);
        if (((cur | prev) & 0x80808080u) != 0) {
            const uint64_t win = ((uint64_t)cur << 32) | prev;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t x = (uint32_t)(win >> (32 + 8 * j)) & 0xFFu;
                const uint32_t a = (uint32_t)(win >> (24 + 8 * j)) & 0xFFu;
                const uint32_t b = (uint32_t)(win >> (16 + 8 * j)) & 0xFFu;
                const uint32_t c = (uint32_t)(win >> (8 + 8 * j)) & 0xFFu;
                if (o + j < n) bad |= bad_at(x, a, b, c);
            }
        }
        prev = cur;
    }
    if (end == n) {  // the three zero bytes past the end
        const uint32_t b1 = n >= 1 ? base[n - 1] : 0u, b2 = n >= 2 ? base[n - 2] : 0u,
                       b3 = n >= 3 ? base[n - 3] : 0u;
        bad |= bad_at(0, b1, b2, b3) | bad_at(0, 0, b1, b2) | bad_at(0, 0, 0, b1);
    }
    if (bad) atomicOr(&uo[u].utf8_err, 1u);
}

hipError_t launch_utf8_strict(hipStream_t s, const LaunchBufs &b, const uint8_t *d_in,
                              uint32_t ntiles, uint32_t n_small) {
    if (ntiles + n_small == 0) return hipSuccess;
    hipLaunchKernelGGL(k_utf8_strict, dim3(ntiles + n_small), dim3(kU8Threads), 0, s, d_in,
                       b.units, b.tile_unit, b.small_units, ntiles, b.uout);
    return hipGetLastError();
}

}  // namespace cjx
