// cjx_values.hip - typed value table of a parse.
//
// The reference returns values as raw substrings: its iterator's getValue
// cuts the text between two structural positions
// (cujson/query/query_iterator_standard_json.cpp:475-567).  Here every gap
// between consecutive structural characters is decoded once on the device:
// gap h (h = 0 .. N) is the text after structural h (host layout; 0 = before
// the first) up to structural h + 1 (the end of the text after the last).
// Trimmed of whitespace it is empty, a key (a string before ':'), a string,
// a number, true / false / null, or invalid (cjx_values.h decodes).
//   k_values        one thread per gap; numbers whose rounding needs the
//                   exact comparison are listed (gap index, atomic append)
//   k_values_exact  those numbers, grid-stride over the list
// String bodies are unescaped into `strings` at the offset of their raw
// body, which is never shorter, so no scan is needed: strings[text_off ..
// text_off + length) holds the UTF-8 of a key or string.  A block of
// k_values first copies its text span into LDS (coalesced 16-byte loads)
// when it fits.
#include "cjx_device.h"
#include "cjx_values.h"

namespace cjx {

namespace {

struct GapText {
    uint32_t a, b;   // trimmed text [a, b)
    uint8_t next;    // the structural character after the gap (0 after the last)
};

// Byte x of the text is src[x - off]: src is the text itself (off = 0) or
// the block's LDS copy of the span that starts at text byte off.
__device__ __forceinline__ GapText gap_of(const uint8_t *src, uint64_t off, uint64_t n,
                                          const int32_t *__restrict__ st, uint64_t N,
                                          uint64_t h) {
    // st[i] is the 1-based position of structural i + 1 (host index); rows
    // that point outside the text are clamped to it
    uint64_t a = h == 0 ? 0 : (uint64_t)(uint32_t)st[h - 1];
    uint64_t b = h == N ? n : (uint64_t)(uint32_t)st[h] - 1u;
    if (b > n) b = n;
    if (a > b) a = b;
    GapText g;
    g.next = (h < N && b < n) ? src[b - off] : 0;
    while (a < b && is_ws(src[a - off])) a++;
    while (b > a && is_ws(src[b - 1 - off])) b--;
    g.a = (uint32_t)a;
    g.b = (uint32_t)b;
    return g;
}

__device__ __forceinline__ bool literal(const uint8_t *s, uint32_t len, const char *lit,
                                        uint32_t k) {
    if (len != k) return false;
    for (uint32_t i = 0; i < k; i++)
        if (s[i] != (uint8_t)lit[i]) return false;
    return true;
}

}  // namespace

// Text bytes a block of k_values stages in LDS: the span from its first
// gap's start to the structural character after its last gap, when that
// fits and the text is 16-byte aligned (else the block reads the text in
// place).  The decoders then walk LDS instead of waiting on a global load
// per byte.
#ifndef CJX_VAL_STAGE
#define CJX_VAL_STAGE 16384
#endif
constexpr uint32_t kValStage = CJX_VAL_STAGE;

__global__ __launch_bounds__(256) void k_values(const uint8_t *__restrict__ in, uint64_t n,
                                                const int32_t *__restrict__ st, uint64_t N,
                                                ValueOut *__restrict__ out,
                                                uint8_t *__restrict__ strings,
                                                uint32_t *__restrict__ exact_list,
                                                uint32_t *__restrict__ exact_count) {
    const uint64_t h0 = (uint64_t)blockIdx.x * 256u;
    const uint64_t h = h0 + threadIdx.x;
    const uint8_t *src = in;
    uint64_t off = 0;
#if CJX_VAL_STAGE > 0
    __shared__ uint4 stage[kValStage / 16];
    {
        const uint64_t hl = h0 + 255 < N ? h0 + 255 : N;
        const uint64_t A = h0 == 0 ? 0 : (uint64_t)(uint32_t)st[h0 - 1];
        uint64_t B = hl == N ? n : (uint64_t)(uint32_t)st[hl];  // through the next structural
        if (B > n) B = n;
        const uint64_t a0 = (A < B ? A : B) & ~15ull;
        // Every row the block's gaps use (h0 - 1 .. h0 + 255), clamped to the
        // text as gap_of clamps it, must lie in [a0, B]: rows that are not
        // ascending (caller-built rows, gapped regions) read the text in place.
        bool outside = false;
        if (h0 + threadIdx.x >= 1 && h0 + threadIdx.x - 1 < N) {
            uint64_t v = (uint64_t)(uint32_t)st[h0 + threadIdx.x - 1];
            if (v > n) v = n;
            outside = v < a0 || v > B;
        }
        const bool any_outside = __syncthreads_or(outside) != 0;
        if (!any_outside && ((uintptr_t)in & 15u) == 0 && B - a0 <= kValStage) {  // uniform
            const uint32_t words = (uint32_t)((B - a0 + 15) / 16);
            for (uint32_t w = threadIdx.x; w < words; w += 256) {
                const uint64_t p = a0 + 16ull * w;
                if (p + 16 <= n) {
                    stage[w] = *reinterpret_cast<const uint4 *>(in + p);
                } else {  // the text's last partial word
                    uint8_t *d = reinterpret_cast<uint8_t *>(&stage[w]);
                    for (uint32_t k = 0; k < 16; k++) d[k] = p + k < n ? in[p + k] : 0;
                }
            }
            __syncthreads();
            src = reinterpret_cast<const uint8_t *>(stage);
            off = a0;
        }
    }
#endif
    // Classify every gap; empty gaps and literals are done at once.  Strings
    // (in four lists by length) and numbers go to block lists, then threads
    // 0.. of the block take one list entry each: a wave runs one decoder for
    // all its lanes instead of every decoder for some of them, and strings of
    // similar length share waves (a wave takes as long as its longest one).
    __shared__ uint32_t s_a[256], s_b[256];
    __shared__ uint8_t s_next[256];
    // lists 0-3: strings by length (under 16, 32, 48 bytes, longer); 4: numbers
    constexpr int kStrLists = 4, kNumList = 4;
    __shared__ uint8_t s_list[kStrLists + 1][256];
    __shared__ uint32_t s_cnt[kStrLists + 1];
    if (threadIdx.x < kStrLists + 1) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    if (h <= N) {
        const GapText g = gap_of(src, off, n, st, N, h);
        const uint8_t *s = src + (g.a - off);
        const uint32_t len = g.b - g.a;
        int list = -1;
        ValueOut v;
        v.v = 0;
        v.row = (uint32_t)h;
        v.text_off = g.a;
        v.text_len = len;
        v.type = kValNone;
        v.flags = 0;
        v.reserved = 0;
        if (len == 0) {
            // nothing between the two structural characters
        } else if (s[0] == '"') {
            list = len < 48 ? (int)(len >> 4) : 3;  // similar lengths share waves
        } else if (literal(s, len, "true", 4)) {
            v.type = kValTrue;
        } else if (literal(s, len, "false", 5)) {
            v.type = kValFalse;
        } else if (literal(s, len, "null", 4)) {
            v.type = kValNull;
        } else {
            list = kNumList;
        }
        if (list < 0) {
            out[h] = v;
        } else {
            const uint32_t k = atomicAdd(&s_cnt[list], 1u);
            s_list[list][k] = (uint8_t)threadIdx.x;
            s_a[threadIdx.x] = g.a;
            s_b[threadIdx.x] = g.b;
            s_next[threadIdx.x] = g.next;
        }
    }
    __syncthreads();
    const uint32_t c0 = s_cnt[0], c1 = c0 + s_cnt[1], c2 = c1 + s_cnt[2], c3 = c2 + s_cnt[3];
    if (threadIdx.x < c3) {  // the string lists in order over the block's threads
        const uint32_t i = threadIdx.x;
        const uint32_t t = i < c0   ? s_list[0][i]
                           : i < c1 ? s_list[1][i - c0]
                           : i < c2 ? s_list[2][i - c1]
                                    : s_list[3][i - c2];
        const uint64_t hs = h0 + t;
        const uint32_t a = s_a[t], len = s_b[t] - a;
        const uint8_t *s = src + (a - off);
        ValueOut v;
        v.v = 0;
        v.row = (uint32_t)hs;
        v.type = kValInvalid;
        v.flags = 0;
        v.reserved = 0;
        uint32_t out_len = 0;
        uint8_t flags = 0;
        if (len >= 2 && s[len - 1] == '"' &&
            unescape(s + 1, len - 2, strings ? strings + a + 1 : nullptr, out_len, flags)) {
            v.type = s_next[t] == ':' ? kValKey : kValString;
            v.flags = flags;
            v.v = out_len;
        }
        v.text_off = a + 1;
        v.text_len = len >= 2 ? len - 2 : 0;
        out[hs] = v;
    }
    if (threadIdx.x < s_cnt[kNumList]) {  // a number (or invalid text)
        const uint32_t t = s_list[kNumList][threadIdx.x];
        const uint64_t hn = h0 + t;
        const uint32_t a = s_a[t], len = s_b[t] - a;
        ValueOut v;
        v.row = (uint32_t)hn;
        v.text_off = a;
        v.text_len = len;
        v.reserved = 0;
        NumText nt;
        uint64_t m_lo = 0;
        int64_t e_ulp = 0, iv = 0;
        uint8_t flags = 0;
        v.type = parse_number(src + (a - off), len, iv, flags, nt, m_lo, e_ulp);
        v.v = iv;
        if (flags & kValExact) {
            const uint32_t k = atomicAdd(exact_count, 1u);
            exact_list[k] = (uint32_t)hn;
        }
        v.flags = flags & (uint8_t)~kValExact;
        out[hn] = v;
    }
}

__global__ __launch_bounds__(64) void k_values_exact(const uint8_t *__restrict__ in, uint64_t n,
                                                     const int32_t *__restrict__ st, uint64_t N,
                                                     ValueOut *__restrict__ out,
                                                     const uint32_t *__restrict__ exact_list,
                                                     const uint32_t *__restrict__ exact_count) {
    const uint32_t cnt = *exact_count;
    for (uint32_t i = blockIdx.x * 64u + threadIdx.x; i < cnt; i += gridDim.x * 64u) {
        const uint64_t h = exact_list[i];
        const GapText g = gap_of(in, 0, n, st, N, h);
        NumText t;
        uint64_t m_lo = 0;
        int64_t e_ulp = 0, iv = 0;
        uint8_t flags = 0;
        parse_number(in + g.a, g.b - g.a, iv, flags, t, m_lo, e_ulp);
        finish_exact(in + g.a, t, m_lo, e_ulp, iv, flags);
        out[h].v = iv;
        out[h].flags = flags;
    }
}

__global__ void k_values_reset(uint32_t *__restrict__ exact_count) { *exact_count = 0; }

hipError_t launch_values(hipStream_t s, const uint8_t *d_in, uint64_t n, const int32_t *st,
                         uint64_t N, ValueOut *out, uint8_t *strings, uint32_t *exact_list,
                         uint32_t *exact_count) {
    hipLaunchKernelGGL(k_values_reset, dim3(1), dim3(1), 0, s, exact_count);
    hipLaunchKernelGGL(k_values, dim3((uint32_t)((N + 1 + 255) / 256)), dim3(256), 0, s, d_in, n, st,
                       N, out, strings, exact_list, exact_count);
    hipLaunchKernelGGL(k_values_exact, dim3(64), dim3(64), 0, s, d_in, n, st, N, out, exact_list,
                       exact_count);
    return hipGetLastError();
}

}  // namespace cjx

// Host runs of the decoders, for the self-tests that run without a GPU.
extern "C" int cjx_debug_parse_number(const uint8_t *s, size_t len, int64_t *value,
                                      uint8_t *flags) {
    using namespace cjx;
    NumText t;
    uint64_t m_lo = 0;
    int64_t e_ulp = 0, iv = 0;
    uint8_t f = 0;
    const uint8_t type = parse_number(s, (uint32_t)len, iv, f, t, m_lo, e_ulp);
    if (f & kValExact) finish_exact(s, t, m_lo, e_ulp, iv, f);
    *value = iv;
    *flags = f;
    return type;
}

extern "C" int cjx_debug_unescape(const uint8_t *s, size_t len, uint8_t *out, size_t *out_len,
                                  uint8_t *flags) {
    uint32_t k = 0;
    uint8_t f = 0;
    const bool ok = cjx::unescape(s, (uint32_t)len, out, k, f);
    *out_len = k;
    *flags = f;
    return ok ? 1 : 0;
}
