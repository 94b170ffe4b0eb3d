// cjx_utf8.hip - strict (RFC 3629) UTF-8 validation, the CJX_FLAG_UTF8_STRICT
// mode.  The reference only has its weaker per-word check (UTF8Validation,
// parse_standard_json.cu:264-327), which the tokenizer computes by default;
// strict mode replaces it with this kernel.  The result equals
// simdjson::validate_utf8 (related_works/simdjson/simdjson.h:3212).
//
// Validity of a byte string equals these local conditions at every position
// p, with a = b[p-1], b = b[p-2], c = b[p-3] (zero before the unit start) and
// three zero bytes after the end:
//   (1) b[p] is a continuation byte (10xxxxxx) exactly when a lead expects
//       one: a >= C0, or b >= E0, or c >= F0;
//   (2) b[p] is none of C0, C1, F5..FF;
//   (3) second-byte ranges: after E0 at least A0 (no overlong), after ED at
//       most 9F (no surrogate), after F0 at least 90 (no overlong), after F4
//       at most 8F (at most U+10FFFF).
// Where b[p] is not a continuation byte after a lead, (1) already fails, so
// (3) is tested on the byte's bits 4 and 5 alone: below A0 <=> bit 5 clear,
// below 90 <=> bits 4 and 5 clear.  The zero bytes past the end make a
// sequence cut by the end fail (1).
//
// Layout: a lane checks 16 bytes per step (one coalesced 16-byte load), a
// wave 1 KiB.  The lane's four words are byte-transposed (word d, byte b =
// input byte 4b + d), so the byte before another is the previous word, or
// across the slot edge the previous lane's word shifted in by DPP.  Two
// three-way v_perm classifications per word give the classes: the
// tokenizer's tables (CONT at bit 1, E >= E0 at bit 5, F >= F0 at bit 6) and
// the lead table below; the conditions are then a few bit operations per
// word.  A step whose wave holds no byte >= 80, after a step that ended in no
// lead, is skipped.
#include "cjx_device.h"

namespace cjx {

// Lead classes by byte bits 0-2 / 3-5 / 6-7, ANDed (each bit a product set):
//   bit 0 C0, C1    bit 3 F4    bit 6 ED
//   bit 1 F5..F7    bit 4 F0    bit 7 L (>= C0)
//   bit 2 F8..FF    bit 5 E0
constexpr uint64_t kLeadBits02 = 0x8686C68C848485B5ull;
constexpr uint64_t kLeadBits35 = 0x849AC0A080808081ull;
constexpr uint32_t kLeadBits67 = 0xFF000000u;

__host__ __device__ __forceinline__ uint32_t lead_class(uint32_t r) {
    const uint32_t a = perm((uint32_t)(kLeadBits02 >> 32), (uint32_t)kLeadBits02, r & 0x07070707u);
    const uint32_t b = perm((uint32_t)(kLeadBits35 >> 32), (uint32_t)kLeadBits35,
                            (r >> 3) & 0x07070707u);
    const uint32_t c = perm(0u, kLeadBits67, (r >> 6) & 0x03030303u);
    return a & b & c;
}

// Error bits (nonzero = invalid) at the four bytes of transposed word t with
// classes kx (tokenizer tables) and sx (lead table); sa: lead classes of the
// bytes before them, kb / kc: tokenizer classes of the bytes two and three
// before.
__host__ __device__ __forceinline__ uint32_t strict_errors(uint32_t t, uint32_t kx, uint32_t sx,
                                                           uint32_t sa, uint32_t kb, uint32_t kc) {
    const uint32_t must = sa | (kb << 2) | (kc << 1);        // bit 7: L(a) | E(b) | F(c)
    const uint32_t e1 = (must ^ (kx << 6)) & 0x80808080u;    // (1), CONT(x) at bit 7
    const uint32_t e2 = sx & 0x07070707u;                    // (2)
    // (3) at bit 5: bit 5 of x selects ED (lead bit 6) over E0 (lead bit 5);
    // at bit 4: bit 4 or 5 of x selects F4 (lead bit 3) over F0 (lead bit 4)
    const uint32_t e3 = ((t & (sa >> 1)) | (~t & sa)) & 0x20202020u;
    const uint32_t t45 = t | (t >> 1);
    const uint32_t e4 = ((t45 & (sa << 1)) | (~t45 & sa)) & 0x10101010u;
    return e1 | e2 | e3 | e4;
}

// One lane's 16 bytes: the four transposed words and their classes.
struct StrictWords {
    uint32_t t[4], k[4], s[4];
};
__host__ __device__ __forceinline__ StrictWords strict_words(uint32_t w0, uint32_t w1, uint32_t w2,
                                                             uint32_t w3) {
    StrictWords x;
    transpose4(w0, w1, w2, w3, x.t[0], x.t[1], x.t[2], x.t[3]);
#pragma unroll
    for (int d = 0; d < 4; d++) {
        x.k[d] = classify<false>(x.t[d]);
        x.s[d] = lead_class(x.t[d]);
    }
    return x;
}

// What the next lane needs of this one: tokenizer classes of transposed
// words 1-3 and lead classes of word 3 (only their top bytes, the lane's
// last bytes 13-15, are used).
struct StrictCarry {
    uint32_t k1, k2, k3, s3;
};
__host__ __device__ __forceinline__ StrictCarry strict_carry(const StrictWords &x) {
    return StrictCarry{x.k[1], x.k[2], x.k[3], x.s[3]};
}

// Error bits of a lane's 16 bytes given the previous lane's carry.  Byte
// 4b + d follows byte 4b + d - 1: word d - 1 of the same slot, or for d = 0
// word 3 of the slot below (the previous lane's for slot 0).
__host__ __device__ __forceinline__ uint32_t strict_lane_errors(const StrictWords &x,
                                                                const StrictCarry &prev) {
    const uint32_t ak1 = alignbyte3(x.k[1], prev.k1), ak2 = alignbyte3(x.k[2], prev.k2);
    const uint32_t ak3 = alignbyte3(x.k[3], prev.k3), as3 = alignbyte3(x.s[3], prev.s3);
    uint32_t err = strict_errors(x.t[0], x.k[0], x.s[0], as3, ak2, ak1);
    err |= strict_errors(x.t[1], x.k[1], x.s[1], x.s[0], ak3, ak2);
    err |= strict_errors(x.t[2], x.k[2], x.s[2], x.s[1], x.k[0], ak3);
    err |= strict_errors(x.t[3], x.k[3], x.s[3], x.s[2], x.k[1], x.k[0]);
    return err;
}

// The carry of the four bytes before a wave's first byte (word pw, byte j =
// byte j - 4 relative to it), as the previous lane's words would hold them.
__host__ __device__ __forceinline__ StrictCarry strict_carry_of(uint32_t pw) {
    const uint32_t k = classify<false>(pw);
    return StrictCarry{k << 16, k << 8, k, lead_class(pw)};
}

// Whether the top bytes of a carry hold a lead the next bytes must follow.
__host__ __device__ __forceinline__ bool strict_carry_hot(const StrictCarry &c) {
    return ((c.s3 & 0x80000000u) | (c.k2 & 0x60000000u) | (c.k1 & 0x40000000u)) != 0;
}

namespace {

constexpr int kU8Threads = 256;
constexpr uint32_t kU8Step = 1024;  // bytes per wave step

}  // namespace

// One block per tokenizer tile (four waves of 16 KiB), then one block per
// small unit (four waves of 4 KiB).
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
    const uint32_t w = threadIdx.x >> 6, lane = lane_id();
    const uint64_t span = small ? (uint64_t)kWaveBytes / kWaves : (uint64_t)kTileBytes / kWaves;
    const uint64_t s0 = (small ? 0ull : (uint64_t)(blk - U.tile0) * kTileBytes) + w * span;
    if (s0 >= n + 3) return;  // the whole wave lies past the end and its three zero bytes
    // The four bytes before the wave (the input is readable up to n rounded
    // up to 16, and s0 - 4 < n here); bytes at or past n read as zero.
    uint32_t pw = 0;
    if (s0 >= 4) {
        pw = *reinterpret_cast<const uint32_t *>(base + s0 - 4);
        if (s0 > n) pw &= (1u << (8 * (uint32_t)(n - (s0 - 4)))) - 1u;
    }
    StrictCarry carry = strict_carry_of(pw);
    bool hot = strict_carry_hot(carry);
    uint32_t err = 0;
    for (uint64_t st = s0; st < s0 + span && st < n + 3; st += kU8Step) {
        const uint64_t o = st + 16ull * lane;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (o < n) {  // [o, o + 16) lies below n rounded up to 16
            v = *reinterpret_cast<const uint4 *>(base + o);
            if (o + 16 > n) {
                uint32_t x[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int64_t keep = (int64_t)(n - o) - 4 * k;
                    if (keep <= 0) x[k] = 0;
                    else if (keep < 4) x[k] &= (1u << (8 * keep)) - 1u;
                }
                v = make_uint4(x[0], x[1], x[2], x[3]);
            }
        }
        const uint32_t high = (v.x | v.y | v.z | v.w) & 0x80808080u;
        if (!__ballot(high != 0) && !hot) {  // all ASCII, no lead before: nothing to check
            carry = StrictCarry{0, 0, 0, 0};
            continue;
        }
        // the previous lane's carry (lane 0: the previous step's lane 63, or
        // the bytes before the wave)
        const StrictWords x = strict_words(v.x, v.y, v.z, v.w);
        const StrictCarry mine = strict_carry(x);
        StrictCarry prev;
        prev.k1 = wave_shr1(mine.k1, carry.k1);
        prev.k2 = wave_shr1(mine.k2, carry.k2);
        prev.k3 = wave_shr1(mine.k3, carry.k3);
        prev.s3 = wave_shr1(mine.s3, carry.s3);
        err |= strict_lane_errors(x, prev);
        carry = StrictCarry{wave_read(mine.k1, 63), wave_read(mine.k2, 63), wave_read(mine.k3, 63),
                            wave_read(mine.s3, 63)};
        hot = strict_carry_hot(carry);
    }
    if (__ballot(err != 0) && lane == 0) atomicOr(&uo[u].utf8_err, 1u);
}

hipError_t launch_utf8_strict(hipStream_t s, const LaunchBufs &b, const uint8_t *d_in,
                              uint32_t ntiles, uint32_t n_small) {
    if (ntiles + n_small == 0) return hipSuccess;
    hipLaunchKernelGGL(k_utf8_strict, dim3(ntiles + n_small), dim3(kU8Threads), 0, s, d_in,
                       b.units, b.tile_unit, b.small_units, ntiles, b.uout);
    return hipGetLastError();
}

}  // namespace cjx

// Host run of the strict check's word-level steps over data[0..n), lane by
// lane in order (16 bytes each, zero-padded past the end, plus one lane of
// zeros for the three bytes past the end), for the test that runs without a
// GPU; returns 1 if the bytes are invalid.  The device kernel runs the same
// steps with the lanes side by side.
extern "C" int cjx_debug_strict_utf8(const uint8_t *data, size_t n) {
    using namespace cjx;
    StrictCarry prev{0, 0, 0, 0};
    uint32_t err = 0;
    for (size_t o = 0; o < n + 3; o += 16) {
        uint32_t w[4] = {0, 0, 0, 0};
        for (int i = 0; i < 16; i++)
            if (o + i < n) w[i / 4] |= (uint32_t)data[o + i] << (8 * (i % 4));
        const StrictWords x = strict_words(w[0], w[1], w[2], w[3]);
        err |= strict_lane_errors(x, prev);
        prev = strict_carry(x);
    }
    return err != 0;
}

extern "C" uint32_t cjx_debug_lead_class(uint32_t byte) { return cjx::lead_class(byte & 0xFFu) & 0xFFu; }
