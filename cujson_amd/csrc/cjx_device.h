// cjx_device.h - wave64 helpers shared by the kernels (DPP scans, lane
// shifts, agent-scope status words, the escape/string carry map).
#pragma once
#include "cjx_internal.h"

namespace cjx {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << __lane_id()) - 1ull; }

// DPP control codes (GFX9 encoding).
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114, kDppRowShr8 = 0x118;
constexpr int kDppWaveShr1 = 0x138, kDppWaveRor1 = 0x13C, kDppRowBcast15 = 0x142,
              kDppRowBcast31 = 0x143;

template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ uint32_t dpp(uint32_t old, uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, ROW_MASK, 0xF, false);
}
// Row shift whose lanes without a source read 0 (bound_ctrl), so no old
// value is needed and the shift can fold into the consuming add.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_zero(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, true);
}

// Inclusive prefix sum over the 64 lanes of a wave (row shifts, then row
// broadcasts 15 and 31).
// The result is opaque to the optimizer: otherwise a caller's `incl - x`
// becomes the sum of the six shifted partials, which keeps them alive and
// stops each shift from folding into its add.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    x += dpp_zero<kDppRowShr1>(x);
    x += dpp_zero<kDppRowShr2>(x);
    x += dpp_zero<kDppRowShr4>(x);
    x += dpp_zero<kDppRowShr8>(x);
    x += dpp<kDppRowBcast15, 0xA>(0u, x);
    x += dpp<kDppRowBcast31, 0xC>(0u, x);
    __asm__("" : "+v"(x));
    return x;
}

// Value of lane l-1 in lane l; lane 0 receives `carry`.
__device__ __forceinline__ uint32_t wave_shr1(uint32_t v, uint32_t carry) {
    return dpp<kDppWaveShr1>(carry, v);
}

// Value of lane l-1 in lane l; lane 0 receives lane 63 of `prev` (the same
// register one step earlier): a rotate of prev, then a shift of v over it.
__device__ __forceinline__ uint32_t wave_shr1_from(uint32_t v, uint32_t prev) {
    const uint32_t rot = (uint32_t)__builtin_amdgcn_mov_dpp((int)prev, kDppWaveRor1, 0xF, 0xF, false);
    return dpp<kDppWaveShr1>(rot, v);
}

// Set bits of a wave-wide mask in the lanes below this one (v_mbcnt).
__device__ __forceinline__ uint32_t mbcnt64(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint32_t wave_read(uint32_t x, int lane) {
    return (uint32_t)__builtin_amdgcn_readlane((int)x, lane);
}

__device__ __forceinline__ uint32_t agent_load(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void agent_store(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long agent_load64(const unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void agent_store64(unsigned long long *p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Byte classes by three v_perm table lookups over a word's four bytes (the
// tokenizer and the strict UTF-8 check): lookups indexed by byte bits 0-2,
// 3-5 and 6-7 are ANDed; every class bit is a product set in those three
// coordinates, so the AND is exact.  Class bits of the tokenizer's tables:
//   bit 0 ':' (and '\n' for JSON Lines)   bit 4 '\\'
//   bit 1 CONT (0x80-0xBF)                 bit 5 E    (>= 0xE0)
//   bit 2 '[' ']' '{' '}'                  bit 6 F    (>= 0xF0)
//   bit 3 ','                              bit 7 '"'
// ---------------------------------------------------------------------------
// Table entries, byte i = value for index i.
constexpr uint64_t kLutBits02 = 0x6262667A66E36262ull;        // by bits 0-2
constexpr uint64_t kLutBits35Std = 0x67622AA216020202ull;     // by bits 3-5
constexpr uint64_t kLutBits35Lines = 0x67622AA216020302ull;   // '\n' joins ':'
constexpr uint64_t kLutBits67 = 0x0000000060021489ull;        // by bits 6-7

// v_perm: byte i of the result is byte sel_i of hi:lo (sel_i < 8); the host
// pass models it for the host runs of the classifiers in the tests (only
// table selectors 0-7 and 12 (zero) occur in these uses).
__host__ __device__ __forceinline__ uint32_t perm(uint32_t hi, uint32_t lo, uint32_t sel) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_perm(hi, lo, sel);
#else
    const uint64_t data = ((uint64_t)hi << 32) | lo;
    uint32_t out = 0;
    for (int i = 0; i < 4; i++) {
        const uint32_t k = (sel >> (8 * i)) & 0xFFu;
        const uint32_t v = k < 8 ? (uint32_t)(data >> (8 * k)) & 0xFFu : (k == 12 ? 0u : 0xFFu);
        out |= v << (8 * i);
    }
    return out;
#endif
}

template <bool JSONL>
__host__ __device__ __forceinline__ uint32_t classify(uint32_t r) {
    constexpr uint64_t t35 = JSONL ? kLutBits35Lines : kLutBits35Std;
    const uint32_t a = perm((uint32_t)(kLutBits02 >> 32), (uint32_t)kLutBits02, r & 0x07070707u);
    const uint32_t b = perm((uint32_t)(t35 >> 32), (uint32_t)t35, (r >> 3) & 0x07070707u);
    const uint32_t c = perm(0u, (uint32_t)kLutBits67, (r >> 6) & 0x03030303u);
    return a & b & c;
}

// 4x4 byte transpose: o_k = [a.k, b.k, c.k, e.k].
__host__ __device__ __forceinline__ void transpose4(uint32_t a, uint32_t b, uint32_t c, uint32_t e,
                                           uint32_t &o0, uint32_t &o1, uint32_t &o2,
                                           uint32_t &o3) {
    const uint32_t ab_lo = perm(b, a, 0x05010400u), ab_hi = perm(b, a, 0x07030602u);
    const uint32_t ce_lo = perm(e, c, 0x05010400u), ce_hi = perm(e, c, 0x07030602u);
    o0 = perm(ce_lo, ab_lo, 0x05040100u);
    o1 = perm(ce_lo, ab_lo, 0x07060302u);
    o2 = perm(ce_hi, ab_hi, 0x05040100u);
    o3 = perm(ce_hi, ab_hi, 0x07060302u);
}

// Bytes 3-6 of hi:lo, i.e. (hi << 8) | (lo >> 24): the previous word's top
// byte shifted in below.
__host__ __device__ __forceinline__ uint32_t alignbyte3(uint32_t hi, uint32_t lo) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_alignbyte(hi, lo, 3u);
#else
    return (hi << 8) | (lo >> 24);
#endif
}

// ---------------------------------------------------------------------------
// Escape / string carry as an affine map over GF(2).
// A span of input maps the state (c = parity of the backslash run ending at
// its start, s = in-string parity at its start) to the state at its end:
//   c' = a.c ^ b,   s' = s ^ q ^ d.c
// a: the span is all backslashes; b: its trailing run parity; q: parity of
// its unescaped quotes computed with c = s = 0; d: its first non-backslash
// byte is a quote (only that quote's escape status depends on c).
// Bits: a=1, b=2, d=4, q=8.  States: c=1, s=2.
// ---------------------------------------------------------------------------
constexpr uint32_t kMapIdentity = 1u;

__device__ __forceinline__ uint32_t map_compose(uint32_t t1, uint32_t t2) {  // t1 then t2
    const uint32_t a1 = t1 & 1, b1 = (t1 >> 1) & 1, d1 = (t1 >> 2) & 1, q1 = (t1 >> 3) & 1;
    const uint32_t a2 = t2 & 1, b2 = (t2 >> 1) & 1, d2 = (t2 >> 2) & 1, q2 = (t2 >> 3) & 1;
    const uint32_t a = a1 & a2, b = (a2 & b1) ^ b2, d = d1 ^ (d2 & a1), q = q1 ^ q2 ^ (d2 & b1);
    return a | (b << 1) | (d << 2) | (q << 3);
}
__device__ __forceinline__ uint32_t map_apply(uint32_t t, uint32_t st) {
    const uint32_t a = t & 1, b = (t >> 1) & 1, d = (t >> 2) & 1, q = (t >> 3) & 1;
    const uint32_t c = st & 1, s = (st >> 1) & 1;
    return ((a & c) ^ b) | ((s ^ q ^ (d & c)) << 1);
}

}  // namespace cjx
