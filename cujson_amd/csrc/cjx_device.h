// cjx_device.h - wave64 helpers shared by the kernels (DPP scans, lane
// shifts, agent-scope status words, the escape/string carry map).
#pragma once
#include "cjx_internal.h"

namespace cjx {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << __lane_id()) - 1ull; }

// DPP control codes (GFX9 encoding).
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114, kDppRowShr8 = 0x118;
constexpr int kDppWaveShr1 = 0x138, kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143;

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
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    x += dpp_zero<kDppRowShr1>(x);
    x += dpp_zero<kDppRowShr2>(x);
    x += dpp_zero<kDppRowShr4>(x);
    x += dpp_zero<kDppRowShr8>(x);
    x += dpp<kDppRowBcast15, 0xA>(0u, x);
    x += dpp<kDppRowBcast31, 0xC>(0u, x);
    return x;
}

// Value of lane l-1 in lane l; lane 0 receives `carry`.
__device__ __forceinline__ uint32_t wave_shr1(uint32_t v, uint32_t carry) {
    return dpp<kDppWaveShr1>(carry, v);
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
