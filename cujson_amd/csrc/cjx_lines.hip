// cjx_lines.hip - newline scan of JSON Lines text in device memory.
//
// The loader splits JSON Lines text at every '\n' byte (strings are not
// looked at: a raw newline cannot occur inside a valid JSON string) and cuts
// chunks at line starts (cujson/load_file.cu:72-81 scans the host buffer,
// :83-103 cuts by count, :153-176 by size).  Here the newline offsets, i.e.
// the record ends, come from the device:
//   k_nl_count   '\n' bytes of each 256 KiB segment (one block per segment)
//   k_nl_scan    one block: segment bases (exclusive prefix, 64-bit) and the
//                total, plus the last byte of the text
//   k_nl_write   the offset of every '\n', in order
//   k_nl_gather  line starts at given line indices (chunk boundaries)
// HBM-bound: the text is read twice (count, then write) and 8 bytes are
// written per newline.
#include "cjx_device.h"

namespace cjx {

namespace {

constexpr int kNlThreads = 256;
constexpr uint32_t kNlStep = kNlThreads * 16;               // bytes per block step
constexpr uint32_t kNlSteps = 64;
constexpr uint64_t kNlSeg = (uint64_t)kNlStep * kNlSteps;  // 256 KiB per block
constexpr int kNlScanThreads = 1024;

// Bit 7 of byte b is set iff byte b of w is '\n'.  (t & 0x7F) + 0x7F stays
// below 0x100, so no carry crosses a byte.
__device__ __forceinline__ uint32_t nl_bits(uint32_t w) {
    const uint32_t t = w ^ 0x0A0A0A0Au;
    return ~(((t & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | t) & 0x80808080u;
}
// The four byte flags of nl_bits as bits 0-3: bits 7, 15, 23, 31 shifted to
// 0, 8, 16, 24 and multiplied by 2^28 + 2^21 + 2^14 + 2^7, which puts them at
// bits 28-31; every other partial product lands on its own bit below 28.
__device__ __forceinline__ uint32_t nl_nibble(uint32_t w) {
    return ((nl_bits(w) >> 7) * 0x10204080u) >> 28;
}

// Newline flags of the 16 bytes at `off` (bit i: byte off + i), bytes at or
// past n cleared.  Whole 16-byte groups are read as one aligned load; the
// group holding byte n - 1, when n is not a multiple of 16, byte by byte.
__device__ __forceinline__ uint32_t nl_mask16(const uint8_t *__restrict__ in, uint64_t off,
                                              uint64_t n) {
    if (off >= n) return 0u;
    if (off + 16u <= n) {
        const uint4 v = *reinterpret_cast<const uint4 *>(in + off);
        return nl_nibble(v.x) | (nl_nibble(v.y) << 4) | (nl_nibble(v.z) << 8) |
               (nl_nibble(v.w) << 12);
    }
    uint32_t m = 0;
    for (uint32_t i = 0; off + i < n; i++) m |= (in[off + i] == '\n' ? 1u : 0u) << i;
    return m;
}

}  // namespace

__global__ __launch_bounds__(kNlThreads) void k_nl_count(const uint8_t *__restrict__ in,
                                                         uint64_t n,
                                                         uint32_t *__restrict__ seg_count) {
    __shared__ uint32_t red[kNlThreads / 64];
    const uint64_t seg = (uint64_t)blockIdx.x * kNlSeg;
    uint32_t c = 0;
#pragma unroll 4
    for (uint32_t s = 0; s < kNlSteps; s++)
        c += (uint32_t)__popc(nl_mask16(in, seg + (uint64_t)s * kNlStep + 16u * threadIdx.x, n));
    for (int d = 32; d >= 1; d >>= 1) c += (uint32_t)__shfl_xor((int)c, d, 64);
    if (lane_id() == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int k = 0; k < kNlThreads / 64; k++) t += red[k];
        seg_count[blockIdx.x] = t;
    }
}

// One block: seg_base[s] = newlines before segment s (s <= n_seg, the last
// entry is the total); stat[0] = the total, stat[1] = the last byte (0 for
// empty text).  Tiles of 1024 counts (coalesced loads) are scanned in turn:
// wave scans, the waves' totals in LDS, and a 64-bit carry between tiles (a
// tile's counts sum to at most 2^28).
__global__ __launch_bounds__(kNlScanThreads) void k_nl_scan(const uint8_t *__restrict__ in,
                                                            uint64_t n,
                                                            const uint32_t *__restrict__ seg_count,
                                                            uint32_t n_seg,
                                                            uint64_t *__restrict__ seg_base,
                                                            uint64_t *__restrict__ stat) {
    constexpr uint32_t kW = kNlScanThreads / 64;
    __shared__ uint32_t wsum[kW];
    const uint32_t t = threadIdx.x, w = t >> 6;
    uint64_t carry = 0;
    for (uint32_t base = 0; base < n_seg; base += kNlScanThreads) {
        const uint32_t i = base + t;
        const uint32_t v = i < n_seg ? seg_count[i] : 0u;
        const uint32_t incl = wave_incl_scan(v);
        if (lane_id() == 63) wsum[w] = incl;
        __syncthreads();
        uint32_t pre = incl - v, tile = 0;
        for (uint32_t k = 0; k < kW; k++) {
            if (k < w) pre += wsum[k];
            tile += wsum[k];
        }
        if (i < n_seg) seg_base[i] = carry + pre;
        carry += tile;
        __syncthreads();  // wsum is rewritten by the next tile
    }
    if (t == 0) {
        seg_base[n_seg] = carry;
        stat[0] = carry;
        stat[1] = n ? in[n - 1] : 0u;
    }
}

// Block s writes the offsets of the newlines of segment s from seg_base[s]
// on; entries at or past cap are not written.
__global__ __launch_bounds__(kNlThreads) void k_nl_write(const uint8_t *__restrict__ in,
                                                         uint64_t n,
                                                         const uint64_t *__restrict__ seg_base,
                                                         uint64_t *__restrict__ out,
                                                         uint64_t cap) {
    __shared__ uint32_t wsum[kNlThreads / 64];
    const uint64_t seg = (uint64_t)blockIdx.x * kNlSeg;
    const uint32_t w = threadIdx.x >> 6;
    uint64_t base = seg_base[blockIdx.x];
    for (uint32_t s = 0; s < kNlSteps; s++) {
        const uint64_t off = seg + (uint64_t)s * kNlStep + 16u * threadIdx.x;
        uint32_t m = nl_mask16(in, off, n);
        const uint32_t c = (uint32_t)__popc(m);
        const uint32_t incl = wave_incl_scan(c);
        if (lane_id() == 63) wsum[w] = incl;
        __syncthreads();
        uint64_t idx = base + incl - c;
        uint32_t step_total = 0;
        for (uint32_t k = 0; k < (uint32_t)(kNlThreads / 64); k++) {
            if (k < w) idx += wsum[k];
            step_total += wsum[k];
        }
        while (m) {
            const uint32_t j = (uint32_t)__builtin_ctz(m);
            m &= m - 1u;
            if (idx < cap) out[idx] = off + j;
            idx++;
        }
        base += step_total;
        __syncthreads();  // wsum is rewritten by the next step
    }
}

// out[i] = start of line line_idx[i]: 0 for line 0, the byte after the
// (j-1)-th newline for line j <= n_nl, n beyond.
__global__ __launch_bounds__(256) void k_nl_gather(const uint64_t *__restrict__ nl, uint64_t n_nl,
                                                   uint64_t n,
                                                   const uint64_t *__restrict__ line_idx,
                                                   uint64_t count, uint64_t *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= count) return;
    const uint64_t j = line_idx[i];
    out[i] = j == 0 ? 0ull : (j <= n_nl ? nl[j - 1] + 1u : n);
}

// Placement of units that do not start at a 16-byte aligned offset: block b
// fills bytes [64 KiB b, 64 KiB (b+1)) of the packed layout dst, where unit
// u occupies [units[u].in_off, + units[u].n) (ascending, 16-byte aligned) and
// comes from src + src_off[u] (src 16-byte aligned).  A thread moves 16
// bytes per step, cut from the two aligned 16-byte source words that hold
// them; the misalignment (src_off[u] mod 16) is the same for every step of a
// unit, so the cut is a uniform branch.  For a whole group the second word
// ends at round16(s + 16) <= src_limit (the readable end of the source,
// rounded up to 16); a unit's last, partial group moves byte by byte.
// Bytes between units are left as they are.
constexpr uint64_t kPlacePiece = 64 * 1024;
__global__ __launch_bounds__(256) void k_place_units(const uint8_t *__restrict__ src,
                                                     uint64_t src_limit,
                                                     const uint64_t *__restrict__ src_off,
                                                     const UnitDesc *__restrict__ units,
                                                     uint32_t n_units, uint8_t *__restrict__ dst,
                                                     uint64_t dst_bytes) {
    __shared__ uint32_t s_u0;
    const uint64_t p0 = (uint64_t)blockIdx.x * kPlacePiece;
    const uint64_t p1 = min(p0 + kPlacePiece, dst_bytes);
    if (threadIdx.x == 0) {  // last unit starting at or before p0
        uint32_t lo = 0, hi = n_units;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (units[mid].in_off <= p0) lo = mid; else hi = mid;
        }
        s_u0 = lo;
    }
    __syncthreads();
    for (uint32_t u = s_u0; u < n_units && units[u].in_off < p1; u++) {
        const uint64_t d0 = units[u].in_off, d1 = d0 + units[u].n;
        const uint64_t a = max(p0, d0), b = min(p1, d1);
        for (uint64_t d = a + 16u * threadIdx.x; d < b; d += 16u * 256u) {
            const uint64_t s = src_off[u] + (d - d0);
            const uint64_t len = min((uint64_t)16, b - d);
            const uint64_t sa = s & ~15ull;
            const uint32_t mis = (uint32_t)(s & 15u);
            if (len == 16 && (mis == 0 || sa + 32u <= src_limit)) {
                const uint4 v0 = *reinterpret_cast<const uint4 *>(src + sa);
                uint4 o = v0;
                if (mis) {
                    const uint4 v1 = *reinterpret_cast<const uint4 *>(src + sa + 16);
                    const uint32_t r = mis & 3u;
                    // dwords mis/4 .. mis/4 + 4 of the 32-byte window, cut at byte r
                    uint32_t w0, w1, w2, w3, w4;
                    switch (mis >> 2) {
                        case 0: w0 = v0.x; w1 = v0.y; w2 = v0.z; w3 = v0.w; w4 = v1.x; break;
                        case 1: w0 = v0.y; w1 = v0.z; w2 = v0.w; w3 = v1.x; w4 = v1.y; break;
                        case 2: w0 = v0.z; w1 = v0.w; w2 = v1.x; w3 = v1.y; w4 = v1.z; break;
                        default: w0 = v0.w; w1 = v1.x; w2 = v1.y; w3 = v1.z; w4 = v1.w; break;
                    }
                    o.x = __builtin_amdgcn_alignbyte(w1, w0, r);
                    o.y = __builtin_amdgcn_alignbyte(w2, w1, r);
                    o.z = __builtin_amdgcn_alignbyte(w3, w2, r);
                    o.w = __builtin_amdgcn_alignbyte(w4, w3, r);
                }
                *reinterpret_cast<uint4 *>(dst + d) = o;
            } else {
                for (uint32_t i = 0; i < (uint32_t)len; i++) dst[d + i] = src[s + i];
            }
        }
    }
}

hipError_t launch_place_units(hipStream_t s, const uint8_t *src, uint64_t src_limit,
                              const uint64_t *src_off, const UnitDesc *units, uint32_t n_units,
                              uint8_t *dst, uint64_t dst_bytes) {
    const uint64_t g = (dst_bytes + kPlacePiece - 1) / kPlacePiece;
    if (g && n_units)
        hipLaunchKernelGGL(k_place_units, dim3((uint32_t)g), dim3(256), 0, s, src, src_limit,
                           src_off, units, n_units, dst, dst_bytes);
    return hipGetLastError();
}

uint32_t nl_segments(uint64_t n) { return (uint32_t)((n + kNlSeg - 1) / kNlSeg); }

hipError_t launch_nl_count(hipStream_t s, const uint8_t *d_in, uint64_t n, uint32_t *seg_count,
                           uint64_t *seg_base, uint64_t *stat) {
    const uint32_t g = nl_segments(n);
    if (g) hipLaunchKernelGGL(k_nl_count, dim3(g), dim3(kNlThreads), 0, s, d_in, n, seg_count);
    hipLaunchKernelGGL(k_nl_scan, dim3(1), dim3(kNlScanThreads), 0, s, d_in, n, seg_count, g,
                       seg_base, stat);
    return hipGetLastError();
}

hipError_t launch_nl_write(hipStream_t s, const uint8_t *d_in, uint64_t n,
                           const uint64_t *seg_base, uint64_t *out, uint64_t cap) {
    const uint32_t g = nl_segments(n);
    if (g) hipLaunchKernelGGL(k_nl_write, dim3(g), dim3(kNlThreads), 0, s, d_in, n, seg_base, out, cap);
    return hipGetLastError();
}

hipError_t launch_nl_gather(hipStream_t s, const uint64_t *nl, uint64_t n_nl, uint64_t n,
                            const uint64_t *line_idx, uint64_t count, uint64_t *out) {
    if (count)
        hipLaunchKernelGGL(k_nl_gather, dim3((uint32_t)((count + 255) / 256)), dim3(256), 0, s, nl,
                           n_nl, n, line_idx, count, out);
    return hipGetLastError();
}

}  // namespace cjx
