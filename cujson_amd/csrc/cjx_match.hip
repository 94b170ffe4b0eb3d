// cjx_match.hip - phase 3 for well-nested units: bracket matching without
// the global sort.
//
// The reference pairs brackets by stably sorting them on a depth key (an
// opener takes the depth before it, a closer the depth after it, modulo 2^8
// or 2^16, parse_standard_json.cu:1537-1546) and pairing sorted positions
// 2j and 2j+1 (validate_expand_MathAPI_new2, parse_standard_json.cu:1431-1514).
// When a unit's depth walk starts and ends at 0, never goes below 0 and
// never above 256, the keys are the exact depths, the brackets of one key
// alternate opener/closer starting with an opener, every key class has an
// even size, and so sorted positions 2j, 2j+1 are an opener and the next
// bracket of the same key: its matching closer.  The order check of the
// pairing rule can then only fail where the xor check fails, so the result
// is nesting matching with an error iff a matched pair mixes '[' and '}' or
// '{' and ']'.  Such units are paired here; every other unit (deeper, or
// unbalanced) goes through the sort path of cjx_kernels.hip, which follows the
// reference step by step.
//
// Brackets are cut into chunks of kMatchChunk.  A chunk matches its own
// brackets with a local stable counting sort in LDS; an opener left open at
// the end of chunk b with key d is closed by the first later chunk whose
// depth walk reaches d, at that chunk's first closer of key d.  Both ends are
// recorded per (chunk, key), so the cross-chunk step is a search over the
// chunks' minimum depths (with 64- and 4096-chunk group minima).
//
// The pair row cells of all structurals are filled with -1 by the tokenizer;
// these kernels write only the openers' cells.
#include "cjx_device.h"

namespace cjx {

namespace {

__device__ __forceinline__ DepthAgg agg_compose(const DepthAgg &a, const DepthAgg &b) {
    DepthAgg r;
    r.d = a.d + b.d;
    r.mn = min(a.mn, a.d + b.mn);
    r.mx = max(a.mx, a.d + b.mx);
    return r;
}

__device__ __forceinline__ int32_t delta_of(uint8_t c) {
    // '[' 0x5B, '{' 0x7B open; ']' 0x5D, '}' 0x7D close (bit 1 tells them apart).
    return (c & 2u) ? 1 : -1;
}

// Last unit whose chunk0 <= c.
__device__ __forceinline__ uint32_t unit_of_chunk(const UnitOut *__restrict__ uo,
                                                  uint32_t n_units, uint32_t c) {
    uint32_t lo = 0, hi = n_units - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (uo[mid].chunk0 <= c) lo = mid;
        else hi = mid - 1;
    }
    // A unit without brackets has no chunks and the same chunk0 as the next
    // unit, so the last unit with chunk0 <= c is the one that owns chunk c.
    return lo;
}

// Ordered reduction of a DepthAgg over a wave (lane order = walk order).
__device__ __forceinline__ DepthAgg wave_reduce_agg(DepthAgg a) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        DepthAgg o;
        o.d = __shfl_xor(a.d, off, 64);
        o.mn = __shfl_xor(a.mn, off, 64);
        o.mx = __shfl_xor(a.mx, off, 64);
        a = (lane & off) ? agg_compose(o, a) : agg_compose(a, o);
    }
    return a;
}

// Inclusive ordered scan of a DepthAgg over a wave.
__device__ __forceinline__ DepthAgg wave_scan_agg(DepthAgg a) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        DepthAgg o;
        o.d = __shfl_up(a.d, off, 64);
        o.mn = __shfl_up(a.mn, off, 64);
        o.mx = __shfl_up(a.mx, off, 64);
        if (lane >= (uint32_t)off) a = agg_compose(o, a);
    }
    return a;
}

}  // namespace

// K1: depth walk of every chunk.
__global__ __launch_bounds__(kMatchThreads) void k_chunk_stats(const uint8_t *__restrict__ oc_char,
                                                               const UnitOut *__restrict__ uo,
                                                               uint32_t n_units,
                                                               ChunkStat *__restrict__ ck,
                                                               const LaunchState *__restrict__ ls) {
    __shared__ DepthAgg part[kMatchThreads / 64];
    if (ls->overflow) return;
    const uint32_t c = blockIdx.x;
    if (c >= ls->n_chunks) return;
    const uint32_t u = unit_of_chunk(uo, n_units, c);
    const uint32_t first = uo[u].m_base + (c - uo[u].chunk0) * kMatchChunk;
    const uint32_t cnt = min((uint32_t)kMatchChunk, uo[u].m_end - first);
    constexpr int kPer = kMatchChunk / kMatchThreads;
    DepthAgg a{0, 0, 0};
    const uint32_t e0 = threadIdx.x * kPer;
#pragma unroll
    for (int k = 0; k < kPer; k++) {
        if (e0 + k < cnt) {
            a.d += delta_of(oc_char[first + e0 + k]);
            a.mn = min(a.mn, a.d);
            a.mx = max(a.mx, a.d);
        }
    }
    a = wave_reduce_agg(a);
    if (lane_id() == 0) part[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        DepthAgg t = part[0];
        for (int k = 1; k < kMatchThreads / 64; k++) t = agg_compose(t, part[k]);
        ck[c] = ChunkStat{t.d, t.mn, t.mx, 0};
    }
}

// K2, one block per unit: chunk start depths, chunk minima and their group
// minima, and whether the unit is well nested.
__global__ __launch_bounds__(256) void k_chunk_scan(UnitOut *__restrict__ uo,
                                                    ChunkStat *__restrict__ ck,
                                                    int32_t *__restrict__ gmin,
                                                    int32_t *__restrict__ gmin1,
                                                    int32_t *__restrict__ gmin2,
                                                    const LaunchState *__restrict__ ls) {
    __shared__ DepthAgg part[4];
    if (ls->overflow) return;
    const uint32_t u = blockIdx.x, tid = threadIdx.x;
    const uint32_t c0 = uo[u].chunk0;
    const uint32_t nch = (uo[u].m_end - uo[u].m_base + kMatchChunk - 1) / kMatchChunk;
    const uint32_t per = (nch + 255) / 256;
    const uint32_t a0 = c0 + min(nch, tid * per), a1 = c0 + min(nch, (tid + 1) * per);
    DepthAgg a{0, 0, 0};
    for (uint32_t c = a0; c < a1; c++) {
        const ChunkStat s = ck[c];
        a = agg_compose(a, DepthAgg{s.d, s.mn, s.mx});
    }
    DepthAgg incl = wave_scan_agg(a);
    if (lane_id() == 63) part[tid >> 6] = incl;
    __syncthreads();
    DepthAgg pre{0, 0, 0};
    for (uint32_t k = 0; k < (tid >> 6); k++) pre = agg_compose(pre, part[k]);
    DepthAgg excl;
    {
        DepthAgg o;
        o.d = __shfl_up(incl.d, 1, 64);
        o.mn = __shfl_up(incl.mn, 1, 64);
        o.mx = __shfl_up(incl.mx, 1, 64);
        excl = lane_id() ? agg_compose(pre, o) : pre;
    }
    int32_t depth = excl.d;
    for (uint32_t c = a0; c < a1; c++) {
        const ChunkStat s = ck[c];
        ck[c].d0 = depth;
        gmin[c] = depth + s.mn;
        depth += s.d;
    }
    if (tid == 255) {
        const DepthAgg t = agg_compose(pre, incl);
        uo[u].fast = (t.d == 0 && t.mn >= 0 && t.mx <= kMatchMaxDepth) ? 1u : 0u;
    }
    __syncthreads();
    // Group minima for groups that lie inside the unit.
    const uint32_t g1a = (c0 + 63) / 64, g1b = (c0 + nch) / 64;
    for (uint32_t g = g1a + tid; g < g1b; g += 256) {
        int32_t m = gmin[64 * g];
        for (uint32_t k = 1; k < 64; k++) m = min(m, gmin[64 * g + k]);
        gmin1[g] = m;
    }
    __syncthreads();
    const uint32_t g2a = (c0 + 4095) / 4096, g2b = (c0 + nch) / 4096;
    for (uint32_t g = g2a + tid; g < g2b; g += 256) {
        int32_t m = gmin1[64 * g];
        for (uint32_t k = 1; k < 64; k++) m = min(m, gmin1[64 * g + k]);
        gmin2[g] = m;
    }
}

// K3: local matching of one chunk.  The chunk's brackets are stably
// counting-sorted by key in LDS; an opener followed in its key class by
// another bracket is matched to it, an opener last in its class is recorded
// in highs, and a closer first in its class in lows.
__global__ __launch_bounds__(kMatchThreads) void k_match_local(
    const uint8_t *__restrict__ oc_char, const uint32_t *__restrict__ oc_idx,
    UnitOut *__restrict__ uo, uint32_t n_units, const ChunkStat *__restrict__ ck,
    uint32_t *__restrict__ lows, uint32_t *__restrict__ highs, int32_t *__restrict__ pair_row,
    const LaunchState *__restrict__ ls) {
    constexpr int kPer = kMatchChunk / kMatchThreads;
    constexpr int kW = kMatchThreads / 64;
    __shared__ uint8_t s_ch[kMatchChunk];
    __shared__ uint8_t s_key[kMatchChunk];
    __shared__ uint16_t s_sidx[kMatchChunk];  // sorted position -> element
    __shared__ uint16_t s_spos[kMatchChunk];  // element -> sorted position
    __shared__ uint32_t s_idx[kMatchChunk];
    __shared__ uint32_t s_start[kMatchKeys], s_count[kMatchKeys], s_run[kMatchKeys];
    __shared__ uint32_t s_wcnt[kW][kMatchKeys];
    __shared__ int32_t s_wsum[kW];
    __shared__ int32_t s_red[kW];

    if (ls->overflow) return;
    const uint32_t c = blockIdx.x;
    if (c >= ls->n_chunks) return;
    const uint32_t u = unit_of_chunk(uo, n_units, c);
    if (!uo[u].fast) return;
    const uint32_t tid = threadIdx.x, w = tid >> 6;
    const uint32_t first = uo[u].m_base + (c - uo[u].chunk0) * kMatchChunk;
    const uint32_t cnt = min((uint32_t)kMatchChunk, uo[u].m_end - first);
    const uint32_t nb = uo[u].n_base, pbase = uo[u].pair_base + 1u;
    const int32_t d0 = ck[c].d0;

    for (uint32_t e = tid; e < kMatchChunk; e += kMatchThreads) {
        const bool v = e < cnt;
        s_ch[e] = v ? oc_char[first + e] : 0;
        s_idx[e] = v ? oc_idx[first + e] : 0;
    }
    s_count[tid] = 0;
    __syncthreads();

    // Depth keys: thread t walks elements [16t, 16t + 16).
    const uint32_t e0 = tid * kPer;
    int32_t mine = 0;
#pragma unroll
    for (int k = 0; k < kPer; k++)
        if (e0 + k < cnt) mine += delta_of(s_ch[e0 + k]);
    const int32_t incl = (int32_t)wave_incl_scan((uint32_t)mine);
    if (lane_id() == 63) s_wsum[w] = incl;
    __syncthreads();
    int32_t depth = d0 + incl - mine;
    for (uint32_t k = 0; k < w; k++) depth += s_wsum[k];
#pragma unroll
    for (int k = 0; k < kPer; k++) {
        if (e0 + k < cnt) {
            const int32_t dl = delta_of(s_ch[e0 + k]);
            const uint32_t key = (uint32_t)(dl > 0 ? depth : depth - 1);  // in [0, 255]
            depth += dl;
            s_key[e0 + k] = (uint8_t)key;
            atomicAdd(&s_count[key], 1u);
        }
    }
    __syncthreads();
    {
        const uint32_t v = s_count[tid];
        const uint32_t in = wave_incl_scan(v);
        if (lane_id() == 63) s_wsum[w] = (int32_t)in;
        __syncthreads();
        uint32_t ex = in - v;
        for (uint32_t k = 0; k < w; k++) ex += (uint32_t)s_wsum[k];
        s_start[tid] = ex;
        s_run[tid] = ex;
    }
    __syncthreads();

    // Stable ranks: rounds of 256 consecutive elements; inside a round, ranks
    // among equal keys come from a wave match (8 ballots) and per-wave counts.
    const uint64_t lt = lanemask_lt();
    for (uint32_t base = 0; base < cnt; base += kMatchThreads) {
        const uint32_t e = base + tid;
        const bool valid = e < cnt;
        const uint32_t key = valid ? s_key[e] : 0u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
            const uint32_t bv = (key >> bit) & 1u;
            const uint64_t bm = __ballot(bv);
            peers &= bv ? bm : ~bm;
        }
#pragma unroll
        for (int k = 0; k < kW; k++) s_wcnt[k][tid] = 0;
        __syncthreads();
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (valid && rank == 0) s_wcnt[w][key] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = s_run[key] + rank;
            for (uint32_t k = 0; k < w; k++) pos += s_wcnt[k][key];
            s_sidx[pos] = (uint16_t)e;
            s_spos[e] = (uint16_t)pos;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int k = 0; k < kW; k++) add += s_wcnt[k][tid];
        s_run[tid] += add;
        __syncthreads();
    }

    // Pairs, in element order so that the row writes stay close together.
    int32_t err = 0;
    for (uint32_t e = tid; e < cnt; e += kMatchThreads) {
        const uint32_t key = s_key[e], p = s_spos[e];
        const uint8_t ch = s_ch[e];
        if (ch & 2u) {
            if (p + 1 < s_start[key] + s_count[key]) {
                const uint32_t e2 = s_sidx[p + 1];
                if ((s_ch[e2] ^ ch) != 6) err = 1;
                pair_row[s_idx[e]] = (int32_t)(s_idx[e2] - nb + pbase);
            } else {
                highs[(uint64_t)c * kMatchKeys + key] = first + e;
            }
        } else if (p == s_start[key]) {
            lows[(uint64_t)c * kMatchKeys + key] = first + e;
        }
    }
    for (int off = 32; off >= 1; off >>= 1) err |= __shfl_xor(err, off, 64);
    if (lane_id() == 0) s_red[w] = err;
    __syncthreads();
    if (tid == 0) {
        int32_t any = 0;
        for (int k = 0; k < kW; k++) any |= s_red[k];
        if (any) atomicOr(&uo[u].struct_err, 1u);
    }
}

// First chunk in (b, last] whose walk reaches depth d or below.  Whole
// 64- and 4096-chunk groups inside the range are skipped by their minima.
__device__ __forceinline__ uint32_t next_reaching(const int32_t *__restrict__ gmin,
                                                  const int32_t *__restrict__ gmin1,
                                                  const int32_t *__restrict__ gmin2, uint32_t b,
                                                  uint32_t last, int32_t d) {
    uint32_t c = b + 1;
    while (c <= last) {
        if ((c & 4095u) == 0 && c + 4095 <= last && gmin2[c >> 12] > d) {
            c += 4096;
            continue;
        }
        if ((c & 63u) == 0 && c + 63 <= last && gmin1[c >> 6] > d) {
            c += 64;
            continue;
        }
        if (gmin[c] <= d) return c;
        c++;
    }
    return last;  // not reached for a well-nested unit
}

// K4: openers left open by their chunk, one wave per chunk.
__global__ __launch_bounds__(64) void k_match_cross(
    const uint8_t *__restrict__ oc_char, const uint32_t *__restrict__ oc_idx,
    UnitOut *__restrict__ uo, uint32_t n_units, const ChunkStat *__restrict__ ck,
    const int32_t *__restrict__ gmin, const int32_t *__restrict__ gmin1,
    const int32_t *__restrict__ gmin2, const uint32_t *__restrict__ lows,
    const uint32_t *__restrict__ highs, int32_t *__restrict__ pair_row,
    const LaunchState *__restrict__ ls) {
    if (ls->overflow) return;
    const uint32_t c = blockIdx.x;
    if (c >= ls->n_chunks) return;
    const uint32_t u = unit_of_chunk(uo, n_units, c);
    if (!uo[u].fast) return;
    const uint32_t nch = (uo[u].m_end - uo[u].m_base + kMatchChunk - 1) / kMatchChunk;
    const uint32_t last = uo[u].chunk0 + nch - 1;
    const uint32_t nb = uo[u].n_base, pbase = uo[u].pair_base + 1u;
    const uint32_t mb = uo[u].m_base, me = uo[u].m_end;
    const ChunkStat s = ck[c];
    const int32_t dend = s.d0 + s.d;
    int32_t err = 0;
    for (int32_t d = gmin[c] + (int32_t)threadIdx.x; d < dend; d += 64) {
        const uint32_t c2 = next_reaching(gmin, gmin1, gmin2, c, last, d);
        const uint32_t m = highs[(uint64_t)c * kMatchKeys + d];
        const uint32_t m2 = lows[(uint64_t)c2 * kMatchKeys + d];
        // Both slots are written for a well-nested unit; the range check keeps
        // a stale slot from ever indexing outside the unit.
        if (m < mb || m >= me || m2 < mb || m2 >= me) {
            err = 1;
            continue;
        }
        if ((oc_char[m] ^ oc_char[m2]) != 6) err = 1;
        pair_row[oc_idx[m]] = (int32_t)(oc_idx[m2] - nb + pbase);
    }
    if (__ballot(err) && threadIdx.x == 0) atomicOr(&uo[u].struct_err, 1u);
}

hipError_t launch_match(hipStream_t s, const LaunchBufs &b, uint32_t n_units, uint32_t chunk_bound,
                        int32_t *pair_row) {
    if (chunk_bound == 0) return hipSuccess;
    hipLaunchKernelGGL(k_chunk_stats, dim3(chunk_bound), dim3(kMatchThreads), 0, s, b.oc_char,
                       b.uout, n_units, b.ck, b.ls);
    hipLaunchKernelGGL(k_chunk_scan, dim3(n_units), dim3(256), 0, s, b.uout, b.ck, b.gmin,
                       b.gmin1, b.gmin2, b.ls);
    hipLaunchKernelGGL(k_match_local, dim3(chunk_bound), dim3(kMatchThreads), 0, s, b.oc_char,
                       b.oc_idx, b.uout, n_units, b.ck, b.lows, b.highs, pair_row, b.ls);
    hipLaunchKernelGGL(k_match_cross, dim3(chunk_bound), dim3(64), 0, s, b.oc_char, b.oc_idx,
                       b.uout, n_units, b.ck, b.gmin, b.gmin1, b.gmin2, b.lows, b.highs,
                       pair_row, b.ls);
    return hipGetLastError();
}

}  // namespace cjx
