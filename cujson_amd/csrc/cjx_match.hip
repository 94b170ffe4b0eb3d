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
// The same two-level rule runs at two scales.  Brackets are cut into chunks
// of kMatchChunk, and a chunk into 16-bracket lanes.  A lane matches its own
// brackets with a stack and is left with some closers that reach new lows
// and some openers left open, at most one per depth.  An opener left open at
// depth d is closed by the first later lane (or chunk) whose walk reaches d,
// at that lane's (chunk's) first closer to reach d.  The search runs over the
// lanes' (chunks') minimum depths, skipping whole groups by their minima.
//
// Each chunk writes the pair-row cells of its rows (from its first bracket's
// row to the next chunk's) densely, -1 where no pair starts, through LDS
// tiles; openers closed in a later chunk are patched afterwards.
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

// Lists a unit that is not well nested for the sort path (k_sort_unit).
__device__ __forceinline__ void list_slow(LaunchState *ls, uint32_t *slow_units, uint32_t u) {
    slow_units[atomicAdd(&ls->n_slow, 1u)] = u;
}

__device__ __forceinline__ int32_t delta_of(uint8_t c) {
    // '[' 0x5B, '{' 0x7B open; ']' 0x5D, '}' 0x7D close (bit 1 tells them apart).
    return (c & 2u) ? 1 : -1;
}

__device__ __forceinline__ DepthAgg shfl_agg_up(const DepthAgg &a, int off) {
    DepthAgg o;
    o.d = __shfl_up(a.d, off, 64);
    o.mn = __shfl_up(a.mn, off, 64);
    o.mx = __shfl_up(a.mx, off, 64);
    return o;
}

// Running minimum / maximum over the lanes up to this one (DPP row shifts
// and broadcasts; lanes without a source keep their own value).
template <bool MAX>
__device__ __forceinline__ int32_t wave_incl_minmax(int32_t x) {
    auto step = [&](int32_t o) { x = MAX ? max(x, o) : min(x, o); };
    step((int32_t)dpp<kDppRowShr1>((uint32_t)x, (uint32_t)x));
    step((int32_t)dpp<kDppRowShr2>((uint32_t)x, (uint32_t)x));
    step((int32_t)dpp<kDppRowShr4>((uint32_t)x, (uint32_t)x));
    step((int32_t)dpp<kDppRowShr8>((uint32_t)x, (uint32_t)x));
    step((int32_t)dpp<kDppRowBcast15, 0xA>((uint32_t)x, (uint32_t)x));
    step((int32_t)dpp<kDppRowBcast31, 0xC>((uint32_t)x, (uint32_t)x));
    return x;
}

// Ordered composition of a DepthAgg over a wave (lane order = walk order),
// the total in lane 63: the lanes' start depths are a DPP prefix sum, and
// the minimum / maximum are running ones of start + mn / start + mx.
__device__ __forceinline__ DepthAgg wave_total_agg(const DepthAgg &a) {
    const int32_t incl = (int32_t)wave_incl_scan((uint32_t)a.d);
    const int32_t ex = incl - a.d;
    return DepthAgg{incl, wave_incl_minmax<false>(ex + a.mn), wave_incl_minmax<true>(ex + a.mx)};
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

// Lane t's 16 brackets [16t, 16t + 16) of a chunk whose brackets start at
// oc_char[first] (cnt of them): one 16-byte load where the chunk's bytes are
// 16-byte aligned and the lane's 16 are all inside it, byte loads otherwise
// (zero past the end).
__device__ __forceinline__ uint4 chunk_lane_bytes(const uint8_t *__restrict__ oc_char,
                                                  uint32_t first, uint32_t cnt, uint32_t t) {
    const uint32_t e0 = 16u * t;
    if ((first & 15u) == 0 && cnt == (uint32_t)kMatchChunk)  // wave-uniform: a full aligned chunk
        return *reinterpret_cast<const uint4 *>(oc_char + first + e0);
    if ((first & 15u) == 0 && e0 + 16u <= cnt)
        return *reinterpret_cast<const uint4 *>(oc_char + first + e0);
    uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (uint32_t k = 0; k < 16; k++)
        if (e0 + k < cnt) w[k >> 2] |= (uint32_t)oc_char[first + e0 + k] << (8 * (k & 3));
    return make_uint4(w[0], w[1], w[2], w[3]);
}

// Depth walk of a lane's 16 brackets when all 16 are there: one byte permute
// per word turns its brackets into +1 / -1 bytes (bit 1 tells openers from
// closers: selector byte 2 picks 0x01, selector byte 0 picks 0xFF), and the
// running minimum and maximum take two depths per step.
__device__ __forceinline__ DepthAgg lane_walk_full(uint4 v) {
    const uint32_t wd[4] = {v.x, v.y, v.z, v.w};
    int32_t d = 0, mn = 0, mx = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint32_t y = __builtin_amdgcn_perm(0u, 0x000100FFu, wd[q] & 0x02020202u);
        const int32_t d1 = d + (int32_t)(int8_t)y;
        const int32_t d2 = d1 + (int32_t)(int8_t)(y >> 8);
        const int32_t d3 = d2 + (int32_t)(int8_t)(y >> 16);
        const int32_t d4 = d3 + ((int32_t)y >> 24);
        mn = min(mn, min(d1, d2));
        mn = min(mn, min(d3, d4));
        mx = max(mx, max(d1, d2));
        mx = max(mx, max(d3, d4));
        d = d4;
    }
    return DepthAgg{d, mn, mx};
}

// K1: depth walk of every chunk.  Thread t walks brackets [16t, 16t + 16)
// of a chunk straight from registers; the next chunk's bytes are loaded
// before this one is walked.  The block's chunks are bid, bid + nblk, ...;
// lane i of every wave resolves the unit and bracket range of the block's
// chunk i up front (64 at a time), so the walk over them waits for no chain
// of dependent loads per chunk.
__device__ __forceinline__ void chunk_stats_blocks(uint32_t bid, uint32_t nblk,
                                                   const uint8_t *__restrict__ oc_char,
                                                   const UnitOut *__restrict__ uo,
                                                   const uint32_t *__restrict__ chunk_unit,
                                                   ChunkStat *__restrict__ ck,
                                                   const LaunchState *__restrict__ ls) {
    static_assert(kMatchChunk == 16 * kMatchThreads, "16 brackets per thread");
    __shared__ DepthAgg part[kMatchThreads / 64];
    __shared__ DepthAgg res[64];  // walks of the block's current 64 chunks
    if (ls->overflow) return;
    const uint32_t n_chunks = ls->n_chunks;
    if (bid >= n_chunks) return;
    const uint32_t t = threadIdx.x, lane = lane_id();
    const uint32_t e0 = 16u * t;
    const uint32_t nper = (n_chunks - bid + nblk - 1) / nblk;  // chunks of this block
    for (uint32_t g0 = 0; g0 < nper; g0 += 64) {
        // Lane i: chunk g0 + i of the block.
        uint32_t fv = 0, cv_n = 0;
        if (g0 + lane < nper) {
            const uint32_t c = bid + (g0 + lane) * nblk;
            const uint32_t u = chunk_unit[c];
            fv = uo[u].m_base + (c - uo[u].chunk0) * kMatchChunk;
            cv_n = uo[u].m_end > fv ? min((uint32_t)kMatchChunk, uo[u].m_end - fv) : 0u;
        }
        const uint32_t gn = min(64u, nper - g0);
        // Wave composition, then the block's: chunk i's walk into res[i].
        auto reduce = [&](DepthAgg a, uint32_t i) {
            a = wave_total_agg(a);
            if (lane == 63) part[t >> 6] = a;
            __syncthreads();
            if (t == 0) {
                DepthAgg r = part[0];
                for (int k = 1; k < kMatchThreads / 64; k++) r = agg_compose(r, part[k]);
                res[i] = r;
            }
            __syncthreads();
        };
        const bool all_full = __builtin_amdgcn_ballot_w64(
                                  lane < gn && ((fv & 15u) != 0 || cv_n != (uint32_t)kMatchChunk)) == 0;
        if (all_full) {
            // Every chunk of the group full and 16-byte aligned: one 16-byte
            // load per lane and chunk with no branches around it, so the walk
            // waits only for its own chunk (the last step loads its chunk again).
            uint4 cur = *reinterpret_cast<const uint4 *>(oc_char + __builtin_amdgcn_readlane(fv, 0) + e0);
            for (uint32_t i = 0; i < gn; i++) {
                const uint32_t in = min(i + 1, gn - 1);
                const uint4 nxt =
                    *reinterpret_cast<const uint4 *>(oc_char + __builtin_amdgcn_readlane(fv, (int)in) + e0);
                reduce(lane_walk_full(cur), i);
                cur = nxt;
            }
        } else {
        uint32_t first = __builtin_amdgcn_readlane(fv, 0), cnt = __builtin_amdgcn_readlane(cv_n, 0);
        uint4 cur = chunk_lane_bytes(oc_char, first, cnt, t);
        for (uint32_t i = 0; i < gn; i++) {
            uint32_t cnt_n = 0;
            uint4 nxt = make_uint4(0u, 0u, 0u, 0u);
            if (i + 1 < gn) {
                const uint32_t first_n = __builtin_amdgcn_readlane(fv, (int)(i + 1));
                cnt_n = __builtin_amdgcn_readlane(cv_n, (int)(i + 1));
                nxt = chunk_lane_bytes(oc_char, first_n, cnt_n, t);
            }
            DepthAgg a{0, 0, 0};
            if (cnt == (uint32_t)kMatchChunk) {
                a = lane_walk_full(cur);
            } else {
                const uint32_t cv[4] = {cur.x, cur.y, cur.z, cur.w};
#pragma unroll
                for (uint32_t k = 0; k < 16; k++) {
                    if (e0 + k < cnt) {
                        a.d += delta_of((uint8_t)(cv[k >> 2] >> (8 * (k & 3))));
                        a.mn = min(a.mn, a.d);
                        a.mx = max(a.mx, a.d);
                    }
                }
            }
            reduce(a, i);
            cur = nxt;
            cnt = cnt_n;
        }
        }
        // The group's results are stored after its walks: a store among the
        // walks' loads would make the next chunk's walk wait for every load
        // in flight (stores and loads share one counter and may complete out
        // of order), the prefetched chunk included.
        if (t < gn) {
            const DepthAgg r = res[t];
            ck[bid + (g0 + t) * nblk] = ChunkStat{r.d, r.mn, r.mx, 0};
        }
    }
}

// K2, one block per unit, in three wave-parallel passes over the unit's
// 64-chunk groups: (a) each group's depth walk, (b) an ordered scan of the
// groups giving their start depths and minima, the 4096-chunk minima and
// whether the unit is well nested, (c) each chunk's start depth and minimum.
constexpr int kScanThreads = 1024;
constexpr int kScanBatch = 2;  // 64-chunk groups per wave step of the scan passes
constexpr int kScanSlots = 12;  // groups per wave of the resident pass (units of <= 12288 chunks)

// Resident pass of K2 for a unit of at most kScanThreads / 64 * kScanSlots
// groups.  Wave w holds groups w, w + kW, ... (lane l: chunk 64 j + l), all
// loaded at once.  Within a group the chunks' start depths are a DPP prefix
// sum, and the group's walk is (sum of d, minimum and maximum over chunks of
// start + mn / mx); the group scan is one pass over threads (thread j: group
// j); chunk start depths and minima then come from the registers.  One
// round trip to memory in all, and no LDS shuffles.
__device__ __forceinline__ void chunk_scan_resident(uint32_t u, uint32_t c0, uint32_t nch, uint32_t g0,
                                                    uint32_t h0, UnitOut *__restrict__ uo,
                                                    ChunkStat *__restrict__ ck,
                                                    int32_t *__restrict__ gmin,
                                                    int32_t *__restrict__ gmin1,
                                                    int32_t *__restrict__ gmin2,
                                                    LaunchState *__restrict__ ls,
                                                    uint32_t *__restrict__ slow_units) {
    constexpr int kW = kScanThreads / 64;
    __shared__ DepthAgg s_g[kW * kScanSlots];
    __shared__ DepthAgg s_part[kW];
    __shared__ int32_t s_gd0[kW * kScanSlots];
    const uint32_t tid = threadIdx.x, w = tid >> 6, lane = lane_id();
    const uint32_t G = (nch + 63) / 64;
    ChunkStat st[kScanSlots];
#pragma unroll
    for (int i = 0; i < kScanSlots; i++) {
        const uint32_t j = w + kW * i, x = 64 * j + lane;
        st[i] = (j < G && x < nch) ? ck[c0 + x] : ChunkStat{0, 0, 0, 0};
    }
    int32_t ex[kScanSlots], lo[kScanSlots];  // start depth in the group; start + mn
#pragma unroll
    for (int i = 0; i < kScanSlots; i++) {
        const uint32_t j = w + kW * i;
        if (j < G) {
            const DepthAgg g = wave_total_agg(DepthAgg{st[i].d, st[i].mn, st[i].mx});
            ex[i] = (int32_t)g.d - st[i].d;  // lane's running sum less its own
            lo[i] = ex[i] + st[i].mn;
            if (lane == 63) s_g[j] = g;
        }
    }
    __syncthreads();
    // Ordered scan of the groups, thread j holding group j.
    const DepthAgg a = tid < G ? s_g[tid] : DepthAgg{0, 0, 0};
    const DepthAgg incl = wave_scan_agg(a);
    if (lane == 63) s_part[w] = incl;
    __syncthreads();
    DepthAgg pre{0, 0, 0};
    for (uint32_t k = 0; k < w; k++) pre = agg_compose(pre, s_part[k]);
    const DepthAgg up = shfl_agg_up(incl, 1);
    const DepthAgg excl = lane ? agg_compose(pre, up) : pre;
    int32_t m1 = INT32_MAX;
    if (tid < G) {
        s_gd0[tid] = excl.d;
        m1 = excl.d + a.mn;
        gmin1[g0 + tid] = m1;
    }
    if (tid == kScanThreads - 1) {
        const DepthAgg t = agg_compose(pre, incl);
        const bool fast = t.d == 0 && t.mn >= 0 && t.mx <= kMatchMaxDepth;
        uo[u].fast = fast ? 1u : 0u;
        if (!fast) list_slow(ls, slow_units, u);
    }
    // 4096-chunk minima: wave h holds groups [64 h, 64 h + 64).
    for (int off = 32; off >= 1; off >>= 1) m1 = min(m1, __shfl_xor(m1, off, 64));
    if (lane == 0 && w < (nch + 4095) / 4096) gmin2[h0 + w] = m1;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kScanSlots; i++) {
        const uint32_t j = w + kW * i, x = 64 * j + lane;
        if (j < G && x < nch) {
            const int32_t gd = s_gd0[j];
            ck[c0 + x].d0 = gd + ex[i];
            gmin[c0 + x] = gd + lo[i];
        }
    }
}

__device__ __forceinline__ void chunk_scan_blocks(
    uint32_t bid, uint32_t nblk, uint32_t n_units, UnitOut *__restrict__ uo, ChunkStat *__restrict__ ck,
    int32_t *__restrict__ gmin,
    int32_t *__restrict__ gmin1, int32_t *__restrict__ gmin2, DepthAgg *__restrict__ gagg,
    int32_t *__restrict__ gd0, const uint32_t *__restrict__ big_units,
    LaunchState *__restrict__ ls, uint32_t *__restrict__ slow_units) {
    constexpr int kW = kScanThreads / 64;
    __shared__ DepthAgg part[kW];
    // The flag, the count of big units and this block's first entry of the
    // list are read at once (the list has room for n_units entries and
    // bid < n_units; the entry is used only when bid < n_big).
    const uint32_t overflow = ls->overflow, n_big = ls->n_big;
    const uint32_t u_first = min(big_units[bid], n_units - 1);
    asm volatile("" ::"s"(u_first));  // keeps that read here, beside the other two
    if (overflow) return;
    for (uint32_t bi = bid; bi < n_big; bi += nblk) {
    __syncthreads();  // part[] of the previous unit no longer read
    const uint32_t u = bi == bid ? u_first : big_units[bi], tid = threadIdx.x, w = tid >> 6, lane = lane_id();
    const uint32_t c0 = uo[u].chunk0, nch = uo[u].nchunks;
    const uint32_t g0 = uo[u].grp0, h0 = uo[u].sgrp0;
    const uint32_t G = (nch + 63) / 64, H = (nch + 4095) / 4096;
    if (G <= (uint32_t)(kW * kScanSlots)) {
        chunk_scan_resident(u, c0, nch, g0, h0, uo, ck, gmin, gmin1, gmin2, ls, slow_units);
        continue;
    }
    // (a) group walks: kScanBatch groups per wave step, their loads issued
    // together.
    for (uint32_t j0 = w; j0 < G; j0 += kScanBatch * kW) {
        DepthAgg v[kScanBatch];
#pragma unroll
        for (int i = 0; i < kScanBatch; i++) {
            const uint32_t j = j0 + i * kW, x = 64 * j + lane;
            v[i] = DepthAgg{0, 0, 0};
            if (j < G && x < nch) {
                const ChunkStat st = ck[c0 + x];
                v[i] = DepthAgg{st.d, st.mn, st.mx};
            }
        }
#pragma unroll
        for (int i = 0; i < kScanBatch; i++) {
            const uint32_t j = j0 + i * kW;
            const DepthAgg r = wave_total_agg(v[i]);
            if (lane == 63 && j < G) gagg[g0 + j] = r;
        }
    }
    __syncthreads();
    // (b) ordered scan over the groups, up to 8 per thread.
    const uint32_t per = (G + kScanThreads - 1) / kScanThreads;
    const uint32_t a0 = min(G, tid * per), a1 = min(G, (tid + 1) * per);
    DepthAgg a{0, 0, 0};
    for (uint32_t jb = a0; jb < a1; jb += 8) {
        DepthAgg v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = gagg[g0 + (jb + k < a1 ? jb + k : a0)];
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (jb + k < a1) a = agg_compose(a, v[k]);
    }
    const DepthAgg incl = wave_scan_agg(a);
    if (lane == 63) part[w] = incl;
    __syncthreads();
    DepthAgg pre{0, 0, 0};
    for (uint32_t k = 0; k < w; k++) pre = agg_compose(pre, part[k]);
    const DepthAgg up = shfl_agg_up(incl, 1);
    const DepthAgg excl = lane ? agg_compose(pre, up) : pre;
    int32_t depth = excl.d;
    for (uint32_t jb = a0; jb < a1; jb += 8) {
        DepthAgg v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = gagg[g0 + (jb + k < a1 ? jb + k : a0)];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            if (jb + k < a1) {
                gd0[g0 + jb + k] = depth;
                gmin1[g0 + jb + k] = depth + v[k].mn;
                depth += v[k].d;
            }
        }
    }
    if (tid == kScanThreads - 1) {
        const DepthAgg t = agg_compose(pre, incl);
        const bool fast = t.d == 0 && t.mn >= 0 && t.mx <= kMatchMaxDepth;
        uo[u].fast = fast ? 1u : 0u;
        if (!fast) list_slow(ls, slow_units, u);
    }
    __syncthreads();
    // 4096-chunk minima.
    for (uint32_t h = w; h < H; h += kW) {
        int32_t m = (64 * h + lane < G) ? gmin1[g0 + 64 * h + lane] : INT32_MAX;
        for (int off = 32; off >= 1; off >>= 1) m = min(m, __shfl_xor(m, off, 64));
        if (lane == 0) gmin2[h0 + h] = m;
    }
    // (c) chunk start depths and minima, kScanBatch groups per wave step.
    for (uint32_t j0 = w; j0 < G; j0 += kScanBatch * kW) {
        int32_t bd[kScanBatch], sd[kScanBatch], smn[kScanBatch];
#pragma unroll
        for (int i = 0; i < kScanBatch; i++) {
            const uint32_t j = j0 + i * kW, x = 64 * j + lane;
            bd[i] = sd[i] = smn[i] = 0;
            if (j < G) {
                bd[i] = gd0[g0 + j];
                if (x < nch) {
                    const ChunkStat st = ck[c0 + x];
                    sd[i] = st.d;
                    smn[i] = st.mn;
                }
            }
        }
#pragma unroll
        for (int i = 0; i < kScanBatch; i++) {
            const uint32_t j = j0 + i * kW, x = 64 * j + lane;
            const int32_t in = (int32_t)wave_incl_scan((uint32_t)sd[i]);
            if (j < G && x < nch) {
                const int32_t d0 = bd[i] + in - sd[i];
                ck[c0 + x].d0 = d0;
                gmin[c0 + x] = d0 + smn[i];
            }
        }
    }
    }
}

// K2 for units of at most 64 chunks, one wave per unit: the ordered scan of
// the unit's chunk walks, chunk start depths and minima, the (single) group
// minima, and the well-nested test.
__device__ __forceinline__ void chunk_scan_small_block(uint32_t bid, UnitOut *__restrict__ uo,
                                                          uint32_t n_units,
                                                          ChunkStat *__restrict__ ck,
                                                          int32_t *__restrict__ gmin,
                                                          int32_t *__restrict__ gmin1,
                                                          int32_t *__restrict__ gmin2,
                                                          LaunchState *__restrict__ ls,
                                                          uint32_t *__restrict__ slow_units) {
    if (ls->overflow) return;
    const uint32_t lane = lane_id();
    const uint32_t u = bid * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (u >= n_units) return;
    const uint32_t nch = uo[u].nchunks;
    if (nch > 64) return;  // chunk_scan_blocks
    if (nch == 0) return;  // match_small_block decides these units
    const uint32_t c0 = uo[u].chunk0;
    ChunkStat st{0, 0, 0, 0};
    if (lane < nch) st = ck[c0 + lane];
    const DepthAgg a{st.d, st.mn, st.mx};
    const DepthAgg incl = wave_scan_agg(a);
    const DepthAgg t = {__shfl(incl.d, 63, 64), __shfl(incl.mn, 63, 64), __shfl(incl.mx, 63, 64)};
    const int32_t d0 = incl.d - st.d;
    int32_t m = INT32_MAX;
    if (lane < nch) {
        ck[c0 + lane].d0 = d0;
        gmin[c0 + lane] = d0 + st.mn;
        m = d0 + st.mn;
    }
    for (int off = 32; off >= 1; off >>= 1) m = min(m, __shfl_xor(m, off, 64));
    if (lane == 0) {
        gmin1[uo[u].grp0] = m;
        gmin2[uo[u].sgrp0] = m;
        const bool fast = t.d == 0 && t.mn >= 0 && t.mx <= kMatchMaxDepth;
        uo[u].fast = fast ? 1u : 0u;
        if (!fast) list_slow(ls, slow_units, u);
    }
}

// K3: matching inside one chunk.  Thread t owns the 16 brackets
// [16t, 16t + 16) of the chunk and matches them with a stack of 4-bit lane
// indices packed in a 64-bit register (no LDS round trips).  What is left
// is, per lane, closers that reach the lane's successive new lows (s_low) and
// openers left open at the lane minimum, minimum + 1, ... (the stack).  Each
// left-open opener is closed by the first later lane whose minimum reaches
// its depth; one that no lane of the chunk closes goes to highs, and a low
// closer that no earlier lane of the chunk opened goes to lows.  Partners are
// kept in LDS (s_part) and the chunk's pair-row cells are then written as
// dense tiles, each wave over the rows of its quarter of the brackets.
constexpr int kLaneBrackets = kMatchChunk / kMatchThreads;  // 16
constexpr int kRowTile = 4096;
constexpr uint32_t kWaveElems = kMatchChunk / (kMatchThreads / 64);  // a wave's brackets
constexpr uint32_t kWaveTile = kRowTile / (kMatchThreads / 64);      // a wave's tile rows
static_assert(kWaveElems == 64 * kLaneBrackets, "a wave's elements: 64 lanes x 16");
// Element of register k of lane l in wave w: 1024 w + 256 (k / 4) + 4 l + k % 4
// (a contiguous range per wave, four consecutive elements per 16-byte access).
__device__ __forceinline__ uint32_t elem_of(uint32_t w, uint32_t l, int k) {
    return w * kWaveElems + 256u * (uint32_t)(k >> 2) + 4u * l + (uint32_t)(k & 3);
}
constexpr uint16_t kNoPart = 0xFFFF;

// LDS writes of the wave are visible to all its lanes after this.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
#ifndef CJX_MATCH_WPE
#define CJX_MATCH_WPE 5
#endif
__global__ __attribute__((amdgpu_waves_per_eu(CJX_MATCH_WPE, CJX_MATCH_WPE)))
__launch_bounds__(kMatchThreads) void k_match_local(
    const uint8_t *__restrict__ oc_char, const uint32_t *__restrict__ oc_idx,
    UnitOut *__restrict__ uo, const uint32_t *__restrict__ chunk_unit,
    const ChunkStat *__restrict__ ck, uint32_t *__restrict__ lows, uint32_t *__restrict__ highs,
    int32_t *__restrict__ pair_row, LaunchState *__restrict__ ls,
    const UnitOut *__restrict__ uor) {
    constexpr int kW = kMatchThreads / 64;
    constexpr int kGroups = kMatchThreads / 16;
    // The chunk's bracket rows stay in registers (vi); for the partners'
    // rows they are copied into the matching state's LDS once matching is
    // done, so no LDS of their own limits the blocks per CU.
    __shared__ __align__(16) uint8_t s_ch[kMatchChunk];
    __shared__ __align__(16) uint16_t s_part[kMatchChunk];
    // Low closers during matching; the row tile afterwards (16-byte
    // aligned: the tile is filled and read as int4).
    __shared__ __align__(16) union {
        struct {
            uint16_t low[kLaneBrackets][kMatchThreads];
        } m;
        int32_t row[kRowTile];
    } s_u;
    __shared__ int32_t s_lmin[kMatchThreads];  // lane minimum depth (start included)
    __shared__ int32_t s_lst[kMatchThreads];   // lane start depth
    __shared__ int32_t s_gmin[kGroups];        // minimum over 16 lanes
    __shared__ int32_t s_wsum[kW], s_wmin[kW];
    __shared__ uint32_t s_wrow[kW];  // row of each wave's first bracket
    __shared__ uint32_t s_next;

    if (ls->overflow) return;
    const uint32_t tid = threadIdx.x, w = tid >> 6, lane = lane_id();
    const uint32_t n_chunks = ls->n_chunks;
    // Chunks are taken from a ticket counter, so blocks that start late or
    // meet cheap chunks take more of them.  The next ticket is drawn when a
    // chunk starts and handed over once the chunk's loads are in, so its
    // round trip overlaps theirs.  (Fetch-and-increment: the compiler gives
    // it no wave-wide rewrite that would wait for the result at once.)
    // With several chunks per block, blocks that start together stay in
    // step: their tile writes come in bursts that the memory serves one
    // after the other while the matching leaves it idle.  Blocks start in
    // four groups about 7 us apart (s_sleep 127 = 8128 cycles, twice per
    // step), which keeps the writes of some blocks under the matching of
    // others (k_match_local 223 -> 215 us per 2 GB).  Small launches start
    // at once.
    if (n_chunks >= 4u * gridDim.x)
        for (uint32_t k = 0; k < 2u * (blockIdx.x & 3u); k++) __builtin_amdgcn_s_sleep(127);
    if (tid == 0) s_next = atomicInc(&ls->chunk_ticket, 0xFFFFFFFFu);
    __syncthreads();
    for (uint32_t c = s_next; c < n_chunks;) {
        uint32_t next = 0;
        if (tid == 0) next = atomicInc(&ls->chunk_ticket, 0xFFFFFFFFu);
        const uint32_t u = chunk_unit[c];
        const UnitOut U = uor[u];  // uniform: scalar loads, not queued behind vector stores
        if (!U.fast) {
            __syncthreads();  // every thread has read this chunk's ticket
            if (tid == 0) s_next = next;
        } else {
            const uint32_t xc = c - U.chunk0;
            const uint32_t first = U.m_base + xc * kMatchChunk;
            const uint32_t cnt = U.m_end > first ? min((uint32_t)kMatchChunk, U.m_end - first) : 0u;
            const uint32_t nb = U.n_base, pbase = U.pair_base + 1u;
            const bool last_chunk = xc + 1 == U.nchunks;
            const uint32_t r1 = last_chunk ? U.n_end : oc_idx[first + kMatchChunk];
            const int32_t d0 = ck[c].d0;
            // All loads in flight before the first LDS store: the thread's 16
            // characters [16 tid, 16 tid + 16) in one load, and the rows of
            // the wave's elements, wave w holding [1024 w, 1024 w + 1024) for
            // the pair-row tiles at the end (register k of lane l: element
            // elem_of(w, l, k), four consecutive rows per 16-byte load).
            const uint4 cw = chunk_lane_bytes(oc_char, first, cnt, tid);
            uint32_t vi[kLaneBrackets];
            if ((first & 3u) == 0 && (w + 1) * kWaveElems <= cnt) {
#pragma unroll
                for (int q = 0; q < kLaneBrackets / 4; q++) {
                    const uint4 r = *reinterpret_cast<const uint4 *>(oc_idx + first + elem_of(w, lane, 4 * q));
                    vi[4 * q] = r.x;
                    vi[4 * q + 1] = r.y;
                    vi[4 * q + 2] = r.z;
                    vi[4 * q + 3] = r.w;
                }
            } else {
#pragma unroll
                for (int k = 0; k < kLaneBrackets; k++) {
                    const uint32_t e = elem_of(w, lane, k);
                    vi[k] = e < cnt ? oc_idx[first + e] : 0;
                }
            }
            __syncthreads();  // LDS of the previous chunk is no longer read
            if (tid == 0) s_next = next;
            *reinterpret_cast<uint4 *>(&s_ch[16u * tid]) = cw;
            {
                uint4 *sp4 = reinterpret_cast<uint4 *>(&s_part[16u * tid]);
                const uint32_t none = 0xFFFFFFFFu;  // two kNoPart per word
                sp4[0] = make_uint4(none, none, none, none);
                sp4[1] = make_uint4(none, none, none, none);
            }
            if (lane == 0) s_wrow[w] = vi[0];  // the wave's first bracket row
            __syncthreads();
            const uint32_t r0 = xc == 0 ? nb : s_wrow[0];  // row of the chunk's first bracket

            const uint32_t e0 = tid * kLaneBrackets;
            const uint32_t ne = cnt > e0 ? min((uint32_t)kLaneBrackets, cnt - e0) : 0u;
            const uint32_t cv[4] = {cw.x, cw.y, cw.z, cw.w};
            int32_t mine = 0;
#pragma unroll
            for (int k = 0; k < kLaneBrackets; k++)
                if ((uint32_t)k < ne) mine += delta_of((uint8_t)(cv[k >> 2] >> (8 * (k & 3))));
            // Lane start depth: exclusive scan of the lanes' net changes.
            const int32_t incl = (int32_t)wave_incl_scan((uint32_t)mine);
            if (lane == 63) s_wsum[w] = incl;
            __syncthreads();
            int32_t ds = d0 + incl - mine;
            for (uint32_t k = 0; k < w; k++) ds += s_wsum[k];

            // Stack walk of the lane, in registers: the open brackets' lane
            // indices, 4 bits each, packed into a 64-bit stack (top in the low
            // bits), and a mask of the curly ones; partners and low closers go
            // to LDS, which the walk never reads back.  An opener and a closer
            // pair up exactly when both are curly or both square (byte bit 5).
            uint64_t stk = 0;
            uint32_t sp = 0, nlow = 0, curly = 0;
            int32_t err = 0;
#pragma unroll
            for (int k = 0; k < kLaneBrackets; k++) {
                if ((uint32_t)k >= ne) break;
                const uint32_t ch = cv[k >> 2] >> (8 * (k & 3));
                const uint32_t e = e0 + k;
                curly |= ((ch >> 5) & 1u) << k;
                if (ch & 2u) {
                    stk = (stk << 4) | (uint64_t)k;
                    sp++;
                } else if (sp) {
                    sp--;
                    const uint32_t i = (uint32_t)stk & 15u;
                    stk >>= 4;
                    if ((curly >> i ^ curly >> k) & 1u) err = 1;
                    s_part[e0 + i] = (uint16_t)e;
                } else {
                    s_u.m.low[nlow][tid] = (uint16_t)e;
                    nlow++;
                }
            }
            const int32_t dm = ds - (int32_t)nlow;
            s_lmin[tid] = ne ? dm : INT32_MAX;
            s_lst[tid] = ds;
            // Lowest depth before this lane within the chunk (chunk start included).
            int32_t pm = wave_incl_minmax<false>(ne ? dm : INT32_MAX);
            if (lane == 63) s_wmin[w] = pm;
            pm = (int32_t)dpp<kDppWaveShr1>((uint32_t)INT32_MAX, (uint32_t)pm);  // lanes before this one
            __syncthreads();
            pm = min(pm, d0);
            for (uint32_t k = 0; k < w; k++) pm = min(pm, s_wmin[k]);
            if (tid < kGroups) {
                int32_t g = INT32_MAX;
                for (int k = 0; k < 16; k++) g = min(g, s_lmin[16 * tid + k]);
                s_gmin[tid] = g;
            }
            __syncthreads();

            // Openers left open by the lane, the deepest first.  The first later
            // lane that reaches depth d never comes before the first that
            // reaches d + 1, so one forward walk over the lanes serves them all
            // (a nested run of openers costs one walk, not one per opener).
            uint32_t L = tid + 1;
            for (int32_t j = (int32_t)sp - 1; j >= 0; j--) {
                const int32_t d = dm + j;
                const uint32_t i = e0 + ((uint32_t)stk & 15u);
                stk >>= 4;
                while (L < kMatchThreads) {
                    if ((L & 15u) == 0 && s_gmin[L >> 4] > d) {
                        L += 16;
                        continue;
                    }
                    if (s_lmin[L] <= d) break;
                    L++;
                }
                if (L < kMatchThreads) {
                    const int32_t k = s_lst[L] - 1 - d;
                    const uint32_t e2 = s_u.m.low[k & (kLaneBrackets - 1)][L];
                    if ((s_ch[i] ^ s_ch[e2]) != 6) err = 1;
                    s_part[i] = (uint16_t)e2;
                } else if ((uint32_t)d < (uint32_t)kMatchKeys) {
                    highs[(uint64_t)c * kMatchKeys + d] = first + i;
                }
            }
            // Low closers that no earlier lane of the chunk opened.
            for (uint32_t k = 0; k < nlow; k++) {
                const int32_t d = ds - 1 - (int32_t)k;
                if (d < pm && (uint32_t)d < (uint32_t)kMatchKeys)
                    lows[(uint64_t)c * kMatchKeys + d] = first + s_u.m.low[k][tid];
            }
            if (err) atomicOr(&uo[u].struct_err, 1u);

            // Dense pair-row tiles over the chunk's rows [r0, r1): wave w
            // writes the rows from its first bracket's row to the next wave's
            // (from r0 for wave 0, to r1 for the last wave with brackets) in
            // tiles of 1024 rows of its own, so the tiles need no workgroup
            // barrier.  The lanes keep the rows and values of their openers in
            // registers and drop those that fall in the current tile.
            __syncthreads();  // s_part complete, the matching state no longer read
            // The chunk's bracket rows go to LDS over the matching state, for
            // the partners' rows.
#pragma unroll
            for (int q = 0; q < kLaneBrackets / 4; q++)
                *reinterpret_cast<int4 *>(&s_u.row[elem_of(w, lane, 4 * q)]) =
                    make_int4((int32_t)vi[4 * q], (int32_t)vi[4 * q + 1], (int32_t)vi[4 * q + 2],
                              (int32_t)vi[4 * q + 3]);
            __syncthreads();
            int32_t pv[kLaneBrackets];
#pragma unroll
            for (int k = 0; k < kLaneBrackets; k++) {
                const uint32_t e = elem_of(w, lane, k);
                const uint32_t p = s_part[e];
                pv[k] = p != kNoPart ? (int32_t)((uint32_t)s_u.row[p] - nb + pbase) : -1;
            }
            __syncthreads();  // the rows are read; the tiles reuse their space
            const uint32_t wbeg = w * kWaveElems;
            const uint32_t wr0 = w == 0 ? r0 : (wbeg < cnt ? s_wrow[w] : r1);
            const uint32_t wr1 = wbeg + kWaveElems < cnt ? s_wrow[w + 1] : r1;
            // Tiles start at a multiple of 4 rows so that the fill and the store
            // move 16 bytes per access; rows before wr0 or from wr1 on belong to
            // the neighbouring waves or chunks and are skipped (edge groups are
            // stored by row).  (Filling the cells in HBM and storing the values
            // over them instead measured slower: 0.51 against 0.35 ms of phase 3
            // per 2 GB.)
            const bool vec = (reinterpret_cast<uintptr_t>(pair_row) & 15u) == 0;
            int32_t *const tile = s_u.row + w * kWaveTile;
            int4 *const tw = reinterpret_cast<int4 *>(tile);
            for (uint32_t t0 = vec ? (wr0 & ~3u) : wr0; t0 < wr1; t0 += kWaveTile) {
                wave_lds_sync();  // the previous tile is no longer read
#pragma unroll
                for (uint32_t i = lane; i < (uint32_t)kWaveTile / 4; i += 64)
                    tw[i] = make_int4(-1, -1, -1, -1);
                wave_lds_sync();
#pragma unroll
                for (int k = 0; k < kLaneBrackets; k++) {
                    const uint32_t off = vi[k] - t0;
                    if (pv[k] >= 0 && off < (uint32_t)kWaveTile) tile[off] = pv[k];
                }
                wave_lds_sync();
                const uint32_t hi = min(wr1, t0 + (uint32_t)kWaveTile);
                if (vec) {
                    for (uint32_t g = lane; g < (uint32_t)kWaveTile / 4; g += 64) {
                        const uint32_t a = t0 + 4u * g;
                        if (a >= hi) break;
                        if (a >= wr0 && a + 4u <= hi) {
                            *reinterpret_cast<int4 *>(pair_row + a) = tw[g];
                        } else {
#pragma unroll
                            for (uint32_t j = 0; j < 4; j++)
                                if (a + j >= wr0 && a + j < hi) pair_row[a + j] = tile[4u * g + j];
                        }
                    }
                } else {
                    for (uint32_t i = lane; t0 + i < hi; i += 64) pair_row[t0 + i] = tile[i];
                }
            }
        }
        __syncthreads();  // the next ticket is in s_next
        c = s_next;
    }
}

// Units with at most kSmallBrackets (1024) brackets, one wave per unit.
// Lane l of group g holds bracket 64g + l.  The depth walk is a wave scan
// per group with a carry; a unit whose walk leaves [0, 256] or does not end
// at 0 is not well nested and keeps fast = 0 (the sort path takes it).
// Otherwise the keys are exact depths and an opener's closer is the next
// bracket of its key.  The groups are visited from the last to the first:
// within a group, the lanes of one key are found with eight ballots (one per
// key bit); the first later bracket of a key outside the group is kept per
// key in LDS (its row and character), updated by each group's first lane of
// that key.  The unit's pair-row cells are filled with -1 (16-byte stores)
// and the openers' values are stored over them.
constexpr int kWaveGroups = kSmallBrackets / 64;
__device__ __forceinline__ void match_small_block(uint32_t bid, const uint8_t *__restrict__ oc_char,
                                                     const uint32_t *__restrict__ oc_idx,
                                                     UnitOut *__restrict__ uo, uint32_t n_units,
                                                     int32_t *__restrict__ pair_row,
                                                     LaunchState *__restrict__ ls,
                                                     uint32_t *__restrict__ slow_units) {
    __shared__ uint32_t s_nrow[4][kMatchKeys];  // per key: row of the next bracket after the group
    __shared__ uint8_t s_nch[4][kMatchKeys];    // and its character
    if (ls->overflow) return;
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    const uint32_t u = bid * 4 + w;
    if (u >= n_units) return;
    const uint32_t mb = uo[u].m_base, M = uo[u].m_end - mb;
    if (M > (uint32_t)kSmallBrackets) return;  // the chunk kernels
    const uint32_t nb = uo[u].n_base, ne = uo[u].n_end, pbase = uo[u].pair_base + 1u;
    const uint32_t G = (M + 63) / 64;
    uint32_t ch[kWaveGroups], row[kWaveGroups], key[kWaveGroups];
    int32_t carry = 0;
    bool bad = false;
#pragma unroll
    for (int g = 0; g < kWaveGroups; g++) {
        ch[g] = row[g] = key[g] = 0;
        if ((uint32_t)g >= G) continue;
        const uint32_t e = 64u * g + lane;
        const bool has = e < M;
        ch[g] = has ? oc_char[mb + e] : 0u;
        row[g] = has ? oc_idx[mb + e] : 0u;
        const int32_t d = has ? delta_of((uint8_t)ch[g]) : 0;
        const int32_t incl = carry + (int32_t)wave_incl_scan((uint32_t)d);  // depth after it
        carry = __shfl(incl, 63, 64);
        bad = bad || (has && (incl < 0 || incl > kMatchMaxDepth));
        key[g] = (uint32_t)(d > 0 ? incl - 1 : incl) & 0xFFu;  // exact depth, in [0, 255]
    }
    if (__ballot(bad) || carry != 0) {  // not well nested
        if (lane == 0) list_slow(ls, slow_units, u);
        return;
    }
    for (uint32_t k = lane; k < (uint32_t)kMatchKeys; k += 64) s_nrow[w][k] = 0;
    int32_t pval[kWaveGroups];
    bool err = false;
#pragma unroll
    for (int g = kWaveGroups - 1; g >= 0; g--) {
        pval[g] = -1;
        if ((uint32_t)g >= G) continue;
        const uint32_t e = 64u * g + lane;
        const bool has = e < M;
        uint64_t peers = __ballot(has);
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
            const uint32_t bv = (key[g] >> bit) & 1u;
            const uint64_t bm = __ballot(bv);
            peers &= bv ? bm : ~bm;
        }
        const uint64_t after = lane == 63 ? 0ull : peers & ~((2ull << lane) - 1ull);
        const uint32_t partner = after ? (uint32_t)__ffsll((long long)after) - 1u : lane;
        uint32_t c2 = (uint32_t)__shfl((int)ch[g], (int)partner, 64);
        uint32_t row2 = (uint32_t)__shfl((int)row[g], (int)partner, 64);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (!after) {  // the next bracket of the key lies in a later group
            c2 = s_nch[w][key[g]];
            row2 = s_nrow[w][key[g]];
        }
        const bool opener = has && (ch[g] & 2u);
        if (opener) {
            pval[g] = (int32_t)(row2 - nb + pbase);
            if ((ch[g] ^ c2) != 6u) err = true;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // the group's first lane of each key is the next bracket of that key
        // for the groups before it
        if (has && (peers & ((1ull << lane) - 1ull)) == 0) {
            s_nrow[w][key[g]] = row[g];
            s_nch[w][key[g]] = (uint8_t)ch[g];
        }
    }
    // Pair-row cells [nb, ne): -1, then the openers' values.
    const bool vec = (reinterpret_cast<uintptr_t>(pair_row) & 15u) == 0;
    const uint32_t a0 = vec ? min(ne, (nb + 3u) & ~3u) : ne;
    const uint32_t a1 = vec ? max(a0, ne & ~3u) : a0;
    for (uint32_t t = nb + lane; t < a0; t += 64) pair_row[t] = -1;
    for (uint32_t t = a0 + 4u * lane; t < a1; t += 256)
        *reinterpret_cast<int4 *>(pair_row + t) = make_int4(-1, -1, -1, -1);
    for (uint32_t t = a1 + lane; t < ne; t += 64) pair_row[t] = -1;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // the fill before the values
#pragma unroll
    for (int g = 0; g < kWaveGroups; g++)
        if (pval[g] >= 0) pair_row[row[g]] = pval[g];
    const uint64_t any_err = __ballot(err);
    if (lane == 0) {
        if (any_err) atomicOr(&uo[u].struct_err, 1u);
        uo[u].fast = 1;
    }
}

// First chunk (unit-local index) x >= x0 whose walk reaches depth d or below,
// or nch.  The whole wave runs one search: each step tests 64 entries of one
// level (chunks, 64-chunk groups, 4096-chunk groups) with a ballot.
__device__ __forceinline__ uint32_t wave_next_reaching(const int32_t *__restrict__ gmin,
                                                       const int32_t *__restrict__ gmin1,
                                                       const int32_t *__restrict__ gmin2,
                                                       uint32_t c0, uint32_t g0, uint32_t h0,
                                                       uint32_t nch, uint32_t x, int32_t d) {
    const uint32_t lane = lane_id();
    const uint32_t G = (nch + 63) / 64, H = (nch + 4095) / 4096;
    while (x < nch) {
        if (x & 63u) {  // chunks up to the next group boundary
            const uint32_t end = min(nch, (x | 63u) + 1);
            const uint64_t m = __ballot(x + lane < end && gmin[c0 + x + lane] <= d);
            if (m) return x + (uint32_t)__ffsll((long long)m) - 1;
            x = end;
            continue;
        }
        uint32_t g = x >> 6;
        if (g & 63u) {  // groups up to the next super-group boundary
            const uint32_t gend = min(G, (g | 63u) + 1);
            const uint64_t m = __ballot(g + lane < gend && gmin1[g0 + g + lane] <= d);
            if (!m) {
                x = gend * 64;
                continue;
            }
            g += (uint32_t)__ffsll((long long)m) - 1;
        } else {
            const uint32_t h = g >> 6;
            const uint32_t hend = min(H, h + 64);
            const uint64_t mh = __ballot(h + lane < hend && gmin2[h0 + h + lane] <= d);
            if (!mh) {
                x = hend * 4096;
                continue;
            }
            const uint32_t hh = h + (uint32_t)__ffsll((long long)mh) - 1;
            const uint64_t mg =
                __ballot(64 * hh + lane < G && gmin1[g0 + 64 * hh + lane] <= d);
            g = 64 * hh + (uint32_t)__ffsll((long long)mg) - 1;
        }
        const uint64_t mc = __ballot(64 * g + lane < nch && gmin[c0 + 64 * g + lane] <= d);
        return 64 * g + (uint32_t)__ffsll((long long)mc) - 1;
    }
    return nch;
}

// K4: openers left open by their chunk, one wave per chunk and one lane per
// depth.  A lane first looks at the next few chunks itself; openers closed
// further away are then searched by the whole wave, one at a time.  A block
// is kCrossWaves waves, each taking its own chunks.
constexpr uint32_t kNearChunks = 8;
constexpr uint32_t kCrossWaves = 4;
__global__ __launch_bounds__(64 * kCrossWaves) void k_match_cross(
    const uint8_t *__restrict__ oc_char, const uint32_t *__restrict__ oc_idx,
    UnitOut *__restrict__ uo, const uint32_t *__restrict__ chunk_unit,
    const ChunkStat *__restrict__ ck, const int32_t *__restrict__ gmin, const int32_t *__restrict__ gmin1,
    const int32_t *__restrict__ gmin2, const uint32_t *__restrict__ lows,
    const uint32_t *__restrict__ highs, int32_t *__restrict__ pair_row,
    const LaunchState *__restrict__ ls) {
    if (ls->overflow) return;
    const uint32_t n_chunks = ls->n_chunks;
    const uint32_t lane = lane_id();
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (uint32_t c = blockIdx.x * kCrossWaves + w; c < n_chunks; c += gridDim.x * kCrossWaves) {
        const uint32_t u = chunk_unit[c];
        if (!uo[u].fast) continue;
        const uint32_t c0 = uo[u].chunk0, nch = uo[u].nchunks;
        const uint32_t g0 = uo[u].grp0, h0 = uo[u].sgrp0;
        const uint32_t xc = c - c0;
        const uint32_t nb = uo[u].n_base, pbase = uo[u].pair_base + 1u;
        const uint32_t mb = uo[u].m_base, me = uo[u].m_end;
        const ChunkStat s = ck[c];
        const int32_t dend = min(s.d0 + s.d, kMatchKeys);
        int32_t err = 0;
        for (int32_t dbase = max(gmin[c], 0); dbase < dend; dbase += 64) {
            const int32_t d = dbase + (int32_t)lane;
            const bool active = d < dend;
            uint32_t found = nch;
            if (active) {
                for (uint32_t x = xc + 1; x < nch && x <= xc + kNearChunks; x++) {
                    if (gmin[c0 + x] <= d) {
                        found = x;
                        break;
                    }
                }
            }
            uint64_t pend = __ballot(active && found == nch);
            while (pend) {
                const uint32_t k = (uint32_t)__ffsll((long long)pend) - 1;
                pend &= pend - 1;
                const int32_t dk = __shfl(d, (int)k, 64);
                const uint32_t r = wave_next_reaching(gmin, gmin1, gmin2, c0, g0, h0, nch,
                                                      xc + kNearChunks + 1, dk);
                if (lane == k) found = r;
            }
            if (!active) continue;
            if (found >= nch) {  // not reached for a well-nested unit
                err = 1;
                continue;
            }
            const uint32_t m = highs[(uint64_t)c * kMatchKeys + d];
            const uint32_t m2 = lows[(uint64_t)(c0 + found) * kMatchKeys + d];
            // Both slots are written for a well-nested unit; the range check
            // keeps a stale slot from ever indexing outside the unit.
            if (m < mb || m >= me || m2 < mb || m2 >= me) {
                err = 1;
                continue;
            }
            if ((oc_char[m] ^ oc_char[m2]) != 6) err = 1;
            pair_row[oc_idx[m]] = (int32_t)(oc_idx[m2] - nb + pbase);
        }
        if (__ballot(err) && lane == 0) atomicOr(&uo[u].struct_err, 1u);
    }
}

// The first matching steps in two launches: blocks [0, g_stats) take the
// chunk walks, the rest the units with at most kSmallBrackets brackets (four
// per block); then blocks [0, g_big) scan the chunk walks of units with more
// than 64 chunks, the rest those of the other units (one per wave).
__global__ __launch_bounds__(kMatchThreads) void k_match_front(
    const uint8_t *__restrict__ oc_char, const uint32_t *__restrict__ oc_idx,
    UnitOut *__restrict__ uo, uint32_t n_units, const uint32_t *__restrict__ chunk_unit,
    ChunkStat *__restrict__ ck, int32_t *__restrict__ pair_row, LaunchState *__restrict__ ls,
    uint32_t g_stats, uint32_t *__restrict__ slow_units) {
    if (blockIdx.x < g_stats)
        chunk_stats_blocks(blockIdx.x, g_stats, oc_char, uo, chunk_unit, ck, ls);
    else
        match_small_block(blockIdx.x - g_stats, oc_char, oc_idx, uo, n_units, pair_row, ls,
                          slow_units);
}

__global__ __launch_bounds__(kScanThreads) void k_chunk_scans(
    UnitOut *__restrict__ uo, uint32_t n_units, ChunkStat *__restrict__ ck,
    int32_t *__restrict__ gmin, int32_t *__restrict__ gmin1, int32_t *__restrict__ gmin2,
    DepthAgg *__restrict__ gagg, int32_t *__restrict__ gd0, const uint32_t *__restrict__ big_units,
    LaunchState *__restrict__ ls, uint32_t g_big, uint32_t *__restrict__ slow_units) {
    if (blockIdx.x < g_big)
        chunk_scan_blocks(blockIdx.x, g_big, n_units, uo, ck, gmin, gmin1, gmin2, gagg, gd0, big_units, ls,
                          slow_units);
    else
        chunk_scan_small_block(blockIdx.x - g_big, uo, n_units, ck, gmin, gmin1, gmin2, ls,
                               slow_units);
}

namespace {
// The chunk count is only known on the device: grid-stride over a grid
// that fills the chip (256 CUs) without launching one block per bound.
uint32_t chunk_grid(uint32_t chunk_bound, uint32_t cap) {
    return chunk_bound < cap ? (chunk_bound ? chunk_bound : 1u) : cap;
}

// Blocks of `kernel` (kMatchThreads threads each) that the device holds at
// once, or 2048 if the runtime cannot tell.
uint32_t resident_blocks(const void *kernel) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kMatchThreads, 0) != hipSuccess ||
        cus <= 0 || per_cu <= 0)
        return 2048u;
    return (uint32_t)(cus * per_cu);
}

// k_match_front's chunk walks take chunks bid, bid + g, ...: with a grid
// above what the device holds at once, the last blocks would start only
// when the first ones end, and the walks would take two turns.
uint32_t front_capacity() {
    static const uint32_t cap = resident_blocks(reinterpret_cast<const void *>(k_match_front));
    return cap;
}
// k_match_local takes chunks from a ticket counter: blocks beyond what the
// device holds at once would only find the tickets gone.
uint32_t local_capacity() {
    static const uint32_t cap = resident_blocks(reinterpret_cast<const void *>(k_match_local));
    return cap;
}
}  // namespace

hipError_t launch_match_front(hipStream_t s, const LaunchBufs &b, uint32_t n_units,
                              uint32_t chunk_bound, int32_t *pair_row) {
    const uint32_t g = chunk_grid(chunk_bound, front_capacity());
    const uint32_t gb = n_units < 256u ? n_units : 256u;
    hipLaunchKernelGGL(k_match_front, dim3(g + (n_units + 3) / 4), dim3(kMatchThreads), 0, s,
                       b.oc_char, b.oc_idx, b.uout, n_units, b.chunk_unit, b.ck, pair_row, b.ls, g,
                       b.slow_units);
    hipLaunchKernelGGL(k_chunk_scans, dim3(gb + (n_units + kScanThreads / 64 - 1) / (kScanThreads / 64)),
                       dim3(kScanThreads), 0, s, b.uout, n_units, b.ck, b.gmin, b.gmin1, b.gmin2,
                       b.gagg, b.gd0, b.big_units, b.ls, gb, b.slow_units);
    return hipGetLastError();
}

hipError_t launch_match_local(hipStream_t s, const LaunchBufs &b, uint32_t chunk_bound,
                              int32_t *pair_row) {
    const uint32_t g = chunk_grid(chunk_bound, local_capacity());
    const uint32_t gc = chunk_grid((chunk_bound + kCrossWaves - 1) / kCrossWaves, 2048u);
    // uo twice: the kernel's fields as read-only (scalar loads) and the
    // error flag it sets; it writes no field that it reads.
    hipLaunchKernelGGL(k_match_local, dim3(g), dim3(kMatchThreads), 0, s, b.oc_char, b.oc_idx,
                       b.uout, b.chunk_unit, b.ck, b.lows, b.highs, pair_row, b.ls,
                       (const UnitOut *)b.uout);
    hipLaunchKernelGGL(k_match_cross, dim3(gc), dim3(64 * kCrossWaves), 0, s, b.oc_char, b.oc_idx, b.uout,
                       b.chunk_unit, b.ck, b.gmin, b.gmin1, b.gmin2, b.lows, b.highs, pair_row,
                       b.ls);
    return hipGetLastError();
}

hipError_t launch_match(hipStream_t s, const LaunchBufs &b, uint32_t n_units, uint32_t chunk_bound,
                        uint32_t group_bound, int32_t *pair_row) {
    (void)group_bound;
    const hipError_t e = launch_match_front(s, b, n_units, chunk_bound, pair_row);
    return e != hipSuccess ? e : launch_match_local(s, b, chunk_bound, pair_row);
}

}  // namespace cjx
