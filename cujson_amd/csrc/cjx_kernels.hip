// cjx_kernels.hip - hand-written gfx950 kernels of the three-phase JSON parse.
//
// Phases 1+2 (the fused tokenizer) live in cjx_tokenize.hip.
// Phase 3 (parse_standard_json.cu:1345-1582) runs over the compact bracket
// array: a segmented +1/-1 depth scan (k_depth_keys, from the chunk start
// depths of the matching scan), a stable LDS counting sort by depth key
// (k_depth_keys' last block scans the histograms, then k_scatter), and a
// dense pair-row writer that applies the reference's group-of-four pairing
// rule (k_pairs).
#include "cjx_device.h"

namespace cjx {

__device__ __forceinline__ int32_t wave_incl_scan_i(int32_t x) {
    return (int32_t)wave_incl_scan((uint32_t)x);
}

// ---------------------------------------------------------------------------
// k_reset: clears look-back state and per-unit results before a parse.
// ---------------------------------------------------------------------------
__global__ void k_reset(unsigned long long *tile_st, uint32_t ntiles, LaunchState *ls,
                        UnitOut *uo, uint32_t n_units) {
    const uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = i0; i < 2 * ntiles; i += stride) tile_st[i] = 0;
    for (uint32_t i = i0; i < n_units; i += stride) uo[i] = UnitOut{};
    if (i0 == 0) *ls = LaunchState{};
}

// k_reset_one: k_reset for a launch of one unit whose descriptor comes as a
// kernel argument: also writes the unit table (the descriptor, its tile ->
// unit map and small-unit entry, all unit 0), so the launch needs no
// host-to-device copy before it.
__global__ void k_reset_one(UnitDesc u, UnitDesc *units, uint32_t *tile_unit, uint32_t ntiles,
                            uint32_t n_small, unsigned long long *tile_st, LaunchState *ls,
                            UnitOut *uo) {
    const uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = i0; i < 2 * ntiles; i += stride) tile_st[i] = 0;
    for (uint32_t i = i0; i < ntiles + n_small; i += stride) tile_unit[i] = 0;
    if (i0 == 0) {
        units[0] = u;
        uo[0] = UnitOut{};
        *ls = LaunchState{};
    }
}

// Per-unit bases: pair_base offsets unit u's pair values by the structural
// count of all earlier units when results are concatenated (the JSON Lines
// lastStructuralIndex carry, parse_json_lines.cu:1222); row_base is that
// count always; chunk0, grp0 and sgrp0 number the units' matching chunks and
// their 64- and 4096-chunk groups consecutively.  Units with at most
// kSmallBrackets brackets have no chunks (k_match_front's wave matcher
// pairs them in one wave).  Three launches: block sums, a scan of the block sums, and a final
// pass that writes the fields, the chunk -> unit map of units with up to 64
// chunks and the list of larger units (whose maps k_chunk_map writes).
constexpr int kBaseThreads = 256;

// Counters of a unit: structurals, chunks, 64- and 4096-chunk groups, brackets.
constexpr int kUnitCounters = 5;
__device__ __forceinline__ void unit_counts(const UnitOut &o, uint32_t v[kUnitCounters]) {
    v[0] = o.n_end - o.n_base;
    const uint32_t m = o.m_end - o.m_base;
    v[1] = m > (uint32_t)kSmallBrackets ? (m + kMatchChunk - 1) / kMatchChunk : 0u;
    v[2] = (v[1] + 63) / 64;
    v[3] = (v[1] + 4095) / 4096;
    v[4] = m;
}

// Block exclusive scan of the unit counters over kBaseThreads threads.
__device__ __forceinline__ void block_scan4(const uint32_t v[kUnitCounters],
                                            uint32_t pre[kUnitCounters],
                                            uint32_t tot[kUnitCounters],
                                            uint32_t (*wsum)[kBaseThreads / 64]) {
    constexpr int kW = kBaseThreads / 64;
    const uint32_t w = threadIdx.x >> 6;
    uint32_t in[kUnitCounters];
#pragma unroll
    for (int k = 0; k < kUnitCounters; k++) {
        in[k] = wave_incl_scan(v[k]);
        if (lane_id() == 63) wsum[k][w] = in[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kUnitCounters; k++) {
        pre[k] = in[k] - v[k];
        tot[k] = 0;
        for (uint32_t j = 0; j < (uint32_t)kW; j++) {
            if (j < w) pre[k] += wsum[k][j];
            tot[k] += wsum[k][j];
        }
    }
}

__global__ __launch_bounds__(kBaseThreads) void k_ub_partial(const UnitOut *__restrict__ uo,
                                                             uint32_t n_units,
                                                             uint32_t *__restrict__ part) {
    __shared__ uint32_t wsum[kUnitCounters][kBaseThreads / 64];
    const uint32_t u = blockIdx.x * kBaseThreads + threadIdx.x;
    uint32_t v[kUnitCounters] = {0, 0, 0, 0, 0};
    if (u < n_units) unit_counts(uo[u], v);
    uint32_t pre[kUnitCounters], tot[kUnitCounters];
    block_scan4(v, pre, tot, wsum);
    if (threadIdx.x < kUnitCounters)
        part[blockIdx.x * kUnitCounters + threadIdx.x] = tot[threadIdx.x];
}

// One block: exclusive prefixes of the block sums (in place) and the totals
// (structurals, chunks and brackets of the launch).
__global__ __launch_bounds__(kBaseThreads) void k_ub_scan(uint32_t *__restrict__ part,
                                                          uint32_t n_blocks,
                                                          LaunchState *__restrict__ ls) {
    __shared__ uint32_t wsum[kUnitCounters][kBaseThreads / 64];
    __shared__ uint32_t carry[kUnitCounters];
    if (threadIdx.x < kUnitCounters) carry[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t base = 0; base < n_blocks; base += kBaseThreads) {
        const uint32_t b = base + threadIdx.x;
        uint32_t v[kUnitCounters] = {0, 0, 0, 0, 0};
        if (b < n_blocks)
            for (int k = 0; k < kUnitCounters; k++) v[k] = part[b * kUnitCounters + k];
        uint32_t pre[kUnitCounters], tot[kUnitCounters];
        block_scan4(v, pre, tot, wsum);
        if (b < n_blocks)
            for (int k = 0; k < kUnitCounters; k++) part[b * kUnitCounters + k] = carry[k] + pre[k];
        __syncthreads();
        if (threadIdx.x < kUnitCounters) carry[threadIdx.x] += tot[threadIdx.x];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ls->n_total = carry[0];
        ls->n_chunks = carry[1];
        ls->m_total = carry[4];
    }
}

__global__ __launch_bounds__(kBaseThreads) void k_ub_final(UnitOut *__restrict__ uo,
                                                           uint32_t n_units, uint32_t concat,
                                                           const uint32_t *__restrict__ part,
                                                           LaunchState *__restrict__ ls,
                                                           uint32_t *__restrict__ chunk_unit,
                                                           uint32_t *__restrict__ big_units) {
    __shared__ uint32_t wsum[kUnitCounters][kBaseThreads / 64];
    const uint32_t u = blockIdx.x * kBaseThreads + threadIdx.x;
    uint32_t v[kUnitCounters] = {0, 0, 0, 0, 0};
    if (u < n_units) unit_counts(uo[u], v);
    uint32_t pre[kUnitCounters], tot[kUnitCounters];
    block_scan4(v, pre, tot, wsum);
    if (u >= n_units) return;
#pragma unroll
    for (int k = 0; k < kUnitCounters; k++) pre[k] += part[blockIdx.x * kUnitCounters + k];
    uo[u].row_base = pre[0];
    uo[u].pair_base = concat ? pre[0] : 0u;
    uo[u].chunk0 = pre[1];
    uo[u].nchunks = v[1];
    uo[u].grp0 = pre[2];
    uo[u].sgrp0 = pre[3];
    if (v[1] > 64) big_units[atomicAdd(&ls->n_big, 1u)] = u;
    else
        for (uint32_t j = 0; j < v[1]; j++) chunk_unit[pre[1] + j] = u;
}

// The same fields in one launch for launches of at most kOneBlockUnits
// units: one block scans tiles of kBaseThreads units in turn with a carry,
// then writes the chunk -> unit maps of the larger units with all threads.
constexpr uint32_t kOneBlockUnits = 1024;
__global__ __launch_bounds__(kBaseThreads) void k_unit_bases_one(UnitOut *__restrict__ uo,
                                                                 uint32_t n_units, uint32_t concat,
                                                                 LaunchState *__restrict__ ls,
                                                                 uint32_t *__restrict__ chunk_unit,
                                                                 uint32_t *__restrict__ big_units) {
    __shared__ uint32_t wsum[kUnitCounters][kBaseThreads / 64];
    __shared__ uint32_t carry[kUnitCounters];
    __shared__ uint32_t s_nbig;
    if (threadIdx.x < kUnitCounters) carry[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t base = 0; base < n_units; base += kBaseThreads) {
        const uint32_t u = base + threadIdx.x;
        uint32_t v[kUnitCounters] = {0, 0, 0, 0, 0};
        if (u < n_units) unit_counts(uo[u], v);
        uint32_t pre[kUnitCounters], tot[kUnitCounters];
        block_scan4(v, pre, tot, wsum);
        if (u < n_units) {
#pragma unroll
            for (int k = 0; k < kUnitCounters; k++) pre[k] += carry[k];
            uo[u].row_base = pre[0];
            uo[u].pair_base = concat ? pre[0] : 0u;
            uo[u].chunk0 = pre[1];
            uo[u].nchunks = v[1];
            uo[u].grp0 = pre[2];
            uo[u].sgrp0 = pre[3];
            if (v[1] > 64) big_units[atomicAdd(&ls->n_big, 1u)] = u;
            else
                for (uint32_t j = 0; j < v[1]; j++) chunk_unit[pre[1] + j] = u;
        }
        __syncthreads();  // carry and wsum are read before they change
        if (threadIdx.x < kUnitCounters) carry[threadIdx.x] += tot[threadIdx.x];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ls->n_total = carry[0];
        ls->n_chunks = carry[1];
        ls->m_total = carry[4];
        s_nbig = atomicAdd(&ls->n_big, 0u);
    }
    __syncthreads();  // the big-unit list and the fields above are written
    for (uint32_t bi = 0; bi < s_nbig; bi++) {
        const uint32_t u = big_units[bi];
        const uint32_t c0 = uo[u].chunk0, nch = uo[u].nchunks;
        for (uint32_t j = threadIdx.x; j < nch; j += kBaseThreads) chunk_unit[c0 + j] = u;
    }
}

// Chunk -> unit map of the units with more than 64 chunks, one block per unit.
__global__ __launch_bounds__(256) void k_chunk_map(const UnitOut *__restrict__ uo,
                                                   const uint32_t *__restrict__ big_units,
                                                   uint32_t *__restrict__ chunk_unit,
                                                   const LaunchState *__restrict__ ls) {
    for (uint32_t bi = blockIdx.x; bi < ls->n_big; bi += gridDim.x) {
        const uint32_t u = big_units[bi];
        const uint32_t c0 = uo[u].chunk0, nch = uo[u].nchunks;
        for (uint32_t j = threadIdx.x; j < nch; j += blockDim.x) chunk_unit[c0 + j] = u;
    }
}

// Packs the per-unit output regions densely (see launch_assemble).  Each
// thread moves kAsmRows consecutive rows and finds their unit once by binary
// search over row_base.
constexpr uint32_t kAsmRows = 16;
__global__ __launch_bounds__(256) void k_assemble(const UnitOut *__restrict__ uo, uint32_t n_units,
                                                  const uint32_t *__restrict__ st,
                                                  const int32_t *__restrict__ pr,
                                                  int32_t *__restrict__ dst_struct,
                                                  int32_t *__restrict__ dst_pair, uint32_t N,
                                                  int32_t *__restrict__ sentinels) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t r0 = t * kAsmRows;
    if (t == 0 && sentinels) {  // host layout, parse_json_lines.cu:1070-1089
        sentinels[0] = 0;
        sentinels[N + 1] = (int32_t)(N + 1);
        sentinels[2 * (uint64_t)N + 2] = -1;
        sentinels[2 * (uint64_t)N + 3] = -1;
    }
    if (r0 >= N) return;
    uint32_t lo = 0, hi = n_units;  // last unit with row_base <= r0
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (uo[mid].row_base <= r0) lo = mid;
        else hi = mid;
    }
    uint32_t u = lo;
    uint32_t ub = uo[u].row_base, ue = ub + (uo[u].n_end - uo[u].n_base);
    const uint64_t r1 = r0 + kAsmRows < N ? r0 + kAsmRows : N;
    for (uint64_t r = r0; r < r1; r++) {
        while (r >= ue) {  // skips units without structurals
            u++;
            ub = uo[u].row_base;
            ue = ub + (uo[u].n_end - uo[u].n_base);
        }
        const uint64_t src = uo[u].n_base + (r - ub);
        dst_struct[r] = (int32_t)st[src];
        dst_pair[r] = pr[src];
    }
}

// Per-document host layout: document u (rows [row_base, row_base + n_u) of
// the packed order) becomes the 2(n_u + 2) cells
//   [0, s_1..s_n, n + 1, p_1..p_n, -1, -1]   (parse_standard_json.cu:1659-1680)
// starting at cell 2 row_base + 4u of dst, so every document of a batch is a
// complete reference-layout buffer inside one copy.  Threads below n_units
// also write their unit's four sentinel cells.
__global__ __launch_bounds__(256) void k_assemble_docs(const UnitOut *__restrict__ uo,
                                                       uint32_t n_units,
                                                       const uint32_t *__restrict__ st,
                                                       const int32_t *__restrict__ pr,
                                                       int32_t *__restrict__ dst, uint32_t N) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n_units) {
        const uint64_t n = uo[t].n_end - uo[t].n_base;
        int32_t *d = dst + 2 * (uint64_t)uo[t].row_base + 4 * t;
        d[0] = 0;
        d[n + 1] = (int32_t)(n + 1);
        d[2 * n + 2] = -1;
        d[2 * n + 3] = -1;
    }
    const uint64_t r0 = t * kAsmRows;
    if (r0 >= N) return;
    uint32_t lo = 0, hi = n_units;  // last unit with row_base <= r0
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (uo[mid].row_base <= r0) lo = mid;
        else hi = mid;
    }
    uint32_t u = lo;
    uint32_t ub = uo[u].row_base, un = uo[u].n_end - uo[u].n_base;
    const uint64_t r1 = r0 + kAsmRows < N ? r0 + kAsmRows : N;
    for (uint64_t r = r0; r < r1; r++) {
        while (r >= (uint64_t)ub + un) {  // skips units without structurals
            u++;
            ub = uo[u].row_base;
            un = uo[u].n_end - uo[u].n_base;
        }
        const uint64_t j = r - ub, src = uo[u].n_base + j;
        int32_t *d = dst + 2 * (uint64_t)ub + 4 * (uint64_t)u;
        d[1 + j] = (int32_t)st[src];
        d[un + 2 + j] = pr[src];
    }
}

// ---------------------------------------------------------------------------
// Phase 3 helpers.  All phase-3 kernels split the unit's bracket range
// [m_base, m_end) into gridDim.x contiguous block ranges the same way: whole
// matching chunks (kMatchChunk brackets) per block, so a block's start depth
// is the start depth of its first chunk (ChunkStat::d0).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void block_range(uint32_t base, uint32_t end, uint32_t b, uint32_t G,
                                            uint32_t &b0, uint32_t &b1) {
    const uint32_t M = end - base;
    const uint64_t per = ((uint64_t)M + G - 1) / G;
    const uint64_t chunk = (per + kMatchChunk - 1) / kMatchChunk * kMatchChunk;
    const uint64_t lo = (uint64_t)base + (uint64_t)b * chunk;
    const uint64_t hi = lo + chunk;
    b0 = (uint32_t)(lo < end ? lo : end);
    b1 = (uint32_t)(hi < end ? hi : end);
}

__device__ __forceinline__ int32_t block_reduce_sum(int32_t x, int32_t *lds /*[4]*/) {
    const uint32_t w = threadIdx.x >> 6;
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
    __syncthreads();
    if (lane_id() == 0) lds[w] = x;
    __syncthreads();
    int32_t r = 0;
    for (int k = 0; k < kP3Threads / 64; k++) r += lds[k];
    return r;
}

__device__ __forceinline__ int bracket_delta(uint8_t c) {
    return (c == 0x5B || c == 0x7B) ? 1 : ((c == 0x5D || c == 0x7D) ? -1 : 0);
}

// Exclusive offsets per (block, digit) of the G blocks' histograms:
// digit-major, block-minor, so that a stable scatter reproduces
// thrust::stable_sort_by_key (parse_standard_json.cu:1546).  Run by the last
// block of the kernel that wrote the histograms (all kP3Threads threads,
// thread k for digit k); `wsum` is kP3Threads / 64 words of LDS.
__device__ __forceinline__ void scan_hist_block(uint32_t *__restrict__ hist, uint32_t G,
                                                uint32_t *wsum) {
    static_assert(kP3Threads == 256, "one thread per digit");
    const uint32_t k = threadIdx.x, w = k >> 6;
    uint32_t total = 0;
    for (uint32_t b = 0; b < G; b++) total += hist[b * 256 + k];
    const uint32_t incl = wave_incl_scan(total);
    if (lane_id() == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t run = incl - total;
    for (uint32_t j = 0; j < w; j++) run += wsum[j];
    for (uint32_t b = 0; b < G; b++) {
        const uint32_t hv = hist[b * 256 + k];
        hist[b * 256 + k] = run;
        run += hv;
    }
}

// Whether this block is the last of the grid to finish writing its
// histogram (its writes and everyone's are then visible to it).  The last
// block resets the ticket for the next histogram kernel of the launch.
__device__ __forceinline__ bool last_hist_block(LaunchState *ls, uint32_t *flag) {
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) *flag = atomicAdd(&ls->hist_ticket, 1u) == gridDim.x - 1 ? 1u : 0u;
    __syncthreads();
    const bool last = *flag != 0;
    if (last) {
        __threadfence();
        if (threadIdx.x == 0) ls->hist_ticket = 0;
    }
    return last;
}

// Depth keys (inclusive scan modulo 2^8 or 2^16, parse_standard_json.cu:1537;
// openers take depth-1, :1540 / utils.cu:104-125) and the histogram of the
// low key byte per block; the last block turns the histograms into scatter
// offsets.  A block starts at the start depth of its first chunk, from the
// matching scan (which runs for every unit, well nested or not).
__global__ __launch_bounds__(kP3Threads) void k_depth_keys(
    const uint8_t *__restrict__ oc_char, const UnitOut *__restrict__ uo, uint32_t u,
    uint32_t kmask, const ChunkStat *__restrict__ ck, uint16_t *__restrict__ keys,
    uint32_t *__restrict__ hist, LaunchState *__restrict__ ls) {
    __shared__ uint32_t s_last;
    if ((ls->overflow | uo[u].fast) != 0u) return;  // both read at once (u < n_units)
    __shared__ int32_t wsum[kP3Threads / 64];
    __shared__ uint32_t h[256];
    const uint32_t tid = threadIdx.x, w = tid >> 6;
    const uint32_t mb = uo[u].m_base;
    uint32_t b0, b1;
    block_range(mb, uo[u].m_end, blockIdx.x, gridDim.x, b0, b1);
    // A block past the first starts a chunk of a unit of more than
    // kMatchChunk brackets, so the unit has chunks and their start depths.
    int32_t pre = b0 > mb && b0 < b1 ? ck[uo[u].chunk0 + (b0 - mb) / kMatchChunk].d0 : 0;
    h[tid] = 0;
    __syncthreads();
    for (uint32_t base = b0; base < b1; base += kP3Threads) {
        const uint32_t i = base + tid;
        const bool valid = i < b1;
        const uint8_t c = valid ? oc_char[i] : 0;
        const int32_t d = valid ? bracket_delta(c) : 0;
        const int32_t incl = wave_incl_scan_i(d);
        if (lane_id() == 63) wsum[w] = incl;
        __syncthreads();
        int32_t woff = 0, tot = 0;
        for (uint32_t k = 0; k < kP3Threads / 64; k++) {
            if (k < w) woff += wsum[k];
            tot += wsum[k];
        }
        if (valid) {
            const uint32_t depth = (uint32_t)(pre + woff + incl) & kmask;
            const uint32_t key = (d > 0) ? ((depth - 1u) & kmask) : depth;
            keys[i] = (uint16_t)key;
            atomicAdd(&h[key & 0xFFu], 1u);
        }
        pre += tot;
        __syncthreads();
    }
    hist[blockIdx.x * 256 + tid] = h[tid];
    if (last_hist_block(ls, &s_last)) scan_hist_block(hist, gridDim.x, (uint32_t *)wsum);
}

// Histogram of key byte `shift` over elements taken in the order perm_in
// (second pass of the 16-bit sort); the last block turns the histograms into
// scatter offsets.
__global__ __launch_bounds__(kP3Threads) void k_hist(const uint16_t *__restrict__ keys,
                                                     const UnitOut *__restrict__ uo, uint32_t u,
                                                     const uint32_t *__restrict__ perm_in,
                                                     uint32_t shift, uint32_t *__restrict__ hist,
                                                     LaunchState *__restrict__ ls) {
    __shared__ uint32_t h[256];
    __shared__ uint32_t s_last, wsum[kP3Threads / 64];
    if ((ls->overflow | uo[u].fast) != 0u) return;  // both read at once (u < n_units)
    const uint32_t tid = threadIdx.x;
    uint32_t b0, b1;
    block_range(uo[u].m_base, uo[u].m_end, blockIdx.x, gridDim.x, b0, b1);
    h[tid] = 0;
    __syncthreads();
    for (uint32_t e = b0 + tid; e < b1; e += kP3Threads)
        atomicAdd(&h[(keys[perm_in[e]] >> shift) & 0xFFu], 1u);
    __syncthreads();
    hist[blockIdx.x * 256 + tid] = h[tid];
    if (last_hist_block(ls, &s_last)) scan_hist_block(hist, gridDim.x, wsum);
}

// Stable scatter by one key byte.  Ranks among equal digits inside a 256-wide
// step come from a wave-level match (8 ballots) plus per-wave digit counts in
// LDS, so the output order is the input order within each digit.
template <bool IDENT, bool FINAL>
__global__ __launch_bounds__(kP3Threads) void k_scatter(
    const uint16_t *__restrict__ keys, const UnitOut *__restrict__ uo, uint32_t u,
    uint32_t shift, const uint32_t *__restrict__ hist_off, const uint32_t *__restrict__ perm_in,
    uint32_t *__restrict__ perm_out, uint32_t *__restrict__ spos,
    const LaunchState *__restrict__ ls) {
    __shared__ uint32_t run[256];
    if ((ls->overflow | uo[u].fast) != 0u) return;  // both read at once (u < n_units)
    __shared__ uint32_t wcnt[kP3Threads / 64][256];
    const uint32_t tid = threadIdx.x, w = tid >> 6;
    const uint64_t lt = lanemask_lt();
    const uint32_t mb = uo[u].m_base;
    uint32_t b0, b1;
    block_range(mb, uo[u].m_end, blockIdx.x, gridDim.x, b0, b1);
    run[tid] = hist_off[blockIdx.x * 256 + tid];
    for (uint32_t base = b0; base < b1; base += kP3Threads) {
        const uint32_t e = base + tid;
        const bool valid = e < b1;
        uint32_t m = 0, digit = 0;
        if (valid) {
            m = IDENT ? e : perm_in[e];
            digit = (keys[m] >> shift) & 0xFFu;
        }
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
            const uint32_t bv = (digit >> bit) & 1u;
            const uint64_t bm = __ballot(bv);
            peers &= bv ? bm : ~bm;
        }
#pragma unroll
        for (int k = 0; k < kP3Threads / 64; k++) wcnt[k][tid] = 0;
        __syncthreads();
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (valid && rank == 0) wcnt[w][digit] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[digit] + rank;
            for (uint32_t k = 0; k < w; k++) pos += wcnt[k][digit];
            perm_out[mb + pos] = m;
            if (FINAL) spos[m] = pos;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int k = 0; k < kP3Threads / 64; k++) add += wcnt[k][tid];
        run[tid] += add;
        __syncthreads();
    }
}

// Pair row writer (validate_expand_MathAPI_new2, parse_standard_json.cu:1431-1514).
// Sorted position s of bracket m pairs with s+1 when s is even, in groups of
// four: non-last groups check xor == 6 and order, the last group checks xor
// only, and a last group of 1 or 3 is an error with no writes.  The row is
// written densely: each block stages a 4096-entry chunk of the row in LDS
// (-1 where nothing is written) and stores it with coalesced writes.
__global__ __launch_bounds__(kP3Threads) void k_pairs(
    const uint8_t *__restrict__ oc_char, const uint32_t *__restrict__ oc_idx,
    const uint32_t *__restrict__ sorted, const uint32_t *__restrict__ spos,
    UnitOut *__restrict__ uo, uint32_t u, int32_t *__restrict__ pair_row,
    const LaunchState *__restrict__ ls) {
    __shared__ int32_t row[kPairChunk];
    if ((ls->overflow | uo[u].fast) != 0u) return;  // both read at once (u < n_units)
    __shared__ uint32_t s_m0, s_m1;
    __shared__ int32_t red[kP3Threads / 64];
    const uint32_t tid = threadIdx.x;
    const uint32_t nb = uo[u].n_base, ne = uo[u].n_end;
    const uint32_t mb = uo[u].m_base, me = uo[u].m_end;
    const uint32_t N = ne - nb, M = me - mb;
    const uint32_t last_group = M ? 4u * ((M + 3u) / 4u - 1u) : 0u;
    const uint32_t pbase = uo[u].pair_base + 1u;
    const uint32_t nchunks = (N + kPairChunk - 1) / kPairChunk;
    int32_t err = 0;
    for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const uint32_t t0 = nb + c * kPairChunk;
        const uint32_t t1 = (t0 + kPairChunk < ne) ? t0 + kPairChunk : ne;
        for (uint32_t i = tid; i < kPairChunk; i += kP3Threads) row[i] = -1;
        if (tid < 2) {  // lower_bound of t0 (tid 0) and t1 (tid 1) in oc_idx[mb, me)
            const uint32_t key = tid ? t1 : t0;
            uint32_t lo = mb, hi = me;
            while (lo < hi) {
                const uint32_t mid = lo + ((hi - lo) >> 1);
                if (oc_idx[mid] < key) lo = mid + 1;
                else hi = mid;
            }
            if (tid) s_m1 = lo;
            else s_m0 = lo;
        }
        __syncthreads();
        for (uint32_t m = s_m0 + tid; m < s_m1; m += kP3Threads) {
            const uint32_t s = spos[m];
            if (s & 1u) continue;
            if (s + 1u >= M) continue;
            if ((M & 3u) == 3u && s == M - 3u) continue;
            const uint32_t m2 = sorted[mb + s + 1u];
            const int c0 = (int)(signed char)oc_char[m];
            const int c1 = (int)(signed char)oc_char[m2];
            if ((c0 ^ c1) != 6) err = 1;
            if (s < last_group && c0 > c1) err = 1;
            row[oc_idx[m] - t0] = (int32_t)(oc_idx[m2] - nb + pbase);
        }
        __syncthreads();
        for (uint32_t i = tid; i < t1 - t0; i += kP3Threads) pair_row[t0 + i] = row[i];
        __syncthreads();
    }
    if (blockIdx.x == 0 && tid == 0 && (M & 1u)) err = 1;
    err = block_reduce_sum(err, red);
    if (tid == 0 && err) atomicOr(&uo[u].struct_err, 1u);
}

}  // namespace cjx

// ---------------------------------------------------------------------------
// Host-side launchers (the only entry points the driver uses).
// ---------------------------------------------------------------------------
namespace cjx {

// Sort path of one unit in one block, for launches with several units (a
// batch of documents, JSON Lines chunks): the steps of k_depth_keys /
// scan_hist_block / k_scatter / k_pairs with the block-level
// ranges replaced by one sequential pass per step.  Units that are well
// nested (paired by cjx_match.hip) are skipped.
template <bool U16>
__global__ __launch_bounds__(kP3Threads) void k_sort_unit(
    const uint8_t *__restrict__ oc_char, const uint32_t *__restrict__ oc_idx,
    uint16_t *__restrict__ keys, uint32_t *__restrict__ sorted, uint32_t *__restrict__ tmp,
    uint32_t *__restrict__ spos, UnitOut *__restrict__ uo, int32_t *__restrict__ pair_row,
    const LaunchState *__restrict__ ls, const uint32_t *__restrict__ slow_units) {
    constexpr int kW = kP3Threads / 64;
    __shared__ uint32_t h[2][256];
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[kW][256];
    __shared__ int32_t wsum[kW];
    __shared__ int32_t row[kPairChunk];
    __shared__ uint32_t s_m0, s_m1;
    if (ls->overflow) return;
    // The units the matching kernels found not well nested, one per block
    // at a time (the others need no sort and cost no block).
    const uint32_t n_slow = ls->n_slow;
    for (uint32_t q = blockIdx.x; q < n_slow; q += gridDim.x) {
    __syncthreads();  // the previous unit's LDS is no longer read
    const uint32_t u = slow_units[q];
    const uint32_t tid = threadIdx.x, w = tid >> 6;
    const uint64_t lt = lanemask_lt();
    const uint32_t kmask = U16 ? 0xFFFFu : 0xFFu;
    const uint32_t mb = uo[u].m_base, me = uo[u].m_end, M = me - mb;
    const uint32_t nb = uo[u].n_base, ne = uo[u].n_end;
    const uint32_t pbase = uo[u].pair_base + 1u;
    h[0][tid] = 0;
    h[1][tid] = 0;
    __syncthreads();

    // Depth keys (PSJ:1345-1358, :1537, :1540) and digit histograms.
    int32_t carry = 0;
    for (uint32_t base = mb; base < me; base += kP3Threads) {
        const uint32_t i = base + tid;
        const bool valid = i < me;
        const int32_t d = valid ? bracket_delta(oc_char[i]) : 0;
        const int32_t incl = wave_incl_scan_i(d);
        if (lane_id() == 63) wsum[w] = incl;
        __syncthreads();
        int32_t woff = 0, tot = 0;
        for (int k = 0; k < kW; k++) {
            if (k < (int)w) woff += wsum[k];
            tot += wsum[k];
        }
        if (valid) {
            const uint32_t depth = (uint32_t)(carry + woff + incl) & kmask;
            const uint32_t key = (d > 0) ? ((depth - 1u) & kmask) : depth;
            keys[i] = (uint16_t)key;
            atomicAdd(&h[0][key & 0xFFu], 1u);
            if (U16) atomicAdd(&h[1][key >> 8], 1u);
        }
        carry += tot;
        __syncthreads();
    }

    // Stable counting sort (PSJ:1546): low key byte, then (16-bit keys) the
    // high byte over the first pass's order.
    constexpr int kPasses = U16 ? 2 : 1;
    for (int pass = 0; pass < kPasses; pass++) {
        {
            const uint32_t v = h[pass][tid];
            const uint32_t in = wave_incl_scan(v);
            if (lane_id() == 63) wsum[w] = (int32_t)in;
            __syncthreads();
            uint32_t ex = in - v;
            for (uint32_t k = 0; k < w; k++) ex += (uint32_t)wsum[k];
            run[tid] = ex;
            __syncthreads();
        }
        const uint32_t shift = pass ? 8u : 0u;
        uint32_t *perm_out = (U16 && pass == 0) ? tmp : sorted;
        const bool final_pass = pass == kPasses - 1;
        for (uint32_t base = mb; base < me; base += kP3Threads) {
            const uint32_t e = base + tid;
            const bool valid = e < me;
            uint32_t m = 0, digit = 0;
            if (valid) {
                m = pass ? tmp[e] : e;
                digit = (keys[m] >> shift) & 0xFFu;
            }
            uint64_t peers = __ballot(valid);
#pragma unroll
            for (int bit = 0; bit < 8; bit++) {
                const uint32_t bv = (digit >> bit) & 1u;
                const uint64_t bm = __ballot(bv);
                peers &= bv ? bm : ~bm;
            }
#pragma unroll
            for (int k = 0; k < kW; k++) wcnt[k][tid] = 0;
            __syncthreads();
            const uint32_t rank = (uint32_t)__popcll(peers & lt);
            if (valid && rank == 0) wcnt[w][digit] = (uint32_t)__popcll(peers);
            __syncthreads();
            if (valid) {
                uint32_t pos = run[digit] + rank;
                for (uint32_t k = 0; k < w; k++) pos += wcnt[k][digit];
                perm_out[mb + pos] = m;
                if (final_pass) spos[m] = pos;
            }
            __syncthreads();
            uint32_t add = 0;
#pragma unroll
            for (int k = 0; k < kW; k++) add += wcnt[k][tid];
            run[tid] += add;
            __syncthreads();
        }
    }

    // Pairing (validate_expand_MathAPI_new2, PSJ:1431-1514) into dense
    // 4096-row tiles of the unit's pair row, as in k_pairs.
    const uint32_t last_group = M ? 4u * ((M + 3u) / 4u - 1u) : 0u;
    int32_t err = (tid == 0 && (M & 1u)) ? 1 : 0;
    for (uint32_t t0 = nb; t0 < ne; t0 += kPairChunk) {
        const uint32_t t1 = (t0 + kPairChunk < ne) ? t0 + kPairChunk : ne;
        for (uint32_t i = tid; i < kPairChunk; i += kP3Threads) row[i] = -1;
        if (tid < 2) {  // lower_bound of t0 (tid 0) and t1 (tid 1) in oc_idx[mb, me)
            const uint32_t key = tid ? t1 : t0;
            uint32_t lo = mb, hi = me;
            while (lo < hi) {
                const uint32_t mid = lo + ((hi - lo) >> 1);
                if (oc_idx[mid] < key) lo = mid + 1;
                else hi = mid;
            }
            if (tid) s_m1 = lo;
            else s_m0 = lo;
        }
        __syncthreads();
        for (uint32_t m = s_m0 + tid; m < s_m1; m += kP3Threads) {
            const uint32_t sp = spos[m];
            if (sp & 1u) continue;
            if (sp + 1u >= M) continue;
            if ((M & 3u) == 3u && sp == M - 3u) continue;
            const uint32_t m2 = sorted[mb + sp + 1u];
            const int c0 = (int)(signed char)oc_char[m];
            const int c1 = (int)(signed char)oc_char[m2];
            if ((c0 ^ c1) != 6) err = 1;
            if (sp < last_group && c0 > c1) err = 1;
            row[oc_idx[m] - t0] = (int32_t)(oc_idx[m2] - nb + pbase);
        }
        __syncthreads();
        for (uint32_t i = tid; i < t1 - t0; i += kP3Threads) pair_row[t0 + i] = row[i];
        __syncthreads();
    }
    err = block_reduce_sum(err, wsum);
    if (tid == 0 && err) atomicOr(&uo[u].struct_err, 1u);
    }
}

hipError_t launch_sort_units(hipStream_t s, const LaunchBufs &b, uint32_t n_units, bool depth_u16,
                             int32_t *pair_row) {
    if (n_units == 0) return hipSuccess;
    // The list's length is known on the device only: a grid that fills the
    // chip walks it.
    const uint32_t g = n_units < 2048u ? n_units : 2048u;
    if (depth_u16)
        hipLaunchKernelGGL(k_sort_unit<true>, dim3(g), dim3(kP3Threads), 0, s, b.oc_char,
                           b.oc_idx, b.keys, b.sorted, b.tmp, b.spos, b.uout, pair_row, b.ls,
                           b.slow_units);
    else
        hipLaunchKernelGGL(k_sort_unit<false>, dim3(g), dim3(kP3Threads), 0, s, b.oc_char,
                           b.oc_idx, b.keys, b.sorted, b.tmp, b.spos, b.uout, pair_row, b.ls,
                           b.slow_units);
    return hipGetLastError();
}

// k_rows_out: copies rows src[0, count) into host memory dst (a pinned
// buffer, written over the host link by the device; the host reads it after
// synchronizing with the stream).  dst need only
// be 4-byte aligned: rows before its first 16-byte boundary and after its
// last one go one by one, the rest as 16-byte stores, so each wave writes
// whole host cache lines.  It moves results out while the DMA engines carry
// input in.
__global__ __launch_bounds__(256) void k_rows_out(const uint32_t *__restrict__ src,
                                                  uint32_t *__restrict__ dst, uint64_t count) {
    const uint64_t tid = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * 256u;
    const uint64_t lead = (uint64_t)((16u - ((uintptr_t)dst & 15u)) & 15u) / 4u;
    const uint64_t head = lead < count ? lead : count;
    const uint64_t body = (count - head) / 4u;
    if (tid < head) dst[tid] = src[tid];
    const uint32_t *s = src + head;
    uint4 *d4 = reinterpret_cast<uint4 *>(dst + head);
    for (uint64_t i = tid; i < body; i += stride) {
        uint4 v;
        v.x = s[4 * i];
        v.y = s[4 * i + 1];
        v.z = s[4 * i + 2];
        v.w = s[4 * i + 3];
        d4[i] = v;
    }
    const uint64_t tail = head + 4 * body;
    if (tid < count - tail) dst[tail + tid] = src[tail + tid];
}

hipError_t launch_rows_out(hipStream_t s, const uint32_t *src, uint32_t *host_dst,
                           uint64_t count) {
    if (count == 0) return hipSuccess;
    uint64_t grid = (count / 4 + 255) / 256;
    grid = grid < 1 ? 1 : (grid > 128 ? 128 : grid);
    hipLaunchKernelGGL(k_rows_out, dim3((uint32_t)grid), dim3(256), 0, s, src, host_dst, count);
    return hipGetLastError();
}

hipError_t launch_reset(hipStream_t s, const LaunchBufs &b, uint32_t ntiles, uint32_t n_units) {
    const uint32_t work = 2 * ntiles > n_units ? 2 * ntiles : n_units;
    uint32_t grid = (work + 255) / 256;
    if (grid < 1) grid = 1;
    if (grid > 4096) grid = 4096;
    hipLaunchKernelGGL(k_reset, dim3(grid), dim3(256), 0, s, b.tile_st, ntiles, b.ls, b.uout,
                       n_units);
    return hipGetLastError();
}

hipError_t launch_reset_one(hipStream_t s, const LaunchBufs &b, const UnitDesc &u, uint32_t ntiles,
                           uint32_t n_small) {
    uint32_t grid = (2 * ntiles + 255) / 256;
    if (grid < 1) grid = 1;
    if (grid > 4096) grid = 4096;
    hipLaunchKernelGGL(k_reset_one, dim3(grid), dim3(256), 0, s, u, b.units, b.tile_unit, ntiles,
                       n_small, b.tile_st, b.ls, b.uout);
    return hipGetLastError();
}

hipError_t launch_assemble(hipStream_t s, const LaunchBufs &b, uint32_t n_units,
                           const uint32_t *d_struct, const int32_t *d_pair, int32_t *dst_struct,
                           int32_t *dst_pair, uint32_t N, bool host_layout) {
    const uint64_t threads = ((uint64_t)N + kAsmRows - 1) / kAsmRows;
    const uint32_t grid = (uint32_t)((threads + 255) / 256) + 1;
    int32_t *sent = nullptr;
    if (host_layout) {
        sent = dst_struct;
        dst_pair = dst_struct + N + 2;
        dst_struct = dst_struct + 1;
    }
    hipLaunchKernelGGL(k_assemble, dim3(grid), dim3(256), 0, s, b.uout, n_units, d_struct, d_pair,
                       dst_struct, dst_pair, N, sent);
    return hipGetLastError();
}

// Copies up to two byte ranges (16-byte aligned ends) with 16-byte loads and
// stores; either side may be pinned host memory, so small transfers run on
// the compute queue instead of waiting behind large DMA copies.
struct CopySeg {
    const uint8_t *src;
    uint8_t *dst;
    uint64_t n;
};

__global__ __launch_bounds__(256) void k_copy_segs(CopySeg a, CopySeg b) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const CopySeg g = k ? b : a;
        const uint64_t n16 = g.n >> 4;
        const uint4 *s = reinterpret_cast<const uint4 *>(g.src);
        uint4 *d = reinterpret_cast<uint4 *>(g.dst);
        for (uint64_t i = t0; i < n16; i += stride) d[i] = s[i];
        for (uint64_t i = (n16 << 4) + t0; i < g.n; i += stride) g.dst[i] = g.src[i];
    }
}

hipError_t launch_copy(hipStream_t s, const void *src0, void *dst0, uint64_t n0, const void *src1,
                       void *dst1, uint64_t n1) {
    const uint64_t words = ((n0 > n1 ? n0 : n1) + 15) / 16;
    // 64 blocks already run a host-link copy at full rate (and leave the CUs
    // to the parse kernels of other streams)
    uint64_t grid = (words + 255) / 256;
    if (grid < 1) grid = 1;
    if (grid > 64) grid = 64;
    CopySeg a{static_cast<const uint8_t *>(src0), static_cast<uint8_t *>(dst0), n0};
    CopySeg b{static_cast<const uint8_t *>(src1), static_cast<uint8_t *>(dst1), n1};
    hipLaunchKernelGGL(k_copy_segs, dim3((uint32_t)grid), dim3(256), 0, s, a, b);
    return hipGetLastError();
}

hipError_t launch_assemble_docs(hipStream_t s, const LaunchBufs &b, uint32_t n_units,
                                const uint32_t *d_struct, const int32_t *d_pair, int32_t *dst,
                                uint32_t N) {
    const uint64_t rows_threads = ((uint64_t)N + kAsmRows - 1) / kAsmRows;
    const uint64_t threads = rows_threads > n_units ? rows_threads : n_units;
    const uint32_t grid = (uint32_t)((threads + 255) / 256) + 1;
    hipLaunchKernelGGL(k_assemble_docs, dim3(grid), dim3(256), 0, s, b.uout, n_units, d_struct,
                       d_pair, dst, N);
    return hipGetLastError();
}

hipError_t launch_unit_bases(hipStream_t s, const LaunchBufs &b, uint32_t n_units, bool concat) {
    if (n_units <= kOneBlockUnits) {
        hipLaunchKernelGGL(k_unit_bases_one, dim3(1), dim3(kBaseThreads), 0, s, b.uout, n_units,
                           concat ? 1u : 0u, b.ls, b.chunk_unit, b.big_units);
        return hipGetLastError();
    }
    const uint32_t nb = (n_units + kBaseThreads - 1) / kBaseThreads;
    hipLaunchKernelGGL(k_ub_partial, dim3(nb), dim3(kBaseThreads), 0, s, b.uout, n_units,
                       b.unit_part);
    hipLaunchKernelGGL(k_ub_scan, dim3(1), dim3(kBaseThreads), 0, s, b.unit_part, nb, b.ls);
    hipLaunchKernelGGL(k_ub_final, dim3(nb), dim3(kBaseThreads), 0, s, b.uout, n_units,
                       concat ? 1u : 0u, b.unit_part, b.ls, b.chunk_unit, b.big_units);
    hipLaunchKernelGGL(k_chunk_map, dim3(n_units < 256u ? n_units : 256u), dim3(256), 0, s,
                       b.uout, b.big_units, b.chunk_unit, b.ls);
    return hipGetLastError();
}

hipError_t launch_phase3(hipStream_t s, const LaunchBufs &b, uint32_t u, uint32_t G,
                         uint32_t pair_grid, bool depth_u16, int32_t *pair_row) {
    const uint32_t kmask = depth_u16 ? 0xFFFFu : 0xFFu;
    hipLaunchKernelGGL(k_depth_keys, dim3(G), dim3(kP3Threads), 0, s, b.oc_char, b.uout, u, kmask,
                       b.ck, b.keys, b.hist, b.ls);
    if (!depth_u16) {
        hipLaunchKernelGGL((k_scatter<true, true>), dim3(G), dim3(kP3Threads), 0, s, b.keys,
                           b.uout, u, 0u, b.hist, (const uint32_t *)nullptr, b.sorted, b.spos,
                           b.ls);
    } else {
        hipLaunchKernelGGL((k_scatter<true, false>), dim3(G), dim3(kP3Threads), 0, s, b.keys,
                           b.uout, u, 0u, b.hist, (const uint32_t *)nullptr, b.tmp, b.spos,
                           b.ls);
        hipLaunchKernelGGL(k_hist, dim3(G), dim3(kP3Threads), 0, s, b.keys, b.uout, u, b.tmp, 8u,
                           b.hist, b.ls);
        hipLaunchKernelGGL((k_scatter<false, true>), dim3(G), dim3(kP3Threads), 0, s, b.keys,
                           b.uout, u, 8u, b.hist, b.tmp, b.sorted, b.spos, b.ls);
    }
    hipLaunchKernelGGL(k_pairs, dim3(pair_grid), dim3(kP3Threads), 0, s, b.oc_char, b.oc_idx,
                       b.sorted, b.spos, b.uout, u, pair_row, b.ls);
    return hipGetLastError();
}

}  // namespace cjx
