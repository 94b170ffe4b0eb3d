// cjx_kernels.hip - hand-written gfx950 kernels of the three-phase JSON parse.
//
// Phase 1+2 run fused in one pass over the input (k_tokenize): every byte is
// read once from HBM, classified in registers, and only the compact outputs
// (structural positions, bracket chars and indices) are written back.  The
// sequential dependencies of the reference's separate kernels and Thrust scans
// (cujson/parse_standard_json.cu:446-1343) are carried by a decoupled
// look-back over tiles:
//   * escape / string state as an affine map over GF(2) per tile (4 bits),
//   * structural and bracket counts as one packed 64-bit prefix sum.
// Phase 3 (parse_standard_json.cu:1345-1582) runs over the compact bracket
// array: a segmented +1/-1 depth scan (k_depth_totals, k_depth_keys), a
// stable LDS counting sort by depth key (k_scan_hist, k_scatter), and a
// dense pair-row writer that applies the reference's group-of-four pairing
// rule (k_pairs).
#include "cjx_internal.h"

namespace cjx {

// ---------------------------------------------------------------------------
// Wave helpers (wave64).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << __lane_id()) - 1ull; }

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    return x;
}

__device__ __forceinline__ int32_t wave_incl_scan_i(int32_t x) {
    return (int32_t)wave_incl_scan((uint32_t)x);
}

__device__ __forceinline__ uint32_t wave_read(uint32_t x, int lane) { return __shfl(x, lane, 64); }

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
// Byte classification (SWAR over 32-bit words).
// ---------------------------------------------------------------------------

// High bit of each byte set iff that byte of t is zero (exact, no borrow leak).
__device__ __forceinline__ uint32_t zero_bytes_hi(uint32_t t) {
    return ~(((t & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | t) & 0x80808080u;
}
// Gather bits 7,15,23,31 into bits 0..3.
__device__ __forceinline__ uint32_t hi_to_nib(uint32_t h) { return (h * 0x00204081u) >> 28; }

// 16-bit masks of one 16-byte segment: bit j <-> byte j.  The character sets
// are those of bitMapCreatorSimd (parse_standard_json.cu:473-490; JSON Lines
// adds '\n', parse_json_lines.cu:453-456).
template <bool JSONL>
__device__ __forceinline__ void classify16(const uint4 v, uint32_t &bs, uint32_t &q, uint32_t &op,
                                           uint32_t &oc) {
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
    bs = q = op = oc = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint32_t x = w[i];
        const uint32_t x20 = x | 0x20202020u;  // folds '[' ']' onto '{' '}'
        const uint32_t hbs = zero_bytes_hi(x ^ 0x5C5C5C5Cu);
        const uint32_t hq = zero_bytes_hi(x ^ 0x22222222u);
        const uint32_t hoc = zero_bytes_hi(x20 ^ 0x7B7B7B7Bu) | zero_bytes_hi(x20 ^ 0x7D7D7D7Du);
        uint32_t hop = hoc | zero_bytes_hi(x ^ 0x3A3A3A3Au) | zero_bytes_hi(x ^ 0x2C2C2C2Cu);
        if (JSONL) hop |= zero_bytes_hi(x ^ 0x0A0A0A0Au);
        bs |= hi_to_nib(hbs) << (4 * i);
        q |= hi_to_nib(hq) << (4 * i);
        op |= hi_to_nib(hop) << (4 * i);
        oc |= hi_to_nib(hoc) << (4 * i);
    }
}

// Reference UTF-8 check of one 32-bit word against its predecessor
// (checkUTF8, parse_standard_json.cu:237-257, reduced to what survives the
// flag reassignment at :122-123): for a word without high bits, the previous
// word must not end in an unfinished lead (:219-223, :239-242); otherwise a
// byte errs iff [prev1 and byte are continuations] != [prev2 >= E0 or prev3 >= F0].
__device__ __forceinline__ uint32_t utf8_word_err(uint32_t cur, uint32_t prev) {
    if ((cur & 0x80808080u) == 0) {
        return (prev >= 0xC0000000u) | (((prev >> 16) & 0xE0u) == 0xE0u) |
               (((prev >> 8) & 0xF0u) == 0xF0u);
    }
    const uint32_t cc = cur & ~(cur << 1) & 0x80808080u;
    const uint32_t cp = prev & ~(prev << 1) & 0x80808080u;
    const uint32_t ec = cur & (cur << 1) & (cur << 2) & 0x80808080u;
    const uint32_t ep = prev & (prev << 1) & (prev << 2) & 0x80808080u;
    const uint32_t fc = ec & (cur << 3);
    const uint32_t fp = ep & (prev << 3);
    const uint32_t p1 = (uint32_t)((((uint64_t)cc << 32) | cp) >> 24);
    const uint32_t p2 = (uint32_t)((((uint64_t)ec << 32) | ep) >> 16);
    const uint32_t p3 = (uint32_t)((((uint64_t)fc << 32) | fp) >> 8);
    return ((p1 & cc) ^ (p2 | p3)) & 0x80808080u;
}

// Unescaped quotes of a 16-bit segment given the parity `ov` of the backslash
// run ending just before it (findEscapedQuoteMerge_NEW,
// parse_standard_json.cu:820-830, on 16-bit words).
__device__ __forceinline__ uint32_t unescaped16(uint32_t bs, uint32_t q, uint32_t ov) {
    const uint32_t even = 0x5555u, odd = 0xAAAAu;
    bs &= ~ov;
    const uint32_t app = ((bs << 1) | ov) & 0xFFFFu;
    const uint32_t oddseq = bs & odd & ~app;
    const uint32_t seq = oddseq + bs;
    const uint32_t inv = (seq << 1) & 0xFFFFu;
    const uint32_t esc = (even ^ inv) & app;
    return q & ~esc & 0xFFFFu;
}

// Inclusive prefix xor over 16 bits (prefix_xor, cujson/utils.cu:141-149).
__device__ __forceinline__ uint32_t prefix_xor16(uint32_t x) {
    x ^= x << 1;
    x ^= x << 2;
    x ^= x << 4;
    x ^= x << 8;
    return x & 0xFFFFu;
}

__device__ __forceinline__ uint4 load_segment(const uint8_t *base, uint64_t o, uint64_t n) {
    if (o >= n) return make_uint4(0, 0, 0, 0);
    uint4 r = *reinterpret_cast<const uint4 *>(base + o);
    const uint64_t rem = n - o;
    if (rem < 16) {
        uint32_t w[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int64_t keep = (int64_t)rem - 4 * k;
            if (keep <= 0) w[k] = 0;
            else if (keep < 4) w[k] &= (1u << (8 * keep)) - 1u;
        }
        r = make_uint4(w[0], w[1], w[2], w[3]);
    }
    return r;
}

__device__ __forceinline__ uint32_t seg_byte(const uint4 v, int j) {
    const uint32_t w = (j < 8) ? ((j < 4) ? v.x : v.y) : ((j < 12) ? v.z : v.w);
    return (w >> (8 * (j & 3))) & 0xFFu;
}

// ---------------------------------------------------------------------------
// Escape / string carry as an affine map over GF(2).
// A span of input maps the state (c = parity of the backslash run ending at
// its start, s = in-string parity at its start) to the state at its end:
//   c' = a.c ^ b,   s' = s ^ q ^ d.c
// a: the span is all backslashes; b: its trailing run parity; q: parity of
// its unescaped quotes computed with c = s = 0; d: its first non-backslash
// byte is a quote (only that quote's escape status depends on c).
// Bits: a=1, b=2, d=4, q=8.
// ---------------------------------------------------------------------------
constexpr uint32_t kMapIdentity = 1u;

__device__ __forceinline__ uint32_t map_compose(uint32_t t1, uint32_t t2) {  // t1 then t2
    const uint32_t a1 = t1 & 1, b1 = (t1 >> 1) & 1, d1 = (t1 >> 2) & 1, q1 = (t1 >> 3) & 1;
    const uint32_t a2 = t2 & 1, b2 = (t2 >> 1) & 1, d2 = (t2 >> 2) & 1, q2 = (t2 >> 3) & 1;
    const uint32_t a = a1 & a2, b = (a2 & b1) ^ b2, d = d1 ^ (d2 & a1), q = q1 ^ q2 ^ (d2 & b1);
    return a | (b << 1) | (d << 2) | (q << 3);
}
__device__ __forceinline__ uint32_t map_apply(uint32_t t, uint32_t st) {  // st: c=1, s=2
    const uint32_t a = t & 1, b = (t >> 1) & 1, d = (t >> 2) & 1, q = (t >> 3) & 1;
    const uint32_t c = st & 1, s = (st >> 1) & 1;
    return ((a & c) ^ b) | ((s ^ q ^ (d & c)) << 1);
}

// ---------------------------------------------------------------------------
// k_reset: clears look-back state and per-unit results before a parse.
// ---------------------------------------------------------------------------
__global__ void k_reset(uint32_t *carry_st, unsigned long long *count_st, uint32_t ntiles,
                        LaunchState *ls, UnitOut *uo, uint32_t n_units) {
    const uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = i0; i < ntiles; i += stride) {
        carry_st[i] = 0;
        count_st[i] = 0;
    }
    for (uint32_t i = i0; i < n_units; i += stride) uo[i] = UnitOut{};
    if (i0 == 0) *ls = LaunchState{};
}

// ---------------------------------------------------------------------------
// k_tokenize: fused phase 1 + phase 2 over 16 KiB tiles.
// ---------------------------------------------------------------------------
template <bool JSONL>
__global__ __launch_bounds__(kTokThreads) void k_tokenize(
    const uint8_t *__restrict__ in, const UnitDesc *__restrict__ units, uint32_t n_units,
    uint32_t *__restrict__ carry_st, unsigned long long *__restrict__ count_st,
    LaunchState *__restrict__ ls, UnitOut *__restrict__ uo, uint32_t *__restrict__ out_struct,
    uint32_t n_cap, uint8_t *__restrict__ oc_char, uint32_t *__restrict__ oc_idx,
    uint32_t m_cap) {
    __shared__ uint32_t s_tile, s_unit;
    __shared__ uint32_t s_map[kWaves];
    __shared__ uint32_t s_seff[kWaves];
    __shared__ uint32_t s_wcnt[kWaves];
    __shared__ uint32_t s_nexcl, s_mexcl;

    const uint32_t tid = threadIdx.x;
    const uint32_t w = tid >> 6, lane = tid & 63;
    const uint64_t lt = lanemask_lt();

    if (tid == 0) {
        const uint32_t t = atomicAdd(&ls->tile_counter, 1u);
        uint32_t lo = 0, hi = n_units - 1;  // last unit with tile0 <= t
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (units[mid].tile0 <= t) lo = mid;
            else hi = mid - 1;
        }
        s_tile = t;
        s_unit = lo;
    }
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t u = s_unit;
    const UnitDesc U = units[u];
    const uint32_t tile_in_unit = tile - U.tile0;
    const uint8_t *ubase = in + U.in_off;
    const uint64_t wave_off = (uint64_t)tile_in_unit * kTileBytes + (uint64_t)w * kWaveBytes;

    uint4 v[kRounds];
#pragma unroll
    for (int i = 0; i < kRounds; i++)
        v[i] = load_segment(ubase, wave_off + (uint64_t)i * 1024 + (uint64_t)lane * 16, U.n);

    // Dword preceding the wave, for the UTF-8 check of its first word.
    uint32_t prev_dw = 0;
    if (wave_off > 0 && wave_off < U.n)
        prev_dw = *reinterpret_cast<const uint32_t *>(ubase + wave_off - 4);

    // ---- local pass: masks with the wave entered at (c = 0, s = 0) -----------
    uint64_t instr64 = 0, op64 = 0, oc64 = 0;
    uint32_t c_run = 0, s_run = 0, seen = 0, dflag = 0;
    uint32_t utf8_bad = 0;
    uint32_t last_dw = prev_dw;  // last dword of the previous round (wave-uniform)
#pragma unroll
    for (int i = 0; i < kRounds; i++) {
        uint32_t bs, q, op, oc;
        classify16<JSONL>(v[i], bs, q, op, oc);

        // Phase 1: UTF-8 over the segment's four words (words at or past n are
        // outside W4 and not checked, parse_standard_json.cu:236).
        {
            const uint32_t up = __shfl_up(v[i].w, 1, 64);
            const uint32_t p0 = (lane == 0) ? last_dw : up;
            const uint64_t seg_off = wave_off + (uint64_t)i * 1024 + (uint64_t)lane * 16;
            const uint32_t anyhigh = ((v[i].x | v[i].y | v[i].z | v[i].w | p0) & 0x80808080u) != 0;
            if (__ballot(anyhigh)) {
                uint32_t e = 0;
                if (seg_off < U.n) e |= utf8_word_err(v[i].x, p0);
                if (seg_off + 4 < U.n) e |= utf8_word_err(v[i].y, v[i].x);
                if (seg_off + 8 < U.n) e |= utf8_word_err(v[i].z, v[i].y);
                if (seg_off + 12 < U.n) e |= utf8_word_err(v[i].w, v[i].z);
                utf8_bad |= (e != 0);
            }
            last_dw = wave_read(v[i].w, 63);
        }

        // Backslash-run carry across lanes: the run entering lane l ends in the
        // nearest lower lane that is not all backslashes (16 is even, so
        // all-backslash lanes do not change the parity).
        const uint32_t all_bs = (bs == 0xFFFFu);
        const uint32_t run = __clz(~(bs << 16));
        const uint64_t NB = __ballot(!all_bs);
        const uint64_t TP = __ballot(run & 1u);
        const uint64_t mm = NB & lt;
        const uint32_t ov = mm ? (uint32_t)((TP >> (63 - __clzll(mm))) & 1ull) : c_run;
        const uint32_t rq = unescaped16(bs, q, ov);

        // In-string parity across lanes.
        const uint32_t px = prefix_xor16(rq);
        const uint64_t QP = __ballot(__popc(rq) & 1u);
        const uint32_t s_lane = s_run ^ (__popcll(QP & lt) & 1u);
        const uint32_t instr = (px ^ (s_lane ? 0xFFFFu : 0u)) & 0xFFFFu;

        if (!seen && NB) {  // first non-backslash byte of the wave: is it a quote?
            const uint32_t k0 = (uint32_t)__ffsll((long long)NB) - 1;
            const uint32_t nbpos = (uint32_t)__ffs(~bs & 0xFFFFu) - 1;
            const uint32_t fq = all_bs ? 0u : ((q >> nbpos) & 1u);
            const uint64_t FQ = __ballot(fq);
            dflag = (uint32_t)((FQ >> k0) & 1ull);
            seen = 1;
        }
        if (NB) c_run = (uint32_t)((TP >> (63 - __clzll(NB))) & 1ull);
        s_run ^= (__popcll(QP) & 1u);

        instr64 |= (uint64_t)instr << (16 * i);
        op64 |= (uint64_t)op << (16 * i);
        oc64 |= (uint64_t)oc << (16 * i);
    }

    const uint64_t bad_lanes = __ballot(utf8_bad);
    if (lane == 0) {
        if (bad_lanes) atomicOr(&uo[u].utf8_err, 1u);
        const uint32_t a = seen ? 0u : 1u;
        s_map[w] = a | (c_run << 1) | (dflag << 2) | (s_run << 3);
    }
    __syncthreads();

    // ---- carry look-back over tiles (restarts at each unit) ------------------
    if (tid == 0) {
        uint32_t T = kMapIdentity;
        for (int k = 0; k < kWaves; k++) T = map_compose(T, s_map[k]);
        uint32_t entry;  // (c, s) entering the tile
        if (tile_in_unit == 0) {
            entry = 0;
            agent_store(&carry_st[tile], (kStInc << 30) | map_apply(T, 0));
        } else {
            agent_store(&carry_st[tile], (kStAgg << 30) | T);
            uint32_t acc = kMapIdentity;
            uint32_t j = tile - 1;
            for (;;) {
                uint32_t st;
                while (((st = agent_load(&carry_st[j])) >> 30) == kStInvalid)
                    __builtin_amdgcn_s_sleep(1);
                if ((st >> 30) == kStInc) {
                    entry = map_apply(acc, st & 3u);
                    break;
                }
                acc = map_compose(st & 0xFu, acc);
                j--;
            }
            agent_store(&carry_st[tile], (kStInc << 30) | map_apply(T, entry));
        }
        uint32_t state = entry;
        for (int k = 0; k < kWaves; k++) {
            const uint32_t m = s_map[k];
            s_seff[k] = ((state >> 1) & 1u) ^ ((state & 1u) & ((m >> 2) & 1u));
            state = map_apply(m, state);
        }
    }
    __syncthreads();

    // ---- select structurals, count ------------------------------------------
    const uint64_t fix = s_seff[w] ? ~0ull : 0ull;
    instr64 ^= fix;
    const uint64_t st64 = op64 & ~instr64;
    const uint64_t ocm64 = oc64 & ~instr64;
    uint32_t excl[kRounds], rtot[kRounds];
    uint32_t wtot = 0;
#pragma unroll
    for (int i = 0; i < kRounds; i++) {
        const uint32_t sti = (uint32_t)(st64 >> (16 * i)) & 0xFFFFu;
        const uint32_t oci = (uint32_t)(ocm64 >> (16 * i)) & 0xFFFFu;
        const uint32_t cnt = (uint32_t)__popc(sti) | ((uint32_t)__popc(oci) << 16);
        const uint32_t incl = wave_incl_scan(cnt);
        excl[i] = incl - cnt;
        rtot[i] = wave_read(incl, 63);
        wtot += rtot[i];
    }
    if (lane == 0) s_wcnt[w] = wtot;
    __syncthreads();

    // ---- count look-back over all tiles of the launch ------------------------
    if (tid == 0) {
        uint32_t nt = 0, mt = 0;
        for (int k = 0; k < kWaves; k++) {
            nt += s_wcnt[k] & 0xFFFFu;
            mt += s_wcnt[k] >> 16;
        }
        const unsigned long long agg = ((unsigned long long)nt << 31) | mt;
        const unsigned long long fl_mask = 3ull << 62;
        unsigned long long excl_sum = 0;
        if (tile == 0) {
            agent_store64(&count_st[tile], ((unsigned long long)kStInc << 62) | agg);
        } else {
            agent_store64(&count_st[tile], ((unsigned long long)kStAgg << 62) | agg);
            uint32_t j = tile - 1;
            for (;;) {
                unsigned long long st;
                while (((st = agent_load64(&count_st[j])) >> 62) == kStInvalid)
                    __builtin_amdgcn_s_sleep(1);
                excl_sum += st & ~fl_mask;
                if ((st >> 62) == kStInc) break;
                j--;
            }
            agent_store64(&count_st[tile], ((unsigned long long)kStInc << 62) | (excl_sum + agg));
        }
        const uint32_t nex = (uint32_t)(excl_sum >> 31);
        const uint32_t mex = (uint32_t)(excl_sum & 0x7FFFFFFFull);
        s_nexcl = nex;
        s_mexcl = mex;
        if (tile_in_unit == 0) {
            uo[u].n_base = nex;
            uo[u].m_base = mex;
        }
        if (tile_in_unit == U.ntiles - 1) {
            uo[u].n_end = nex + nt;
            uo[u].m_end = mex + mt;
        }
        if (tile == gridDim.x - 1) {
            ls->n_total = nex + nt;
            ls->m_total = mex + mt;
        }
    }
    __syncthreads();

    // ---- compaction: write positions and the bracket arrays ------------------
    uint32_t nb = s_nexcl, mb = s_mexcl;
    for (uint32_t k = 0; k < w; k++) {
        nb += s_wcnt[k] & 0xFFFFu;
        mb += s_wcnt[k] >> 16;
    }
    uint32_t overflow = 0;
#pragma unroll
    for (int i = 0; i < kRounds; i++) {
        uint32_t sti = (uint32_t)(st64 >> (16 * i)) & 0xFFFFu;
        const uint32_t oci = (uint32_t)(ocm64 >> (16 * i)) & 0xFFFFu;
        uint32_t idx = nb + (excl[i] & 0xFFFFu);
        uint32_t midx = mb + (excl[i] >> 16);
        const uint64_t pos0 =
            U.pos_base + wave_off + (uint64_t)i * 1024 + (uint64_t)lane * 16 + 1;
        while (sti) {
            const int j = __ffs(sti) - 1;
            sti &= sti - 1;
            if (idx < n_cap) out_struct[idx] = (uint32_t)(pos0 + (uint64_t)j);
            else overflow = 1;
            if ((oci >> j) & 1u) {
                if (midx < m_cap) {
                    oc_char[midx] = (uint8_t)seg_byte(v[i], j);
                    oc_idx[midx] = idx;
                } else {
                    overflow = 1;
                }
                midx++;
            }
            idx++;
        }
        nb += rtot[i] & 0xFFFFu;
        mb += rtot[i] >> 16;
    }
    if (__ballot(overflow) && lane == 0) atomicOr(&ls->overflow, 1u);
}

// ---------------------------------------------------------------------------
// Phase 3 helpers.  All phase-3 kernels split the unit's bracket range
// [m_base, m_end) into gridDim.x contiguous block ranges the same way.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void block_range(uint32_t base, uint32_t end, uint32_t b, uint32_t G,
                                            uint32_t &b0, uint32_t &b1) {
    const uint32_t M = end - base;
    const uint32_t chunk = (M + G - 1) / G;
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

// Sum of +1/-1 per block (depth_init_MathAPI, parse_standard_json.cu:1345-1358).
__global__ __launch_bounds__(kP3Threads) void k_depth_totals(const uint8_t *__restrict__ oc_char,
                                                             const UnitOut *__restrict__ uo,
                                                             uint32_t u,
                                                             int32_t *__restrict__ block_delta,
                                                             const LaunchState *__restrict__ ls) {
    __shared__ int32_t red[kP3Threads / 64];
    if (ls->overflow) return;  // capacity exceeded: ranges are not backed by buffers
    uint32_t b0, b1;
    block_range(uo[u].m_base, uo[u].m_end, blockIdx.x, gridDim.x, b0, b1);
    int32_t s = 0;
    for (uint32_t i = b0 + threadIdx.x; i < b1; i += kP3Threads) s += bracket_delta(oc_char[i]);
    s = block_reduce_sum(s, red);
    if (threadIdx.x == 0) block_delta[blockIdx.x] = s;
}

// Depth keys (inclusive scan modulo 2^8 or 2^16, parse_standard_json.cu:1537;
// openers take depth-1, :1540 / utils.cu:104-125) and the histogram of the
// low key byte per block.
__global__ __launch_bounds__(kP3Threads) void k_depth_keys(
    const uint8_t *__restrict__ oc_char, const UnitOut *__restrict__ uo, uint32_t u,
    uint32_t kmask, const int32_t *__restrict__ block_delta, uint16_t *__restrict__ keys,
    uint32_t *__restrict__ hist, const LaunchState *__restrict__ ls) {
    __shared__ int32_t red[kP3Threads / 64];
    if (ls->overflow) return;
    __shared__ int32_t wsum[kP3Threads / 64];
    __shared__ uint32_t h[256];
    const uint32_t tid = threadIdx.x, w = tid >> 6;
    uint32_t b0, b1;
    block_range(uo[u].m_base, uo[u].m_end, blockIdx.x, gridDim.x, b0, b1);
    int32_t pre = 0;
    for (uint32_t k = tid; k < blockIdx.x; k += kP3Threads) pre += block_delta[k];
    pre = block_reduce_sum(pre, red);
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
}

// Histogram of key byte `shift` over elements taken in the order perm_in
// (second pass of the 16-bit sort).
__global__ __launch_bounds__(kP3Threads) void k_hist(const uint16_t *__restrict__ keys,
                                                     const UnitOut *__restrict__ uo, uint32_t u,
                                                     const uint32_t *__restrict__ perm_in,
                                                     uint32_t shift, uint32_t *__restrict__ hist,
                                                     const LaunchState *__restrict__ ls) {
    __shared__ uint32_t h[256];
    if (ls->overflow) return;
    const uint32_t tid = threadIdx.x;
    uint32_t b0, b1;
    block_range(uo[u].m_base, uo[u].m_end, blockIdx.x, gridDim.x, b0, b1);
    h[tid] = 0;
    __syncthreads();
    for (uint32_t e = b0 + tid; e < b1; e += kP3Threads)
        atomicAdd(&h[(keys[perm_in[e]] >> shift) & 0xFFu], 1u);
    __syncthreads();
    hist[blockIdx.x * 256 + tid] = h[tid];
}

// Exclusive offsets per (block, digit): digit-major, block-minor, so that a
// stable scatter reproduces thrust::stable_sort_by_key (parse_standard_json.cu:1546).
__global__ __launch_bounds__(256) void k_scan_hist(uint32_t *__restrict__ hist, uint32_t G,
                                                  const LaunchState *__restrict__ ls) {
    __shared__ uint32_t wsum[4];
    if (ls->overflow) return;
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
    if (ls->overflow) return;
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
    UnitOut *__restrict__ uo, uint32_t u, int32_t *__restrict__ pair_row, uint32_t local_pairs,
    const LaunchState *__restrict__ ls) {
    __shared__ int32_t row[kPairChunk];
    if (ls->overflow) return;
    __shared__ uint32_t s_m0, s_m1;
    __shared__ int32_t red[kP3Threads / 64];
    const uint32_t tid = threadIdx.x;
    const uint32_t nb = uo[u].n_base, ne = uo[u].n_end;
    const uint32_t mb = uo[u].m_base, me = uo[u].m_end;
    const uint32_t N = ne - nb, M = me - mb;
    const uint32_t last_group = M ? 4u * ((M + 3u) / 4u - 1u) : 0u;
    const uint32_t sub = local_pairs ? nb : 0u;
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
            row[oc_idx[m] - t0] = (int32_t)(oc_idx[m2] + 1u - sub);
        }
        __syncthreads();
        for (uint32_t i = tid; i < t1 - t0; i += kP3Threads) pair_row[t0 + i] = row[i];
        __syncthreads();
    }
    if (blockIdx.x == 0 && tid == 0 && (M & 1u)) err = 1;
    err = block_reduce_sum(err, red);
    if (tid == 0 && err) atomicOr(&uo[u].struct_err, 1u);
}

// Explicit instantiations used by the host driver.
template __global__ void k_tokenize<false>(const uint8_t *, const UnitDesc *, uint32_t, uint32_t *,
                                           unsigned long long *, LaunchState *, UnitOut *,
                                           uint32_t *, uint32_t, uint8_t *, uint32_t *, uint32_t);
template __global__ void k_tokenize<true>(const uint8_t *, const UnitDesc *, uint32_t, uint32_t *,
                                          unsigned long long *, LaunchState *, UnitOut *,
                                          uint32_t *, uint32_t, uint8_t *, uint32_t *, uint32_t);

}  // namespace cjx

// ---------------------------------------------------------------------------
// Host-side launchers (the only entry points the driver uses).
// ---------------------------------------------------------------------------
namespace cjx {

hipError_t launch_reset(hipStream_t s, const LaunchBufs &b, uint32_t ntiles, uint32_t n_units) {
    const uint32_t work = ntiles > n_units ? ntiles : n_units;
    uint32_t grid = (work + 255) / 256;
    if (grid < 1) grid = 1;
    if (grid > 4096) grid = 4096;
    hipLaunchKernelGGL(k_reset, dim3(grid), dim3(256), 0, s, b.carry_st, b.count_st, ntiles, b.ls,
                       b.uout, n_units);
    return hipGetLastError();
}

hipError_t launch_tokenize(hipStream_t s, const LaunchBufs &b, const uint8_t *d_in,
                           uint32_t ntiles, uint32_t n_units, bool jsonl, uint32_t *out_struct,
                           uint32_t n_cap, uint32_t m_cap) {
    if (jsonl)
        hipLaunchKernelGGL(k_tokenize<true>, dim3(ntiles), dim3(kTokThreads), 0, s, d_in, b.units,
                           n_units, b.carry_st, b.count_st, b.ls, b.uout, out_struct, n_cap,
                           b.oc_char, b.oc_idx, m_cap);
    else
        hipLaunchKernelGGL(k_tokenize<false>, dim3(ntiles), dim3(kTokThreads), 0, s, d_in,
                           b.units, n_units, b.carry_st, b.count_st, b.ls, b.uout, out_struct,
                           n_cap, b.oc_char, b.oc_idx, m_cap);
    return hipGetLastError();
}

hipError_t launch_phase3(hipStream_t s, const LaunchBufs &b, uint32_t u, uint32_t G,
                         uint32_t pair_grid, bool depth_u16, int32_t *pair_row,
                         bool local_pairs) {
    const uint32_t kmask = depth_u16 ? 0xFFFFu : 0xFFu;
    hipLaunchKernelGGL(k_depth_totals, dim3(G), dim3(kP3Threads), 0, s, b.oc_char, b.uout, u,
                       b.block_delta, b.ls);
    hipLaunchKernelGGL(k_depth_keys, dim3(G), dim3(kP3Threads), 0, s, b.oc_char, b.uout, u, kmask,
                       b.block_delta, b.keys, b.hist, b.ls);
    hipLaunchKernelGGL(k_scan_hist, dim3(1), dim3(256), 0, s, b.hist, G, b.ls);
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
        hipLaunchKernelGGL(k_scan_hist, dim3(1), dim3(256), 0, s, b.hist, G, b.ls);
        hipLaunchKernelGGL((k_scatter<false, true>), dim3(G), dim3(kP3Threads), 0, s, b.keys,
                           b.uout, u, 8u, b.hist, b.tmp, b.sorted, b.spos, b.ls);
    }
    hipLaunchKernelGGL(k_pairs, dim3(pair_grid), dim3(kP3Threads), 0, s, b.oc_char, b.oc_idx,
                       b.sorted, b.spos, b.uout, u, pair_row, local_pairs ? 1u : 0u, b.ls);
    return hipGetLastError();
}

}  // namespace cjx
