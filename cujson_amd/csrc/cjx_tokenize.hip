// cjx_tokenize.hip - fused phase 1 + phase 2 (UTF-8 check and tokenization).
//
// One pass over the input: every byte is read once from HBM, classified in
// registers, and only the compact outputs (structural positions, bracket chars
// and bracket structural indices) are written back.  The sequential
// dependencies that the reference spreads over separate kernels and Thrust
// scans (cujson/parse_standard_json.cu:446-1343) are carried by
//   * in-lane recurrences on 32-bit masks,
//   * ballots across the 64 lanes of a wave,
//   * a decoupled look-back across 64 KiB tiles: the escape/string state as an
//     affine map over GF(2) (4 bits), and the structural and bracket counts
//     as one packed 64-bit prefix sum.
//
// Lane layout: each lane owns 32 contiguous input bytes per round.  The eight
// words are byte-transposed with v_perm (word d, byte b <- input byte 8b+d),
// so that shifting one class bit per word and OR-ing gives the linear 32-bit
// mask, and so that "previous byte" is the previous word in the UTF-8 rule.
//
// Classification is three v_perm table lookups ANDed together, indexed by
// byte bits 0-2, 3-5 and 6-7.  Every class needed is a product set in those
// three coordinates, so the AND is exact:
//   bit 0 ':' (and '\n' for JSON Lines)   bit 4 '\\'
//   bit 1 CONT (0x80-0xBF)                 bit 5 E    (>= 0xE0)
//   bit 2 '[' ']' '{' '}'                  bit 6 F    (>= 0xF0)
//   bit 3 ','                              bit 7 '"'
// Once the UTF-8 rule has used CONT and E, bits 1 and 5 are replaced by the
// raw byte's bits 1 (set for '[' and '{', clear for ']' and '}') and 5 (set
// for '{' and '}'), and one 8x8 bit-matrix transpose per byte slot turns the
// eight class words into the eight 32-bit class masks at once.
#include "cjx_device.h"

namespace cjx {

// 8x8 bit-matrix transpose in every byte slot: afterwards t[K] holds, in
// byte b, bit d = bit K of byte b of the input t[d].  With transposed input
// words (byte b of word d = input byte 8b+d) t[K] is the linear 32-bit mask
// (bit i <-> byte i of the lane) of class bit K.  Three butterfly stages
// swap the off-diagonal 4x4, 2x2 and 1x1 blocks.
__host__ __device__ __forceinline__ void transpose8x8(uint32_t (&t)[8]) {
#pragma unroll
    for (int d = 0; d < 4; d++) {
        const uint32_t a = t[d], b = t[d + 4];
        t[d] = (a & 0x0F0F0F0Fu) | ((b << 4) & 0xF0F0F0F0u);
        t[d + 4] = ((a >> 4) & 0x0F0F0F0Fu) | (b & 0xF0F0F0F0u);
    }
#pragma unroll
    for (int d = 0; d < 8; d += (d & 1) ? 3 : 1) {  // 0, 1, 4, 5
        const uint32_t a = t[d], b = t[d + 2];
        t[d] = (a & 0x33333333u) | ((b << 2) & 0xCCCCCCCCu);
        t[d + 2] = ((a >> 2) & 0x33333333u) | (b & 0xCCCCCCCCu);
    }
#pragma unroll
    for (int d = 0; d < 8; d += 2) {
        const uint32_t a = t[d], b = t[d + 1];
        t[d] = (a & 0x55555555u) | ((b << 1) & 0xAAAAAAAAu);
        t[d + 1] = ((a >> 1) & 0x55555555u) | (b & 0xAAAAAAAAu);
    }
}

// Unescaped quotes given the parity `ov` of the backslash run ending just
// before the lane (findEscapedQuoteMerge_NEW, parse_standard_json.cu:820-830).
__device__ __forceinline__ uint32_t unescaped32(uint32_t bs, uint32_t q, uint32_t ov) {
    const uint32_t even = 0x55555555u, odd = 0xAAAAAAAAu;
    bs &= ~ov;
    const uint32_t app = (bs << 1) | ov;
    const uint32_t oddseq = bs & odd & ~app;
    const uint32_t seq = oddseq + bs;
    const uint32_t esc = (even ^ (seq << 1)) & app;
    return q & ~esc;
}

// Inclusive prefix xor (prefix_xor, cujson/utils.cu:141-149).
__device__ __forceinline__ uint32_t prefix_xor32(uint32_t x) {
    x ^= x << 1;
    x ^= x << 2;
    x ^= x << 4;
    x ^= x << 8;
    x ^= x << 16;
    return x;
}


// `inside`: the caller knows the wave's whole range lies inside the unit
// (a uniform condition), so no bound checks are needed.
__device__ __forceinline__ void load_lane(const uint8_t *base, uint64_t o, uint64_t n,
                                          bool inside, uint32_t (&x)[8]) {
    if (inside) {
        const uint4 v0 = *reinterpret_cast<const uint4 *>(base + o);
        const uint4 v1 = *reinterpret_cast<const uint4 *>(base + o + 16);
        x[0] = v0.x; x[1] = v0.y; x[2] = v0.z; x[3] = v0.w;
        x[4] = v1.x; x[5] = v1.y; x[6] = v1.z; x[7] = v1.w;
        return;
    }
    if (o >= n) {
#pragma unroll
        for (int k = 0; k < 8; k++) x[k] = 0;
        return;
    }
    const uint4 v0 = *reinterpret_cast<const uint4 *>(base + o);
    uint4 v1 = make_uint4(0, 0, 0, 0);
    if (o + 16 < n) v1 = *reinterpret_cast<const uint4 *>(base + o + 16);
    x[0] = v0.x; x[1] = v0.y; x[2] = v0.z; x[3] = v0.w;
    x[4] = v1.x; x[5] = v1.y; x[6] = v1.z; x[7] = v1.w;
    const uint64_t rem = n - o;
    if (rem < 32) {
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int64_t keep = (int64_t)rem - 4 * k;
            if (keep <= 0) x[k] = 0;
            else if (keep < 4) x[k] &= (1u << (8 * keep)) - 1u;
        }
    }
}

// Bits 0..31 of the 64-bit value hi:lo shifted right by s: (hi << (32 - s)) |
// (lo >> s), one v_alignbit.  With s = 32 - k it shifts hi left by k and
// brings in the top k bits of lo (the bytes before this lane's 32).
__host__ __device__ __forceinline__ uint32_t alignbit(uint32_t hi, uint32_t lo, uint32_t s) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_alignbit(hi, lo, s);
#else
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> s);
#endif
}

// Four dwords at a dword-aligned address (one 16-byte global access).
typedef uint32_t U32x4A4 __attribute__((ext_vector_type(4), aligned(4)));

// Character classes of a lane's 32 bytes from its eight bit planes (R[k] bit
// i = bit k of byte i), as products of plane bits (each a few 3-input bit
// operations for all 32 bytes at once):
//   '"' 0x22 = 0010 0010   ':' 0x3A = 0011 1010   ',' 0x2C = 0010 1100
//   '\\' 0x5C = 0101 1100  '\n' 0x0A = 0000 1010
//   '[' 0x5B, ']' 0x5D, '{' 0x7B, '}' 0x7D = 01c1 1abd with abd = 011 or 101:
//   byte bit 1 tells an opener, byte bit 5 a curly bracket.
// op: the structural characters (':' ',' and brackets; '\n' too for JSON
// Lines); b1 / b5: bit 1 / bit 5 of the bracket bytes;
// for JSON Lines b5 also holds the '\n' bytes (never brackets).
struct LaneClasses {
    uint32_t bs, q, br, op, b1, b5;
};

template <bool JSONL>
__host__ __device__ __forceinline__ LaneClasses classes_from_planes(const uint32_t (&R)[8]) {
    const uint32_t lo010 = ~R[2] & R[1] & ~R[0];  // low bits 010: '"' ':' '\n'
    const uint32_t lo100 = R[2] & ~R[1] & ~R[0];  // low bits 100: ',' '\\'
    const uint32_t h001 = ~R[7] & ~R[6] & R[5];   // high bits 001: '"' ':' ','
    const uint32_t h01x11 = ~R[7] & R[6] & R[4] & R[3];  // brackets and '\\'
    LaneClasses c;
    c.q = h001 & lo010 & ~R[4] & ~R[3];
    const uint32_t colon = h001 & lo010 & R[4] & R[3];
    const uint32_t comma = h001 & lo100 & ~R[4] & R[3];
    c.bs = h01x11 & lo100 & ~R[5];
    c.br = h01x11 & R[0] & (R[1] ^ R[2]);
    c.op = colon | comma | c.br;
    c.b1 = R[1] & c.br;
    c.b5 = R[5] & c.br;
    if (JSONL) {
        const uint32_t nl = ~R[7] & ~R[6] & ~R[5] & ~R[4] & R[3] & lo010;
        c.op |= nl;
        c.b5 |= nl;
    }
    return c;
}

// UTF-8 byte classes as masks: continuation (80-BF), lead of two or more
// (>= C0), of three or more (>= E0), of four or more (>= F0).
struct Utf8Masks {
    uint32_t c, l, e, f;
};
__host__ __device__ __forceinline__ Utf8Masks utf8_masks(const uint32_t (&R)[8]) {
    Utf8Masks u;
    u.c = R[7] & ~R[6];
    u.l = R[7] & R[6];
    u.e = u.l & R[5];
    u.f = u.e & R[4];
    return u;
}

// The masks of the four bytes before a wave (a little-endian word, its last
// byte at bit 31), as the previous lane's masks would hold them.
__host__ __device__ __forceinline__ Utf8Masks utf8_masks_before(uint32_t prev_dw) {
    Utf8Masks u{0u, 0u, 0u, 0u};
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint32_t b = (prev_dw >> (8 * j)) & 0xFFu, bit = 1u << (28 + j);
        if ((b & 0xC0u) == 0x80u) u.c |= bit;
        if (b >= 0xC0u) u.l |= bit;
        if (b >= 0xE0u) u.e |= bit;
        if (b >= 0xF0u) u.f |= bit;
    }
    return u;
}

// Reference UTF-8 rule (checkUTF8, parse_standard_json.cu:210-262, reduced to
// what survives :122-123) on the masks of a lane's 32 bytes; p holds the
// masks of the 32 bytes before them (only their top three bits are read).
// The reference checks 4-byte words (bytes 4w..4w+3 of the lane here):
//  * a word with a high byte: byte i errs iff
//      [i-1 and i are continuations] != [i-2 >= E0 or i-3 >= F0]
//  * a word without one errs iff the previous word ends in a byte >= C0,
//    or its third byte is >= E0, or its second >= F0 (:219-223, :239-242),
//    i.e. the same three tests at the word's first byte.
// Results are gathered at bit 4w+3 of each word; `valid` selects the words
// that lie in the unit (bit 4w+3 set iff byte 4w is inside it).
__host__ __device__ __forceinline__ uint32_t utf8_plane_error(uint32_t high, const Utf8Masks &u,
                                                              const Utf8Masks &p, uint32_t valid) {
    const uint32_t c1 = alignbit(u.c, p.c, 31), l1 = alignbit(u.l, p.l, 31);
    const uint32_t e2 = alignbit(u.e, p.e, 30), f3 = alignbit(u.f, p.f, 29);
    const uint32_t x = e2 | f3;
    const uint32_t eb = (u.c & c1) ^ x;  // per byte, for words with a high byte
    const uint32_t ea = (l1 | x) << 3;   // at a word's first byte, moved to its bit 3
    uint32_t wh = high | (high << 1);
    wh |= wh << 2;  // bit 4w+3: word w has a high byte
    uint32_t we = eb | (eb << 1);
    we |= we << 2;  // bit 4w+3: some byte of word w errs
    return ((wh & we) | (~wh & ea)) & valid;
}

// ---------------------------------------------------------------------------
// Tile status for the decoupled look-back: two 64-bit words per tile.
//   aggregate  A = 1<<62 | map<<58 | n0<<29 | m0     B = 1<<62 | opc<<29 | brc
//   inclusive  A = 2<<62 | state<<60 | n_incl        B = 2<<62 | m_incl
// n0/m0 are the tile's structural/bracket counts when it is entered with
// e = 0, where e = s ^ (c & d) for entry state (c, s) and the tile's d bit;
// with e = 1 every wave's string mask flips, so the counts are opc - n0 and
// brc - m0 (opc/brc count the structural/bracket characters regardless of
// strings).  Words are written B then A; a reader accepts a tile only when
// both words carry the same flag.
// ---------------------------------------------------------------------------
constexpr unsigned long long kMask29 = (1ull << 29) - 1, kMask31 = (1ull << 31) - 1;

__device__ __forceinline__ void publish_agg(unsigned long long *st, uint32_t tile, uint32_t map,
                                            uint32_t n0, uint32_t m0, uint32_t opc,
                                            uint32_t brc) {
    agent_store64(&st[2 * (uint64_t)tile + 1],
                  (1ull << 62) | ((unsigned long long)opc << 29) | brc);
    agent_store64(&st[2 * (uint64_t)tile],
                  (1ull << 62) | ((unsigned long long)map << 58) |
                      ((unsigned long long)n0 << 29) | m0);
}

__device__ __forceinline__ void publish_inc(unsigned long long *st, uint32_t tile, uint32_t state,
                                            uint32_t n_incl, uint32_t m_incl) {
    agent_store64(&st[2 * (uint64_t)tile + 1], (2ull << 62) | m_incl);
    agent_store64(&st[2 * (uint64_t)tile],
                  (2ull << 62) | ((unsigned long long)state << 60) | n_incl);
}

// Waiting in the look-back (s_sleep units of 64 cycles).  Polls are
// agent-scope loads that miss the XCD's L2 and compete with the streaming
// traffic, so the walk first sleeps once (by then the tiles ahead have mostly
// published), then polls a window of kLookLanes tiles, and sleeps between
// polls of a window that is not ready.  Measured per 2 GB, with 5 waves per
// SIMD: windows of 8 / 16 / 32 / 48 / 64 tiles 0.900 / 0.822 / 0.791 / 0.791
// / 0.811 ms (sleep 64); sleep 8 / 16 / 32 / 64 / 127 between polls (64-tile
// windows) 0.883 (4 waves per SIMD) / 0.844 / 0.827 / 0.811 / 0.798 ms; a
// first sleep of 127 before the first poll 0.794 -> 0.779 ms.
#ifndef CJX_TOK_STAGGER
#define CJX_TOK_STAGGER 2
#endif

#ifndef CJX_LB_PRESLEEP
#define CJX_LB_PRESLEEP 127
#endif
#ifndef CJX_LB_SLEEP
#define CJX_LB_SLEEP 64
#endif
#ifndef CJX_LB_LANES
#define CJX_LB_LANES 32
#endif
constexpr int kLookLanes = CJX_LB_LANES;

__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
    return wave_read(wave_incl_scan(x), 63);
}

// Counts of a span of tiles as a function of the state (c, s) it is entered
// in.  Entered at (c, s), tile j of the span is entered at an affine image of
// (c, s), and its e bit is s ^ (a linear term in c) ^ (a constant): s flips
// every tile's e, which turns each tile's n into opc - n.  So the span's count
// entered at (c, s) is t[c] when s = 0 and sum - t[c] when s = 1.
struct SpanCounts {
    uint32_t n0, n1, nsum, m0, m1, msum;
};
// (Selections by masks: written as conditionals, the compiler turned them
// into indexed loads from a private copy of the span in scratch memory.)
__device__ __forceinline__ uint32_t sel_mask(uint32_t a, uint32_t b, uint32_t bit) {
    return a ^ ((a ^ b) & (0u - bit));  // bit ? b : a
}
__device__ __forceinline__ uint32_t span_n(const SpanCounts &f, uint32_t st) {
    const uint32_t t = sel_mask(f.n0, f.n1, st & 1u);
    return sel_mask(t, f.nsum - t, (st >> 1) & 1u);
}
__device__ __forceinline__ uint32_t span_m(const SpanCounts &f, uint32_t st) {
    const uint32_t t = sel_mask(f.m0, f.m1, st & 1u);
    return sel_mask(t, f.msum - t, (st >> 1) & 1u);
}

// Decoupled look-back over windows of kLookLanes tiles (lane i reads tile
// base - i, lane 0 the newest).  A window whose tiles have all published an aggregate
// is folded into the span seen so far (its map, and its counts as a function
// of its entry state) and the walk moves on to the 64 older tiles; it stops
// at the newest inclusive tile, whose state and counts start the chain.  It
// waits only for a tile newer than that one that has published nothing yet;
// tiles are taken in launch order and every running tile publishes its
// aggregate before it looks back, so the wait always ends.  Tiles before the
// unit read as inclusive with state 0 and counts 0.
__device__ __forceinline__ void lookback_tile(const unsigned long long *st, uint32_t tile,
                                              uint32_t unit_tile0, uint32_t &entry,
                                              uint32_t &n_excl, uint32_t &m_excl) {
    const uint32_t lane = lane_id();
    // The span between the window being read and this tile: its map (the
    // older part first) and its counts.
    uint32_t fmap = kMapIdentity;
    SpanCounts f{0u, 0u, 0u, 0u, 0u, 0u};
    int64_t base = (int64_t)tile - 1;
    __builtin_amdgcn_s_sleep(CJX_LB_PRESLEEP);
    for (;;) {
        const int64_t j = base - (int64_t)lane;
        unsigned long long a = 2ull << 62, b = 2ull << 62;
        const bool in_win = lane < (uint32_t)kLookLanes;
        if (in_win && j >= (int64_t)unit_tile0) {
            a = agent_load64(&st[2 * j]);
            b = agent_load64(&st[2 * j + 1]);
        }
        const uint32_t fa = (uint32_t)(a >> 62), fb = (uint32_t)(b >> 62);
        const uint32_t kind = (fa == fb) ? fa : kStInvalid;
        const uint64_t inc_lanes = __ballot(kind == kStInc && in_win);
        const uint32_t li = inc_lanes ? (uint32_t)__ffsll((long long)inc_lanes) - 1 : (uint32_t)kLookLanes;
        const uint64_t newer = (li >= 64) ? ~0ull : ((1ull << li) - 1ull);
        if (__ballot(kind != kStAgg) & newer) {  // a newer tile has published nothing yet
            __builtin_amdgcn_s_sleep(CJX_LB_SLEEP);
            continue;
        }
        // The aggregates newer than the inclusive tile (or the whole window):
        // maps composed from the oldest, P = L(63) then ... then L(lane).
        const bool part = lane < li;
        const uint32_t map = (uint32_t)((a >> 58) & 0xFu);
        uint32_t P = part ? map : kMapIdentity;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t other = (uint32_t)__shfl_down((int)P, off, 64);
            if (lane + off < 64) P = map_compose(other, P);
        }
        uint32_t E = (uint32_t)__shfl_down((int)P, 1, 64);
        if (lane == 63) E = kMapIdentity;
        // The tile's counts with the part entered at (0, 0) and at (1, 0).
        uint32_t cn0 = 0, cn1 = 0, cm0 = 0, cm1 = 0, opc = 0, brc = 0;
        if (part) {
            const uint32_t n0 = (uint32_t)((a >> 29) & kMask29), m0 = (uint32_t)(a & kMask29);
            opc = (uint32_t)((b >> 29) & kMask29);
            brc = (uint32_t)(b & kMask29);
            const uint32_t s0 = map_apply(E, 0u), s1 = map_apply(E, 1u);
            const uint32_t e0 = ((s0 >> 1) ^ (s0 & (map >> 2))) & 1u;
            const uint32_t e1 = ((s1 >> 1) ^ (s1 & (map >> 2))) & 1u;
            cn0 = e0 ? opc - n0 : n0;
            cn1 = e1 ? opc - n0 : n0;
            cm0 = e0 ? brc - m0 : m0;
            cm1 = e1 ? brc - m0 : m0;
        }
        const uint32_t gmap = wave_read(P, 0);
        SpanCounts g{wave_sum(cn0), wave_sum(cn1), wave_sum(opc),
                     wave_sum(cm0), wave_sum(cm1), wave_sum(brc)};
        // The part (older) then the span seen so far.
        const uint32_t x0 = map_apply(gmap, 0u), x1 = map_apply(gmap, 1u);
        SpanCounts h{g.n0 + span_n(f, x0), g.n1 + span_n(f, x1), g.nsum + f.nsum,
                     g.m0 + span_m(f, x0), g.m1 + span_m(f, x1), g.msum + f.msum};
        f = h;
        fmap = map_compose(gmap, fmap);
        if (li < (uint32_t)kLookLanes) {
            const uint32_t sk = (uint32_t)(wave_read((uint32_t)(a >> 32), (int)li) >> 28) & 3u;
            const uint32_t ni = wave_read((uint32_t)(a & kMask31), (int)li);
            const uint32_t mi = wave_read((uint32_t)(b & kMask31), (int)li);
            entry = map_apply(fmap, sk);
            n_excl = ni + span_n(f, sk);
            m_excl = mi + span_m(f, sk);
            return;
        }
        base -= kLookLanes;
    }
}

// State of a wave after phases 1 and 2 (everything before the look-back).
template <int ROUNDS>
struct WaveScan {
    uint32_t instr_r[ROUNDS], op_r[ROUNDS], br_r[ROUNDS];
    uint32_t c_run, s_run, seen, dflag, utf8_bad, lane_cnt0, lane_all;
};

#ifndef CJX_PREFETCH
#define CJX_PREFETCH 2
#endif
// Phases 1 and 2 over the rounds of one wave.  INSIDE: the wave's whole
// range lies inside the unit (a uniform condition, true for all but the last
// waves of a unit), so the rounds need no end checks and the loads no bounds
// checks; a separate instantiation keeps that path free of the branches
// (and the register copies at their joins) that the edge waves need.
template <bool JSONL, bool INSIDE, int ROUNDS>
__device__ __forceinline__ void scan_rounds(const uint8_t *ubase, uint64_t wave_off, uint64_t n,
                                            uint32_t lane, uint64_t lt, uint32_t weak_utf8,
                                            uint2 (&b15)[ROUNDS][64], WaveScan<ROUNDS> &ws) {
    // Rounds stream through PF buffers: round i+PF is loaded as soon as
    // round i has been transposed.
    constexpr int PF = CJX_PREFETCH;  // rounds loaded ahead
    uint32_t x[PF][8];
#pragma unroll
    for (int r = 0; r < PF; r++)
        if (r < ROUNDS)
            load_lane(ubase, wave_off + (uint64_t)r * kRoundBytes + (uint64_t)lane * kLaneBytes, n,
                      INSIDE, x[r]);
    uint32_t prev_dw = 0;
    if (wave_off > 0 && wave_off < n)
        prev_dw = *reinterpret_cast<const uint32_t *>(ubase + wave_off - 4);

    // UTF-8 masks of the bytes before the lane: the previous lane's, for lane
    // 0 the previous round's lane 63; entering the wave, the four bytes
    // before it at bits 28-31.
    Utf8Masks cu = utf8_masks_before(prev_dw);
    bool hi_prev = ((cu.c | cu.l) >> 29) != 0u;

    uint32_t(&instr_r)[ROUNDS] = ws.instr_r;
    uint32_t(&op_r)[ROUNDS] = ws.op_r;
    uint32_t(&br_r)[ROUNDS] = ws.br_r;
    uint32_t &c_run = ws.c_run, &s_run = ws.s_run, &seen = ws.seen, &dflag = ws.dflag;
    uint32_t &utf8_bad = ws.utf8_bad;
    uint32_t &lane_cnt0 = ws.lane_cnt0, &lane_all = ws.lane_all;
    c_run = s_run = seen = dflag = utf8_bad = lane_cnt0 = lane_all = 0;
    // per-lane counts over the rounds (each popcount accumulates in one op)
    uint32_t cnt_op0 = 0, cnt_br0 = 0, cnt_op = 0, cnt_br = 0;

#pragma unroll
    for (int i = 0; i < ROUNDS; i++) {
        // Rounds past the end of the unit hold no bytes: no masks, carries
        // unchanged (nothing after the unit's end depends on them).
        if (!INSIDE && wave_off + (uint64_t)i * kRoundBytes >= n) {
            instr_r[i] = op_r[i] = br_r[i] = 0;
            continue;
        }
        const uint64_t lane_off = wave_off + (uint64_t)i * kRoundBytes + (uint64_t)lane * kLaneBytes;
        // Byte transposes (word d, byte b <- lane byte 8b + d), then one 8x8
        // bit transpose per byte slot: R[k] is the plane of byte bit k.
        uint32_t R[8];
        uint32_t(&xi)[8] = x[i % PF];
        transpose4(xi[0], xi[2], xi[4], xi[6], R[0], R[1], R[2], R[3]);
        transpose4(xi[1], xi[3], xi[5], xi[7], R[4], R[5], R[6], R[7]);
        if (i + PF < ROUNDS)
            load_lane(ubase, lane_off + PF * kRoundBytes, n, INSIDE, xi);
        transpose8x8(R);

        // ---- phase 1: UTF-8 (reference rule) --------------------------------
        // Skipped for a round without high bytes whose carry has none either.
        // cu: the previous round's masks (lane 63 of them reaches lane 0).
        // They are formed only in rounds that run the check: a skipped round
        // has no high bytes (its masks are zero) and follows a round without
        // any, whose masks were zero too, so cu is already right for it.
        if (weak_utf8) {
            const bool hi = __ballot(R[7] != 0u) != 0ull;
            if (hi || hi_prev) {
                const Utf8Masks u = utf8_masks(R);
                const Utf8Masks pm{wave_shr1_from(u.c, cu.c), wave_shr1_from(u.l, cu.l),
                                   wave_shr1_from(u.e, cu.e), wave_shr1_from(u.f, cu.f)};
                uint32_t valid = 0x88888888u;
                if (!INSIDE && lane_off + 32 > n) {
                    valid = 0;
#pragma unroll
                    for (int w4 = 0; w4 < 8; w4++)
                        if (lane_off + 4 * w4 < n) valid |= 8u << (4 * w4);
                }
                if (utf8_plane_error(R[7], u, pm, valid)) utf8_bad = 1;
                cu = u;
            }
            hi_prev = hi;
        }

        // ---- phase 2: masks, escapes, strings ------------------------------
        const LaneClasses cls = classes_from_planes<JSONL>(R);
        const uint32_t bs = cls.bs;
        const uint32_t q = cls.q;
        br_r[i] = cls.br;
        op_r[i] = cls.op;
        b15[i][lane] = make_uint2(cls.b1, cls.b5);

        const uint32_t all_bs = (bs == 0xFFFFFFFFu);
        // Trailing backslash run of the lane.  Its parity is only read for
        // lanes that are not all backslashes (~bs != 0), so the count leading
        // zeros need not define the zero case.
        const uint32_t run = (uint32_t)__builtin_clz(~bs);
        const uint64_t NB = __ballot(!all_bs);
        const uint64_t TP = __ballot(run & 1u);
        // Parity of the backslash run that ends just before the lane: the
        // trailing run of the nearest lower lane that is not all backslashes.
        // When no lane of the round is all backslashes (a uniform test) that
        // is the previous lane's, one lane shift.
        uint32_t ov;
        if (NB == ~0ull) {
            ov = wave_shr1(run & 1u, c_run);
        } else {
            const uint64_t mm = NB & lt;
            ov = mm ? (uint32_t)((TP >> (63 - __clzll(mm))) & 1ull) : c_run;
        }
        const uint32_t rq = unescaped32(bs, q, ov);
        const uint32_t px = prefix_xor32(rq);
        const uint64_t QP = __ballot(__popc(rq) & 1u);
        // String state entering the lane: s_run plus the odd-quote lanes
        // below it (the count's bit 0, sign-extended to a mask).
        const uint32_t s_cnt = __builtin_amdgcn_mbcnt_hi((uint32_t)(QP >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)QP, s_run));
        instr_r[i] = px ^ (uint32_t)__builtin_amdgcn_sbfe((int32_t)s_cnt, 0u, 1u);
        cnt_op0 += (uint32_t)__popc(op_r[i] & ~instr_r[i]);
        cnt_br0 += (uint32_t)__popc(br_r[i] & ~instr_r[i]);
        cnt_op += (uint32_t)__popc(op_r[i]);
        cnt_br += (uint32_t)__popc(br_r[i]);

        if (!seen && NB) {  // first non-backslash byte of the wave: is it a quote?
            const uint32_t k0 = (uint32_t)__ffsll((long long)NB) - 1;
            const uint32_t nbpos = (uint32_t)__ffs(~bs) - 1;
            const uint32_t fq = all_bs ? 0u : ((q >> nbpos) & 1u);
            const uint64_t FQ = __ballot(fq);
            dflag = (uint32_t)((FQ >> k0) & 1ull);
            seen = 1;
        }
        if (NB) c_run = (uint32_t)((TP >> (63 - __clzll(NB))) & 1ull);
        s_run ^= (uint32_t)(__popcll(QP) & 1);
        // The round's masks are formed here, not sunk past the loop: that
        // keeps each round's intermediates from living across later rounds.
        __asm__ volatile("" : "+v"(instr_r[i]), "+v"(op_r[i]), "+v"(br_r[i]));
    }
    lane_cnt0 = cnt_op0 | (cnt_br0 << 16);
    lane_all = cnt_op | (cnt_br << 16);
}

// Large units (SMALL = false): one workgroup per 64 KiB tile; tiles are taken
// from a ticket counter in launch order so that every look-back waits only on
// running workgroups.  Small units (SMALL = true, at most one wave's 16 KiB):
// one wave per unit, which starts in state (0, 0) and owns the unit's whole
// output region, so there is no look-back and no workgroup barrier.
// A tile is WAVES waves of ROUNDS rounds (WAVES x ROUNDS x 2 KiB = 64 KiB);
// a small-unit workgroup is WAVES waves of SROUNDS rounds, one unit per wave:
// kTokRounds rounds cover the largest small unit, one round the units of at
// most 2 KiB (fewer registers, so more of those waves are resident).  WPE is
// the register budget in waves per SIMD (see launch_tiles).
template <bool JSONL, bool SMALL, int WAVES, int WPE, int SROUNDS = kTokRounds>
__global__ __attribute__((amdgpu_waves_per_eu(WPE, WPE))) __launch_bounds__(64 * WAVES) void k_tokenize(
    const uint8_t *__restrict__ in, const UnitDesc *__restrict__ units,
    const uint32_t *__restrict__ tile_unit, const uint32_t *__restrict__ small_units,
    uint32_t n_small, unsigned long long *__restrict__ tile_st, LaunchState *__restrict__ ls,
    UnitOut *__restrict__ uo, uint32_t *__restrict__ out_struct, uint8_t *__restrict__ oc_char,
    uint32_t *__restrict__ oc_idx, uint8_t *__restrict__ row_nl, uint32_t weak_utf8,
    uint32_t stagger_blocks) {
    constexpr int ROUNDS = SMALL ? SROUNDS : kTileBytes / (WAVES * kRoundBytes);
    constexpr uint64_t WAVE_BYTES = (uint64_t)ROUNDS * kRoundBytes;
    static_assert(SMALL || WAVES * ROUNDS * kRoundBytes == kTileBytes, "a tile is 64 KiB");
    static_assert(!SMALL || WAVE_BYTES <= kWaveBytes, "one wave covers a small unit");
    __shared__ uint32_t s_tile, s_unit;
    __shared__ uint32_t s_map[WAVES];
    __shared__ uint32_t s_cnt0[WAVES], s_cnt_all[WAVES];
    __shared__ uint32_t s_fix[WAVES], s_nbase[WAVES], s_mbase[WAVES];
    __shared__ uint32_t s_stage[WAVES][kStageRows];  // one round's positions per wave
    __shared__ uint8_t s_bchar[WAVES][kStageRows];   // and its newline flags (JSON Lines)
    // Bracket characters as two bit masks per lane-round, kept in LDS until
    // compaction: byte bit 1 ('[' / '{' have it, ']' / '}' do not) and byte
    // bit 5 (curly), both restricted to bracket bytes.
    // For JSON Lines the bit-5 mask also holds the '\n' bytes, which are never
    // brackets; they give the per-row newline flags of record tables (row_nl).
    __shared__ uint2 s_b15[WAVES][ROUNDS][64];  // (bit-1 mask, bit-5 mask)
    const bool want_nl = JSONL && row_nl != nullptr;

    const uint32_t tid = threadIdx.x;
    // Wave-uniform values live in scalar registers (readfirstlane), so the
    // addresses and bounds derived from them cost no vector instructions.
    const uint32_t w = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const uint64_t lt = lanemask_lt();

    uint32_t tile = 0, u;
    if constexpr (SMALL) {
        const uint32_t idx = blockIdx.x * WAVES + w;
        if (idx >= n_small) return;  // the whole wave; SMALL code has no workgroup barrier
        u = small_units[idx];
    } else {
        // The first blocks of a long launch start in four groups about 7 us
        // apart (s_sleep 127 = 8128 cycles, twice per step) before they draw
        // a tile: blocks that start together stay in step, their reads, look-
        // backs and writes in bursts.  (Tiles go in ticket order, so a block
        // that sleeps only draws a later tile.)
        if (blockIdx.x < stagger_blocks)
            for (uint32_t k = 0; k < CJX_TOK_STAGGER * (blockIdx.x & 3u); k++) __builtin_amdgcn_s_sleep(127);
        if (tid == 0) {
            const uint32_t t = atomicAdd(&ls->tile_counter, 1u);
            s_tile = t;
            s_unit = tile_unit[t];
        }
        __syncthreads();
        tile = __builtin_amdgcn_readfirstlane(s_tile);
        u = __builtin_amdgcn_readfirstlane(s_unit);
    }
    const UnitDesc U = units[u];
    const uint32_t tile_in_unit = SMALL ? 0u : tile - U.tile0;
    const uint8_t *ubase = in + U.in_off;
    const uint64_t wave_off =
        SMALL ? 0ull : (uint64_t)tile_in_unit * kTileBytes + (uint64_t)w * WAVE_BYTES;

    const bool inside = wave_off + (uint64_t)WAVE_BYTES <= U.n;
    WaveScan<ROUNDS> ws;
    if (inside)
        scan_rounds<JSONL, true, ROUNDS>(ubase, wave_off, U.n, lane, lt, weak_utf8, s_b15[w], ws);
    else
        scan_rounds<JSONL, false, ROUNDS>(ubase, wave_off, U.n, lane, lt, weak_utf8, s_b15[w], ws);
    const uint32_t(&instr_r)[ROUNDS] = ws.instr_r;
    const uint32_t(&op_r)[ROUNDS] = ws.op_r;
    const uint32_t(&br_r)[ROUNDS] = ws.br_r;
    const uint32_t c_run = ws.c_run, s_run = ws.s_run, seen = ws.seen, dflag = ws.dflag;
    const uint32_t utf8_bad = ws.utf8_bad, lane_cnt0 = ws.lane_cnt0, lane_all = ws.lane_all;

    const uint64_t bad_lanes = __ballot(utf8_bad);
    const uint32_t wcnt0 = wave_sum(lane_cnt0), wall = wave_sum(lane_all);
    uint32_t fix, nb, mb, wave_rows;
    if constexpr (SMALL) {
        // The unit is this wave: entered in state (0, 0), so no fix.
        const uint32_t nu = wcnt0 & 0xFFFFu, mu = wcnt0 >> 16;
        if (lane == 0) {
            if (bad_lanes) atomicOr(&uo[u].utf8_err, 1u);
            uo[u].n_base = U.out_off;
            uo[u].m_base = U.out_off;
            uo[u].n_end = U.out_off + nu;
            uo[u].m_end = U.out_off + mu;
            if (nu > U.out_cap) atomicOr(&ls->overflow, 1u);
        }
        fix = 0;
        nb = mb = U.out_off;
        wave_rows = nu;
        (void)wall;
        (void)dflag;
        (void)seen;
    } else {
        if (lane == 0) {
            if (bad_lanes) atomicOr(&uo[u].utf8_err, 1u);
            s_map[w] = (seen ? 0u : 1u) | (c_run << 1) | (dflag << 2) | (s_run << 3);
            s_cnt0[w] = wcnt0;
            s_cnt_all[w] = wall;
        }
        __syncthreads();

        // ---- one look-back: entry state and output offsets -----------------------
        if (w == 0) {
            // The tile entered at (c, s) = (0, 0): wave string fixes and counts.
            uint32_t T = kMapIdentity, state = 0, n0 = 0, m0 = 0, opc = 0, brc = 0;
            uint32_t fix0[WAVES];
#pragma unroll
            for (int kk = 0; kk < WAVES; kk++) {
                const uint32_t m = s_map[kk];
                fix0[kk] = ((state >> 1) & 1u) ^ ((state & 1u) & ((m >> 2) & 1u));
                state = map_apply(m, state);
                T = map_compose(T, m);
                const uint32_t c0 = s_cnt0[kk], ca = s_cnt_all[kk];
                n0 += fix0[kk] ? (ca & 0xFFFFu) - (c0 & 0xFFFFu) : (c0 & 0xFFFFu);
                m0 += fix0[kk] ? (ca >> 16) - (c0 >> 16) : (c0 >> 16);
                opc += ca & 0xFFFFu;
                brc += ca >> 16;
            }
            uint32_t entry = 0, n_excl = 0, m_excl = 0;
            if (tile_in_unit != 0) {
                if (lane == 0) publish_agg(tile_st, tile, T, n0, m0, opc, brc);
                lookback_tile(tile_st, tile, U.tile0, entry, n_excl, m_excl);
            }
            const uint32_t e = ((entry >> 1) ^ (entry & (T >> 2))) & 1u;
            const uint32_t nt = e ? opc - n0 : n0, mt = e ? brc - m0 : m0;
            if (lane == 0) {
                publish_inc(tile_st, tile, map_apply(T, entry), n_excl + nt, m_excl + mt);
                uint32_t nb = U.out_off + n_excl, mb = U.out_off + m_excl;
#pragma unroll
                for (int kk = 0; kk < WAVES; kk++) {
                    const uint32_t f = fix0[kk] ^ e;
                    const uint32_t c0 = s_cnt0[kk], ca = s_cnt_all[kk];
                    s_fix[kk] = f;
                    s_nbase[kk] = nb;
                    s_mbase[kk] = mb;
                    nb += f ? (ca & 0xFFFFu) - (c0 & 0xFFFFu) : (c0 & 0xFFFFu);
                    mb += f ? (ca >> 16) - (c0 >> 16) : (c0 >> 16);
                }
                if (tile_in_unit == 0) {
                    uo[u].n_base = U.out_off;
                    uo[u].m_base = U.out_off;
                }
                if (tile_in_unit == U.ntiles - 1) {
                    const uint32_t nu = n_excl + nt, mu = m_excl + mt;
                    uo[u].n_end = U.out_off + nu;
                    uo[u].m_end = U.out_off + mu;
                    if (nu > U.out_cap) atomicOr(&ls->overflow, 1u);
                }
            }
        }
        __syncthreads();
        const uint32_t fw = s_fix[w], c0 = s_cnt0[w] & 0xFFFFu, ca = s_cnt_all[w] & 0xFFFFu;
        fix = fw ? ~0u : 0u;
        nb = s_nbase[w];
        mb = s_mbase[w];
        wave_rows = fw ? ca - c0 : c0;
    }
    // Uniform values in scalar registers.
    nb = __builtin_amdgcn_readfirstlane(nb);
    mb = __builtin_amdgcn_readfirstlane(mb);
    fix = __builtin_amdgcn_readfirstlane(fix);
    wave_rows = __builtin_amdgcn_readfirstlane(wave_rows);

    // ---- select structurals and compact them ---------------------------------
    const uint32_t cap_end = U.out_off + U.out_cap;
    // Every row of the wave lies in the unit's region (bracket rows too: they
    // are fewer and start no later), so its rounds need no capacity checks.
    const bool fits = (uint64_t)nb + wave_rows <= cap_end;
    uint32_t overflow = 0;
    // The rounds that hold bytes (the others hold no rows), the first row
    // position of the wave (1-based) and the stage of the wave: uniform,
    // computed once in scalar registers.  (With 16 rounds the loop below is
    // not unrolled when it ends at a round count, so it keeps the per-round
    // test of the round's offset.)
    const uint64_t wave_rem = U.n > wave_off ? U.n - wave_off : 0ull;
    const uint32_t nrounds = __builtin_amdgcn_readfirstlane(
        (uint32_t)min((uint64_t)ROUNDS, (wave_rem + kRoundBytes - 1) / kRoundBytes));
    const uint32_t pos_w = __builtin_amdgcn_readfirstlane((uint32_t)(U.pos_base + wave_off + 1));
    const uint32_t lane_pos = lane * kLaneBytes;
    // (The stage's LDS address as one scalar, so an entry's address is one
    // operation from its index.)
    typedef __attribute__((address_space(3))) uint32_t LdsU32;
    const uint32_t stg_addr =
        __builtin_amdgcn_readfirstlane((uint32_t)(uintptr_t)(LdsU32 *)&s_stage[w][0]);
    LdsU32 *const stg = (LdsU32 *)(uintptr_t)stg_addr;
    // Bracket characters by (open, curly) code: ']' 0x5D, '[' 0x5B, '}' 0x7D, '{' 0x7B.
    const uint32_t br_chars = __builtin_amdgcn_readfirstlane(0x7B7D5B5Du);
#pragma unroll
    for (int i = 0; i < ROUNDS; i++) {
        if constexpr (ROUNDS <= 8) {
            if ((uint32_t)i >= nrounds) break;  // no bytes, no rows
        } else {
            if (wave_off + (uint64_t)i * kRoundBytes >= U.n) break;
        }
        const uint32_t sti = op_r[i] & ~(instr_r[i] ^ fix);
        const uint32_t oci = br_r[i] & sti;
        const uint32_t cnt = (uint32_t)__popc(sti) | ((uint32_t)__popc(oci) << 16);
        const uint32_t incl = wave_incl_scan(cnt);
        const uint32_t excl = incl - cnt;
        const uint32_t rtot = __builtin_amdgcn_readfirstlane(wave_read(incl, 63));
        const uint32_t idx0 = nb + (excl & 0xFFFFu);
        uint32_t midx = mb + (excl >> 16);
        const uint32_t pos0 = (pos_w + (uint32_t)i * kRoundBytes) + lane_pos;
        uint32_t m = sti;
        // Structural positions.  The wave's positions of the round are
        // gathered in LDS and stored as one contiguous run; a round with more
        // than kStageRows of them, or one that reaches the end of the region,
        // is stored lane by lane.
        const uint32_t rn = rtot & 0xFFFFu;
        const uint32_t nlm = want_nl ? (s_b15[w][i][lane].y & ~br_r[i]) : 0u;
        if (fits && rn <= (uint32_t)kStageRows) {
            // The lane's structurals go to slots [k, kh] of the stage, taken
            // from both ends of its mask at once (lowest bit to slot k, highest
            // to slot kh), so a wave runs half as many steps as its densest
            // lane has structurals.  A last single bit is both ends: the same
            // value to the same slot twice.
            uint32_t k = excl & 0xFFFFu, kh = k + (cnt & 0xFFFFu) - 1u;
            const uint32_t posh = (pos_w + (uint32_t)i * kRoundBytes + 31u) + lane_pos;
            if (want_nl) {  // and the row's newline flag (in s_bchar, free until the brackets)
                while (m) {
                    const uint32_t jl = (uint32_t)__builtin_ctz(m), jz = (uint32_t)__builtin_clz(m);
                    stg[k] = pos0 + jl;
                    stg[kh] = posh - jz;
                    s_bchar[w][k] = (uint8_t)((nlm >> jl) & 1u);
                    s_bchar[w][kh] = (uint8_t)((nlm >> (31u - jz)) & 1u);
                    m = (m & (m - 1u)) & ~(0x80000000u >> jz);
                    k++;
                    kh--;
                }
            } else {
                while (m) {
                    const uint32_t jl = (uint32_t)__builtin_ctz(m), jz = (uint32_t)__builtin_clz(m);
                    stg[k] = pos0 + jl;
                    stg[kh] = posh - jz;
                    m = (m & (m - 1u)) & ~(0x80000000u >> jz);
                    k++;
                    kh--;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (rn >= 4u) {
                // Four rows per lane and one 16-byte store each (global
                // accesses of a dword or more need only dword alignment).  The
                // last lane's group is moved back to end at row rn; the rows it
                // shares with the lane before are written twice, with the same
                // values.
                for (uint32_t t = 4u * lane; t < rn; t += 256u) {
                    const uint32_t s = min(t, rn - 4u);
                    U32x4A4 v;
                    v.x = stg[s];
                    v.y = stg[s + 1u];
                    v.z = stg[s + 2u];
                    v.w = stg[s + 3u];
                    *reinterpret_cast<U32x4A4 *>(out_struct + nb + s) = v;
                }
            } else {
                for (uint32_t t = lane; t < rn; t += 64) out_struct[nb + t] = stg[t];
            }
            if (want_nl)
                for (uint32_t t = lane; t < rn; t += 64) row_nl[nb + t] = s_bchar[w][t];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        } else if ((uint64_t)idx0 + (uint32_t)__popc(sti) <= cap_end) {
            uint32_t idx = idx0;
            while (m) {
                const uint32_t j = (uint32_t)__builtin_ctz(m);
                m &= m - 1;
                out_struct[idx] = pos0 + j;
                if (want_nl) row_nl[idx] = (uint8_t)((nlm >> j) & 1u);
                idx++;
            }
        } else if (m) {
            overflow = 1;
            uint32_t idx = idx0;
            while (m) {
                const uint32_t j = (uint32_t)__builtin_ctz(m);
                m &= m - 1;
                if (idx < cap_end) {
                    out_struct[idx] = pos0 + j;
                    if (want_nl) row_nl[idx] = (uint8_t)((nlm >> j) & 1u);
                }
                idx++;
            }
        }
        // Brackets: the structural row and the character of each, gathered in
        // LDS and stored as contiguous runs when the round's rows fit.
        // Characters: (open, curly) -> ']' 0x5D, '[' 0x5B, '}' 0x7D, '{' 0x7B.
        m = oci;
        const uint2 b15 = s_b15[w][i][lane];
        const uint32_t b1 = b15.x, b5 = b15.y;
        const uint32_t rm = rtot >> 16;
        if (fits && rn <= (uint32_t)kStageRows) {
            // Both ends of the bracket mask at once, as for the structurals.
            // A stage entry holds the bracket's structural row within the
            // round (below 2^11) and, in bits 16-26, its byte offset within
            // the round (lane * 32 + bit); the character is formed when the
            // run is stored, from the (open, curly) masks of the lane that
            // holds the byte.
            uint32_t k = excl >> 16, kh = k + (cnt >> 16) - 1u;
            const uint32_t r0 = (excl & 0xFFFFu) + (lane << 21);
            while (m) {
                const uint32_t jl = (uint32_t)__builtin_ctz(m), jz = (uint32_t)__builtin_clz(m);
                const uint32_t jh = 31u - jz;
                const uint32_t el = ((uint32_t)__popc(sti & ((1u << jl) - 1u)) + r0) + (jl << 16);
                const uint32_t eh = ((uint32_t)__popc(sti & ((1u << jh) - 1u)) + r0) + (jh << 16);
                stg[k] = el;
                stg[kh] = eh;
                m = (m & (m - 1u)) & ~(0x80000000u >> jz);
                k++;
                kh--;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (uint32_t t = lane; t < rm; t += 64) {
                const uint32_t e = stg[t];
                const uint2 bm = s_b15[w][i][e >> 21];
                const uint32_t j = (e >> 16) & 31u;
                const uint32_t sel = __builtin_amdgcn_ubfe(bm.x, j, 1u) | (__builtin_amdgcn_ubfe(bm.y, j, 1u) << 1);
                oc_idx[mb + t] = nb + (e & 0xFFFFu);
                oc_char[mb + t] = (uint8_t)perm(0u, br_chars, sel);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        } else {
            while (m) {
                const uint32_t j = (uint32_t)__builtin_ctz(m);
                m &= m - 1;
                const uint32_t t = idx0 + (uint32_t)__popc(sti & ((1u << j) - 1u));
                if (midx < cap_end) {
                    oc_idx[midx] = t;
                    const uint32_t sel = ((b1 >> j) & 1u) | (((b5 >> j) & 1u) << 1);
                    oc_char[midx] = (uint8_t)perm(0u, 0x7B7D5B5Du, sel);
                } else {
                    overflow = 1;
                }
                midx++;
            }
        }
        nb += rtot & 0xFFFFu;
        mb += rtot >> 16;
    }
    if (__ballot(overflow) && lane == 0) atomicOr(&ls->overflow, 1u);
}

namespace {

// Workgroups of a kernel (of `threads` threads) that the device holds at
// once, or 0 if the runtime cannot tell.
uint32_t resident_blocks(const void *kernel, int threads) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess ||
        cus <= 0 || per_cu <= 0)
        return 0u;
    return (uint32_t)(cus * per_cu);
}

template <bool JSONL, bool SMALL, int WAVES, int WPE, int SROUNDS = kTokRounds>
void launch_variant(hipStream_t s, uint32_t grid, const LaunchBufs &b, const uint32_t *small_units,
                    const uint8_t *d_in, uint32_t n_small, uint32_t *out_struct, uint32_t weak) {
    // Tile launches of at least four turns of the device stagger their
    // first turn of blocks.
    static const uint32_t resident = SMALL ? 0u : resident_blocks(
        reinterpret_cast<const void *>(k_tokenize<JSONL, SMALL, WAVES, WPE, SROUNDS>), 64 * WAVES);
    const uint32_t stagger = !SMALL && grid >= 4u * resident ? resident : 0u;
    hipLaunchKernelGGL((k_tokenize<JSONL, SMALL, WAVES, WPE, SROUNDS>), dim3(grid),
                       dim3(64 * WAVES), 0, s, d_in, b.units, b.tile_unit, small_units, n_small,
                       b.tile_st, b.ls, b.uout, out_struct, b.oc_char, b.oc_idx, b.row_nl, weak,
                       stagger);
}

// Register budgets: 6 waves per SIMD for 4 and 8 waves of 16 and 8 KiB (79
// VGPRs, JSON Lines 80, no spills; per 2 GB 0.852 ms at 4 waves, 0.811 at 5,
// 0.764 at 6 once the wave-uniform values moved to scalar registers), 4 for
// 16 waves of 4 KiB; the 32 KiB waves of the 2-wave tile carry twice the
// per-round counts and take a 3-wave budget.
#ifndef CJX_WPE_T2
#define CJX_WPE_T2 3
#endif
#ifndef CJX_WPE_T4
#define CJX_WPE_T4 6
#endif
#ifndef CJX_WPE_T8
#define CJX_WPE_T8 6
#endif
#ifndef CJX_WPE_T16
#define CJX_WPE_T16 4
#endif
template <bool JSONL>
void launch_tiles(hipStream_t s, uint32_t grid, uint32_t tile_waves, const LaunchBufs &b,
                  const uint8_t *d_in, uint32_t *out_struct, uint32_t weak) {
    switch (tile_waves) {
    case 2: launch_variant<JSONL, false, 2, CJX_WPE_T2>(s, grid, b, nullptr, d_in, 0u, out_struct, weak); break;
    case 8: launch_variant<JSONL, false, 8, CJX_WPE_T8>(s, grid, b, nullptr, d_in, 0u, out_struct, weak); break;
    case 16: launch_variant<JSONL, false, 16, CJX_WPE_T16>(s, grid, b, nullptr, d_in, 0u, out_struct, weak); break;
    default: launch_variant<JSONL, false, 4, CJX_WPE_T4>(s, grid, b, nullptr, d_in, 0u, out_struct, weak); break;
    }
}

template <bool JSONL, int SROUNDS>
void launch_small(hipStream_t s, uint32_t n_small, uint32_t small_waves, const LaunchBufs &b,
                  const uint32_t *list, const uint8_t *d_in, uint32_t *out_struct, uint32_t weak) {
    // one round needs 47 VGPRs: a budget of 8 waves per SIMD
    constexpr int WPE = SROUNDS == 1 ? 8 : 4;
    const uint32_t sw = small_waves;
    const uint32_t grid = (n_small + sw - 1) / sw;
    switch (sw) {
    case 1:
        launch_variant<JSONL, true, 1, WPE, SROUNDS>(s, grid, b, list, d_in, n_small, out_struct,
                                                weak);
        break;
    case 2:
        launch_variant<JSONL, true, 2, WPE, SROUNDS>(s, grid, b, list, d_in, n_small, out_struct,
                                                weak);
        break;
    case 8:
        launch_variant<JSONL, true, 8, WPE, SROUNDS>(s, grid, b, list, d_in, n_small, out_struct,
                                                weak);
        break;
    default:
        launch_variant<JSONL, true, 4, WPE, SROUNDS>(s, grid, b, list, d_in, n_small, out_struct,
                                                weak);
        break;
    }
}

// The small units: the first n_tiny of the list (units of at most one round)
// with the one-round kernel, the rest with the full small-unit kernel.
template <bool JSONL>
void launch_smalls(hipStream_t s, uint32_t n_small, uint32_t n_tiny, uint32_t small_waves,
                   const LaunchBufs &b, const uint8_t *d_in, uint32_t *out_struct, uint32_t weak) {
    if (n_tiny > n_small) n_tiny = n_small;
    if (n_tiny)
        launch_small<JSONL, 1>(s, n_tiny, small_waves, b, b.small_units, d_in, out_struct, weak);
    if (n_small > n_tiny)
        launch_small<JSONL, kTokRounds>(s, n_small - n_tiny, small_waves, b, b.small_units + n_tiny,
                                        d_in, out_struct, weak);
}

}  // namespace

hipError_t launch_tokenize(hipStream_t s, const LaunchBufs &b, const uint8_t *d_in,
                           uint32_t ntiles, uint32_t n_small, bool jsonl, bool weak_utf8,
                           uint32_t *out_struct, TokShape shape, uint32_t n_tiny) {
    const uint32_t weak = weak_utf8 ? 1u : 0u;
    if (jsonl) {
        if (ntiles) launch_tiles<true>(s, ntiles, shape.tile_waves, b, d_in, out_struct, weak);
        if (n_small)
            launch_smalls<true>(s, n_small, n_tiny, shape.small_waves, b, d_in, out_struct, weak);
    } else {
        if (ntiles) launch_tiles<false>(s, ntiles, shape.tile_waves, b, d_in, out_struct, weak);
        if (n_small)
            launch_smalls<false>(s, n_small, n_tiny, shape.small_waves, b, d_in, out_struct, weak);
    }
    return hipGetLastError();
}

hipError_t launch_tokenize_tiles(hipStream_t s, const LaunchBufs &b, const uint8_t *d_in,
                                 uint32_t count, bool jsonl, bool weak_utf8, uint32_t *out_struct,
                                 TokShape shape) {
    if (count == 0) return hipSuccess;
    const uint32_t weak = weak_utf8 ? 1u : 0u;
    if (jsonl) launch_tiles<true>(s, count, shape.tile_waves, b, d_in, out_struct, weak);
    else launch_tiles<false>(s, count, shape.tile_waves, b, d_in, out_struct, weak);
    return hipGetLastError();
}

}  // namespace cjx

// Host-side model of the byte classifier (same tables, v_perm emulated), for
// the table self-test that runs without a GPU.
static uint32_t host_perm(uint32_t hi, uint32_t lo, uint32_t sel) {
    const uint64_t data = ((uint64_t)hi << 32) | lo;
    uint32_t out = 0;
    for (int i = 0; i < 4; i++) {
        const uint32_t s = (sel >> (8 * i)) & 0xFFu;
        uint32_t v;
        if (s >= 13) v = 0xFF;
        else if (s == 12) v = 0;
        else if (s >= 8) v = ((data >> (15 + 16 * (s - 8))) & 1) ? 0xFF : 0;
        else v = (uint32_t)(data >> (8 * s)) & 0xFFu;
        out |= v << (8 * i);
    }
    return out;
}

// Host run of the mask transpose, for the self-test that runs without a GPU.
extern "C" void cjx_debug_transpose8x8(uint32_t *t) {
    uint32_t v[8];
    for (int d = 0; d < 8; d++) v[d] = t[d];
    cjx::transpose8x8(v);
    for (int d = 0; d < 8; d++) t[d] = v[d];
}

extern "C" uint32_t cjx_debug_byte_class(uint32_t byte, int jsonl) {
    using namespace cjx;
    const uint64_t t35 = jsonl ? kLutBits35Lines : kLutBits35Std;
    const uint32_t r = byte & 0xFFu;
    const uint32_t a = host_perm((uint32_t)(kLutBits02 >> 32), (uint32_t)kLutBits02, r & 7u);
    const uint32_t b = host_perm((uint32_t)(t35 >> 32), (uint32_t)t35, (r >> 3) & 7u);
    const uint32_t c = host_perm(0u, (uint32_t)kLutBits67, (r >> 6) & 3u);
    return (a & b & c) & 0xFFu;
}

// Host run of the tokenizer's lane step: the byte transposes, the bit planes
// and the character classes of 32 bytes (out: bs, q, br, op, b1, b5).
extern "C" void cjx_debug_lane_classes(const uint8_t *b32, int jsonl, uint32_t *out) {
    using namespace cjx;
    uint32_t x[8];
    for (int k = 0; k < 8; k++)
        x[k] = (uint32_t)b32[4 * k] | ((uint32_t)b32[4 * k + 1] << 8) |
               ((uint32_t)b32[4 * k + 2] << 16) | ((uint32_t)b32[4 * k + 3] << 24);
    uint32_t R[8];
    transpose4(x[0], x[2], x[4], x[6], R[0], R[1], R[2], R[3]);
    transpose4(x[1], x[3], x[5], x[7], R[4], R[5], R[6], R[7]);
    transpose8x8(R);
    const LaneClasses c = jsonl ? classes_from_planes<true>(R) : classes_from_planes<false>(R);
    out[0] = c.bs; out[1] = c.q; out[2] = c.br; out[3] = c.op; out[4] = c.b1; out[5] = c.b5;
}

// Host run of the tokenizer's reference-rule UTF-8 check over a unit: lanes
// of 32 bytes in order, each carrying its masks to the next, and every
// `wave_bytes` bytes a new wave whose carry comes from the four bytes before
// it, as on the device.  Returns 1 if the unit fails the check.
extern "C" int cjx_debug_weak_utf8(const uint8_t *d, size_t n, size_t wave_bytes) {
    using namespace cjx;
    int bad = 0;
    Utf8Masks cu{0u, 0u, 0u, 0u};
    for (size_t o = 0; o < n; o += 32) {
        if (wave_bytes && o % wave_bytes == 0) {
            uint32_t prev_dw = 0;
            if (o > 0)
                prev_dw = (uint32_t)d[o - 4] | ((uint32_t)d[o - 3] << 8) |
                          ((uint32_t)d[o - 2] << 16) | ((uint32_t)d[o - 1] << 24);
            cu = utf8_masks_before(prev_dw);
        }
        uint8_t b[32] = {0};
        for (size_t k = 0; k < 32 && o + k < n; k++) b[k] = d[o + k];
        uint32_t x[8];
        for (int k = 0; k < 8; k++)
            x[k] = (uint32_t)b[4 * k] | ((uint32_t)b[4 * k + 1] << 8) |
                   ((uint32_t)b[4 * k + 2] << 16) | ((uint32_t)b[4 * k + 3] << 24);
        uint32_t R[8];
        transpose4(x[0], x[2], x[4], x[6], R[0], R[1], R[2], R[3]);
        transpose4(x[1], x[3], x[5], x[7], R[4], R[5], R[6], R[7]);
        transpose8x8(R);
        const Utf8Masks u = utf8_masks(R);
        uint32_t valid = 0;
        for (int w4 = 0; w4 < 8; w4++)
            if (o + 4 * (size_t)w4 < n) valid |= 8u << (4 * w4);
        if (utf8_plane_error(R[7], u, cu, valid)) bad = 1;
        cu = u;
    }
    return bad;
}
