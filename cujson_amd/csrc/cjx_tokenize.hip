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

// Reference UTF-8 rule (checkUTF8, parse_standard_json.cu:210-262, reduced to
// what survives :122-123) over the lane's eight words, on transposed data:
// word 2s is slot s of transposed words 0-3, word 2s+1 slot s of words 4-7.
//  * a word with a high byte: byte p errs iff
//      [p-1 and p are continuations] != [p-2 >= E0 or p-3 >= F0]
//  * a word without one errs iff the previous word ends in a byte >= C0,
//    or its third byte is >= E0, or its second >= F0 (:219-223, :239-242)
// The class bits are CONT = bit 1, E = bit 5, F = bit 6; the per-byte error
// terms are formed in bit 1 and the ASCII-word terms in bit 7.
// pk5/pk6/pk7/pr7 carry the previous lane's transposed words 5, 6, 7 (classes)
// and 7 (raw).  vA/vB: 0xFF in slot s if word 2s / 2s+1 lies inside the unit.
__device__ __forceinline__ uint32_t utf8_lane_error(const uint32_t (&r)[8], const uint32_t (&k)[8],
                                                    uint32_t pk5, uint32_t pk6, uint32_t pk7,
                                                    uint32_t pr7, uint32_t vA, uint32_t vB) {
    const uint32_t ks5 = alignbyte3(k[5], pk5), ks6 = alignbyte3(k[6], pk6);
    const uint32_t ks7 = alignbyte3(k[7], pk7), rs7 = alignbyte3(r[7], pr7);
    const uint32_t e0 = (ks7 & k[0]) ^ ((ks6 >> 4) | (ks5 >> 5));
    const uint32_t e1 = (k[0] & k[1]) ^ ((ks7 >> 4) | (ks6 >> 5));
    const uint32_t e2 = (k[1] & k[2]) ^ ((k[0] >> 4) | (ks7 >> 5));
    const uint32_t e3 = (k[2] & k[3]) ^ ((k[1] >> 4) | (k[0] >> 5));
    const uint32_t e4 = (k[3] & k[4]) ^ ((k[2] >> 4) | (k[1] >> 5));
    const uint32_t e5 = (k[4] & k[5]) ^ ((k[3] >> 4) | (k[2] >> 5));
    const uint32_t e6 = (k[5] & k[6]) ^ ((k[4] >> 4) | (k[3] >> 5));
    const uint32_t e7 = (k[6] & k[7]) ^ ((k[5] >> 4) | (k[4] >> 5));
    const uint32_t errA = (e0 | e1 | e2 | e3) & 0x02020202u;
    const uint32_t errB = (e4 | e5 | e6 | e7) & 0x02020202u;
    const uint32_t gA = (r[0] | r[1] | r[2] | r[3]) & 0x80808080u;
    const uint32_t gB = (r[4] | r[5] | r[6] | r[7]) & 0x80808080u;
    const uint32_t gen = ((errA & (gA >> 6)) & vA) | ((errB & (gB >> 6)) & vB);
    const uint32_t ascA = ((rs7 & ~(ks7 << 6)) | (ks6 << 2) | (ks5 << 1)) & ~gA & vA;
    const uint32_t ascB = ((r[3] & ~(k[3] << 6)) | (k[2] << 2) | (k[1] << 1)) & ~gB & vB;
    return gen | ((ascA | ascB) & 0x80808080u);
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

// One wave walks back from `tile` over a window of 64 x kLookPer tiles: lane
// i reads tiles tile - 1 - (kLookPer * i + k), k = 0..kLookPer-1 (position
// q = kLookPer * i + k, 0 = the newest).  When an inclusive tile lies in the
// window and every newer tile is an aggregate, the maps of those tiles are
// composed from the oldest to give each one's entry state, hence its e bit
// and its counts; their sum plus the inclusive tile's counts is this tile's
// exclusive prefix.  Tiles before the unit read as inclusive with state 0 and
// counts 0.  The walk waits while no inclusive tile is in the window (the
// oldest running tile always finds one, so waiting cannot deadlock).  The
// inclusive frontier advances up to a window per memory round trip, which is
// what bounds a launch of many tiles.
#ifndef CJX_LOOK_PER
#define CJX_LOOK_PER 1
#endif
// Sleep between polls (s_sleep units of 64 cycles).  Pollers read the status
// lines the frontier tiles are publishing into; fewer polls let the
// inclusive frontier advance faster: 2 -> 48 measured 0.950 -> 0.922 ms per
// 2 GB (16: 0.933, 32: 0.925, 64: 0.928, 127: 0.936; an exponential backoff
// and a first probe of the A words only were slower).
#ifndef CJX_LB_SLEEP
#define CJX_LB_SLEEP 48
#endif
constexpr int kLookPer = CJX_LOOK_PER;
__device__ __forceinline__ void lookback_tile(const unsigned long long *st, uint32_t tile,
                                              uint32_t unit_tile0, uint32_t &entry,
                                              uint32_t &n_excl, uint32_t &m_excl) {
    const uint32_t lane = lane_id();
    for (;;) {
        unsigned long long a[kLookPer], b[kLookPer];
#pragma unroll
        for (int k = 0; k < kLookPer; k++) {
            const int64_t j = (int64_t)tile - 1 - (int64_t)(kLookPer * lane + k);
            a[k] = b[k] = 2ull << 62;
            if (j >= (int64_t)unit_tile0) {
                a[k] = agent_load64(&st[2 * j]);
                b[k] = agent_load64(&st[2 * j + 1]);
            }
        }
        // Newest inclusive entry of the lane, and whether every newer entry
        // of the lane is an aggregate.
        uint32_t k_inc = kLookPer, inc_state = 0;
        bool ok = true;
#pragma unroll
        for (int k = kLookPer - 1; k >= 0; k--) {
            const uint32_t fa = (uint32_t)(a[k] >> 62), fb = (uint32_t)(b[k] >> 62);
            const uint32_t kind = (fa == fb) ? fa : kStInvalid;
            if (kind == kStInc) {
                k_inc = (uint32_t)k;
                inc_state = (uint32_t)(a[k] >> 60) & 3u;
            }
        }
#pragma unroll
        for (int k = 0; k < kLookPer; k++) {
            const uint32_t fa = (uint32_t)(a[k] >> 62), fb = (uint32_t)(b[k] >> 62);
            const uint32_t kind = (fa == fb) ? fa : kStInvalid;
            if ((uint32_t)k < k_inc && kind != kStAgg) ok = false;
        }
        const uint64_t inc_lanes = __ballot(k_inc < (uint32_t)kLookPer);
        if (!inc_lanes) {
            __builtin_amdgcn_s_sleep(CJX_LB_SLEEP);
            continue;
        }
        const uint32_t li = (uint32_t)__ffsll((long long)inc_lanes) - 1;
        const uint64_t upto = (li == 63) ? ~0ull : ((2ull << li) - 1ull);
        if (__ballot(!ok) & upto) {  // a tile newer than the inclusive one is not ready
            __builtin_amdgcn_s_sleep(CJX_LB_SLEEP);
            continue;
        }
        const uint32_t qi = kLookPer * li + wave_read(k_inc, (int)li);
        const uint32_t sk = wave_read(inc_state, (int)li);
        // Lane map: its newer-than-inclusive entries, oldest first.
        uint32_t L = kMapIdentity;
#pragma unroll
        for (int k = kLookPer - 1; k >= 0; k--) {
            const uint32_t q = kLookPer * lane + (uint32_t)k;
            if (q < qi) L = map_compose(L, (uint32_t)((a[k] >> 58) & 0xFu));
        }
        // P = L(lane 63) then ... then L(lane): inclusive scan from the oldest.
        uint32_t P = L;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t other = (uint32_t)__shfl_down((int)P, off, 64);
            if (lane + off < 64) P = map_compose(other, P);
        }
        uint32_t E = (uint32_t)__shfl_down((int)P, 1, 64);
        if (lane == 63) E = kMapIdentity;
        uint32_t stj = map_apply(E, sk);
        uint32_t cn = 0, cm = 0;
#pragma unroll
        for (int k = kLookPer - 1; k >= 0; k--) {
            const uint32_t q = kLookPer * lane + (uint32_t)k;
            if (q < qi) {
                const uint32_t map = (uint32_t)((a[k] >> 58) & 0xFu);
                const uint32_t e = ((stj >> 1) ^ (stj & (map >> 2))) & 1u;
                const uint32_t n0 = (uint32_t)((a[k] >> 29) & kMask29), m0 = (uint32_t)(a[k] & kMask29);
                const uint32_t opc = (uint32_t)((b[k] >> 29) & kMask29);
                const uint32_t brc = (uint32_t)(b[k] & kMask29);
                cn += e ? opc - n0 : n0;
                cm += e ? brc - m0 : m0;
                stj = map_apply(map, stj);
            } else if (q == qi) {
                cn += (uint32_t)(a[k] & kMask31);
                cm += (uint32_t)(b[k] & kMask31);
            }
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            cn += (uint32_t)__shfl_xor((int)cn, off, 64);
            cm += (uint32_t)__shfl_xor((int)cm, off, 64);
        }
        n_excl = cn;
        m_excl = cm;
        entry = map_apply(wave_read(P, 0), sk);
        return;
    }
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
    return wave_read(wave_incl_scan(x), 63);
}

// State of a wave after phases 1 and 2 (everything before the look-back).
template <int ROUNDS>
struct WaveScan {
    uint32_t instr_r[ROUNDS], op_r[ROUNDS], br_r[ROUNDS];
    uint32_t c_run, s_run, seen, dflag, utf8_bad, lane_cnt0, lane_all;
};

// Phases 1 and 2 over the rounds of one wave.  INSIDE: the wave's whole
// range lies inside the unit (a uniform condition, true for all but the last
// waves of a unit), so the rounds need no end checks and the loads no bounds
// checks; a separate instantiation keeps that path free of the branches
// (and the register copies at their joins) that the edge waves need.
template <bool JSONL, bool INSIDE, int ROUNDS>
__device__ __forceinline__ void scan_rounds(const uint8_t *ubase, uint64_t wave_off, uint64_t n,
                                            uint32_t lane, uint64_t lt, uint32_t weak_utf8,
                                            uint32_t (&b1)[ROUNDS][64],
                                            uint32_t (&b5)[ROUNDS][64], WaveScan<ROUNDS> &ws) {
    // Rounds stream through a double buffer: round i+2 is loaded as soon as
    // round i has been transposed.
    uint32_t x[2][8];
    load_lane(ubase, wave_off + (uint64_t)lane * kLaneBytes, n, INSIDE, x[0]);
    if (ROUNDS > 1)
        load_lane(ubase, wave_off + kRoundBytes + (uint64_t)lane * kLaneBytes, n, INSIDE, x[1]);
    uint32_t prev_dw = 0;
    if (wave_off > 0 && wave_off < n)
        prev_dw = *reinterpret_cast<const uint32_t *>(ubase + wave_off - 4);

    // Carries entering the wave: the four bytes before it (UTF-8 context).
    uint32_t c_k5, c_k6, c_k7, c_r7;
    {
        const uint32_t cp = classify<JSONL>(prev_dw);
        c_k7 = cp;
        c_k6 = cp << 8;
        c_k5 = cp << 16;
        c_r7 = prev_dw;
    }

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
        uint32_t r[8], k[8];
        transpose4(x[i & 1][0], x[i & 1][2], x[i & 1][4], x[i & 1][6], r[0], r[1], r[2], r[3]);
        transpose4(x[i & 1][1], x[i & 1][3], x[i & 1][5], x[i & 1][7], r[4], r[5], r[6], r[7]);
        if (i + 2 < ROUNDS)
            load_lane(ubase, lane_off + 2 * kRoundBytes, n, INSIDE, x[i & 1]);
#pragma unroll
        for (int d = 0; d < 8; d++) k[d] = classify<JSONL>(r[d]);

        // ---- phase 1: UTF-8 ---------------------------------------------------
        {
            const uint32_t pk5 = wave_shr1(k[5], c_k5), pk6 = wave_shr1(k[6], c_k6);
            const uint32_t pk7 = wave_shr1(k[7], c_k7), pr7 = wave_shr1(r[7], c_r7);
            const uint32_t high = ((r[0] | r[1] | r[2] | r[3] | r[4] | r[5] | r[6] | r[7]) &
                                   0x80808080u) |
                                  ((pk5 | pk6 | pk7) & 0x62000000u) | (pr7 & 0x80000000u);
            if (weak_utf8 && __ballot(high != 0)) {
                uint32_t vA = ~0u, vB = ~0u;
                if (!INSIDE && lane_off + 32 > n) {
                    vA = vB = 0;
#pragma unroll
                    for (int s = 0; s < 4; s++) {
                        if (lane_off + 8 * s < n) vA |= 0xFFu << (8 * s);
                        if (lane_off + 8 * s + 4 < n) vB |= 0xFFu << (8 * s);
                    }
                }
                if (high && utf8_lane_error(r, k, pk5, pk6, pk7, pr7, vA, vB)) utf8_bad = 1;
            }
            c_k5 = wave_read(k[5], 63);
            c_k6 = wave_read(k[6], 63);
            c_k7 = wave_read(k[7], 63);
            c_r7 = wave_read(r[7], 63);
        }

        // ---- phase 2: masks, escapes, strings ------------------------------
        // Class bits 1 and 5 (CONT, E) give way to raw bits 1 and 5, then one
        // transpose yields every mask.
#pragma unroll
        for (int d = 0; d < 8; d++) k[d] = (k[d] & ~0x22222222u) | (r[d] & 0x22222222u);
        transpose8x8(k);
        const uint32_t bs = k[4];
        const uint32_t q = k[7];
        br_r[i] = k[2];
        op_r[i] = k[0] | k[2] | k[3];
        b1[i][lane] = k[1] & k[2];
        b5[i][lane] = (k[5] & k[2]) | (JSONL ? (k[0] & ~k[5]) : 0u);

        const uint32_t all_bs = (bs == 0xFFFFFFFFu);
        const uint32_t run = (uint32_t)__clz(~bs);  // trailing backslash run of the lane
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
        const uint32_t s_lane = s_run ^ (mbcnt64(QP) & 1u);
        instr_r[i] = px ^ (s_lane ? ~0u : 0u);
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
// most 2 KiB (fewer registers, so more of those waves are resident).  WPE is the register budget in waves per SIMD:
// 4 x 8 at 4 waves per SIMD uses 103 VGPRs without spills; with the interior
// and edge rounds as separate instantiations, the 5-wave budget (96 VGPRs)
// spills and measured 4 % slower.
template <bool JSONL, bool SMALL, int WAVES, int WPE, int SROUNDS = kTokRounds>
__global__ __attribute__((amdgpu_waves_per_eu(WPE, WPE))) __launch_bounds__(64 * WAVES) void k_tokenize(
    const uint8_t *__restrict__ in, const UnitDesc *__restrict__ units,
    const uint32_t *__restrict__ tile_unit, const uint32_t *__restrict__ small_units,
    uint32_t n_small, unsigned long long *__restrict__ tile_st, LaunchState *__restrict__ ls,
    UnitOut *__restrict__ uo, uint32_t *__restrict__ out_struct, uint8_t *__restrict__ oc_char,
    uint32_t *__restrict__ oc_idx, uint8_t *__restrict__ row_nl, uint32_t weak_utf8) {
    constexpr int ROUNDS = SMALL ? SROUNDS : kTileBytes / (WAVES * kRoundBytes);
    constexpr uint64_t WAVE_BYTES = (uint64_t)ROUNDS * kRoundBytes;
    static_assert(SMALL || WAVES * ROUNDS * kRoundBytes == kTileBytes, "a tile is 64 KiB");
    static_assert(!SMALL || WAVE_BYTES <= kWaveBytes, "one wave covers a small unit");
    __shared__ uint32_t s_tile, s_unit;
    __shared__ uint32_t s_map[WAVES];
    __shared__ uint32_t s_cnt0[WAVES], s_cnt_all[WAVES];
    __shared__ uint32_t s_fix[WAVES], s_nbase[WAVES], s_mbase[WAVES];
    __shared__ uint32_t s_stage[WAVES][kStageRows];  // one round's positions per wave
    __shared__ uint8_t s_bchar[WAVES][kStageRows];   // and its bracket characters
    // Bracket characters as two bit masks per lane-round, kept in LDS until
    // compaction: byte bit 1 ('[' / '{' have it, ']' / '}' do not) and byte
    // bit 5 (curly), both restricted to bracket bytes.
    // For JSON Lines s_b5 also holds the '\n' bytes, which are never brackets:
    // class 0 without byte bit 5 (':' has it, '\n' not); they give the
    // per-row newline flags of record tables (row_nl).
    __shared__ uint32_t s_b1[WAVES][ROUNDS][64], s_b5[WAVES][ROUNDS][64];
    const bool want_nl = JSONL && row_nl != nullptr;

    const uint32_t tid = threadIdx.x;
    const uint32_t w = tid >> 6, lane = tid & 63;
    const uint64_t lt = lanemask_lt();

    uint32_t tile = 0, u;
    if constexpr (SMALL) {
        const uint32_t idx = blockIdx.x * WAVES + w;
        if (idx >= n_small) return;  // the whole wave; SMALL code has no workgroup barrier
        u = small_units[idx];
    } else {
        if (tid == 0) {
            const uint32_t t = atomicAdd(&ls->tile_counter, 1u);
            s_tile = t;
            s_unit = tile_unit[t];
        }
        __syncthreads();
        tile = s_tile;
        u = s_unit;
    }
    const UnitDesc U = units[u];
    const uint32_t tile_in_unit = SMALL ? 0u : tile - U.tile0;
    const uint8_t *ubase = in + U.in_off;
    const uint64_t wave_off =
        SMALL ? 0ull : (uint64_t)tile_in_unit * kTileBytes + (uint64_t)w * WAVE_BYTES;

    const bool inside = wave_off + (uint64_t)WAVE_BYTES <= U.n;
    WaveScan<ROUNDS> ws;
    if (inside)
        scan_rounds<JSONL, true, ROUNDS>(ubase, wave_off, U.n, lane, lt, weak_utf8, s_b1[w], s_b5[w], ws);
    else
        scan_rounds<JSONL, false, ROUNDS>(ubase, wave_off, U.n, lane, lt, weak_utf8, s_b1[w], s_b5[w], ws);
    const uint32_t(&instr_r)[ROUNDS] = ws.instr_r;
    const uint32_t(&op_r)[ROUNDS] = ws.op_r;
    const uint32_t(&br_r)[ROUNDS] = ws.br_r;
    const uint32_t c_run = ws.c_run, s_run = ws.s_run, seen = ws.seen, dflag = ws.dflag;
    const uint32_t utf8_bad = ws.utf8_bad, lane_cnt0 = ws.lane_cnt0, lane_all = ws.lane_all;

    const uint64_t bad_lanes = __ballot(utf8_bad);
    const uint32_t wcnt0 = wave_sum(lane_cnt0), wall = wave_sum(lane_all);
    uint32_t fix, nb, mb;
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
        fix = s_fix[w] ? ~0u : 0u;
        nb = s_nbase[w];
        mb = s_mbase[w];
    }

    // ---- select structurals and compact them ---------------------------------
    const uint32_t cap_end = U.out_off + U.out_cap;
    uint32_t overflow = 0;
#pragma unroll
    for (int i = 0; i < ROUNDS; i++) {
        if (wave_off + (uint64_t)i * kRoundBytes >= U.n) break;  // no bytes, no rows
        const uint64_t lane_off = wave_off + (uint64_t)i * kRoundBytes + (uint64_t)lane * kLaneBytes;
        const uint32_t sti = op_r[i] & ~(instr_r[i] ^ fix);
        const uint32_t oci = br_r[i] & sti;
        const uint32_t cnt = (uint32_t)__popc(sti) | ((uint32_t)__popc(oci) << 16);
        const uint32_t incl = wave_incl_scan(cnt);
        const uint32_t excl = incl - cnt;
        const uint32_t rtot = wave_read(incl, 63);
        const uint32_t idx0 = nb + (excl & 0xFFFFu);
        uint32_t midx = mb + (excl >> 16);
        const uint32_t pos0 = (uint32_t)(U.pos_base + lane_off + 1);
        uint32_t m = sti;
        // Structural positions.  The wave's positions of the round are
        // gathered in LDS and stored as one contiguous run; a round with more
        // than kStageRows of them, or one that reaches the end of the region,
        // is stored lane by lane.
        const uint32_t rn = rtot & 0xFFFFu;
        const uint32_t nlm = want_nl ? (s_b5[w][i][lane] & ~br_r[i]) : 0u;
        if (rn <= (uint32_t)kStageRows && (uint64_t)nb + rn <= cap_end) {
            // The lane's structurals go to slots [k, kh] of the stage, taken
            // from both ends of its mask at once (lowest bit to slot k, highest
            // to slot kh), so a wave runs half as many steps as its densest
            // lane has structurals.  A last single bit is both ends: the same
            // value to the same slot twice.
            uint32_t k = excl & 0xFFFFu, kh = k + (cnt & 0xFFFFu) - 1u;
            const uint32_t posh = pos0 + 31u;
            if (want_nl) {  // and the row's newline flag (in s_bchar, free until the brackets)
                while (m) {
                    const uint32_t jl = (uint32_t)__builtin_ctz(m), jz = (uint32_t)__builtin_clz(m);
                    s_stage[w][k] = pos0 + jl;
                    s_stage[w][kh] = posh - jz;
                    s_bchar[w][k] = (uint8_t)((nlm >> jl) & 1u);
                    s_bchar[w][kh] = (uint8_t)((nlm >> (31u - jz)) & 1u);
                    m = (m & (m - 1u)) & ~(0x80000000u >> jz);
                    k++;
                    kh--;
                }
            } else {
                while (m) {
                    const uint32_t jl = (uint32_t)__builtin_ctz(m), jz = (uint32_t)__builtin_clz(m);
                    s_stage[w][k] = pos0 + jl;
                    s_stage[w][kh] = posh - jz;
                    m = (m & (m - 1u)) & ~(0x80000000u >> jz);
                    k++;
                    kh--;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (uint32_t t = lane; t < rn; t += 64) out_struct[nb + t] = s_stage[w][t];
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
        const uint32_t b1 = s_b1[w][i][lane], b5 = s_b5[w][i][lane];
        const uint32_t rm = rtot >> 16;
        if (rn <= (uint32_t)kStageRows && (uint64_t)mb + rm <= cap_end) {
            // Both ends of the bracket mask at once, as for the structurals.
            // A stage entry holds the bracket's structural row within the
            // round (below 2^11) and its (open, curly) code in bits 16-17; the
            // row index and the character are formed when the run is stored.
            uint32_t k = excl >> 16, kh = k + (cnt >> 16) - 1u;
            const uint32_t r0 = excl & 0xFFFFu;  // the lane's first row in the round
            while (m) {
                const uint32_t jl = (uint32_t)__builtin_ctz(m), jz = (uint32_t)__builtin_clz(m);
                const uint32_t jh = 31u - jz;
                const uint32_t el = ((uint32_t)__popc(sti & ((1u << jl) - 1u)) + r0) +
                                    (__builtin_amdgcn_ubfe(b1, jl, 1u) << 16) +
                                    (__builtin_amdgcn_ubfe(b5, jl, 1u) << 17);
                const uint32_t eh = ((uint32_t)__popc(sti & ((1u << jh) - 1u)) + r0) +
                                    (__builtin_amdgcn_ubfe(b1, jh, 1u) << 16) +
                                    (__builtin_amdgcn_ubfe(b5, jh, 1u) << 17);
                s_stage[w][k] = el;
                s_stage[w][kh] = eh;
                m = (m & (m - 1u)) & ~(0x80000000u >> jz);
                k++;
                kh--;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (uint32_t t = lane; t < rm; t += 64) {
                const uint32_t e = s_stage[w][t];
                oc_idx[mb + t] = nb + (e & 0xFFFFu);
                oc_char[mb + t] = (uint8_t)perm(0u, 0x7B7D5B5Du, e >> 16);
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

template <bool JSONL, bool SMALL, int WAVES, int WPE, int SROUNDS = kTokRounds>
void launch_variant(hipStream_t s, uint32_t grid, const LaunchBufs &b, const uint32_t *small_units,
                    const uint8_t *d_in, uint32_t n_small, uint32_t *out_struct, uint32_t weak) {
    hipLaunchKernelGGL((k_tokenize<JSONL, SMALL, WAVES, WPE, SROUNDS>), dim3(grid),
                       dim3(64 * WAVES), 0, s, d_in, b.units, b.tile_unit, small_units, n_small,
                       b.tile_st, b.ls, b.uout, out_struct, b.oc_char, b.oc_idx, b.row_nl, weak);
}

// Register budgets: 4 waves per SIMD for 4, 8 and 16 waves of 16, 8 and
// 4 KiB; the 32 KiB waves of the 2-wave tile carry twice the per-round
// counts and take a 3-wave budget.
template <bool JSONL>
void launch_tiles(hipStream_t s, uint32_t grid, uint32_t tile_waves, const LaunchBufs &b,
                  const uint8_t *d_in, uint32_t *out_struct, uint32_t weak) {
    switch (tile_waves) {
    case 2: launch_variant<JSONL, false, 2, 3>(s, grid, b, nullptr, d_in, 0u, out_struct, weak); break;
    case 8: launch_variant<JSONL, false, 8, 4>(s, grid, b, nullptr, d_in, 0u, out_struct, weak); break;
    case 16: launch_variant<JSONL, false, 16, 4>(s, grid, b, nullptr, d_in, 0u, out_struct, weak); break;
    default: launch_variant<JSONL, false, 4, 4>(s, grid, b, nullptr, d_in, 0u, out_struct, weak); break;
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
