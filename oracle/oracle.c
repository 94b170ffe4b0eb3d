/*
 * oracle.c - single-threaded CPU restatement of cuJSON's three-phase parse.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Word-level transcription of the
 * reference's live kernels, run sequentially:
 *   phase 1  UTF-8 validation      PSJ:9-327   (PJL:9-315)
 *   phase 2  tokenization          PSJ:446-571, 766-1021, 1023-1343 (PJL:422-937)
 *   phase 3  structure recognition PSJ:1345-1358, 1431-1582 (PJL:939-1068)
 * The Thrust scans and the stable sort the reference calls are restated by
 * their mathematical definitions (prefix sums modulo the element width, a
 * stable counting sort by key).
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* CUDA per-byte SIMD intrinsics (semantics from the CUDA math API docs).     */
/* ------------------------------------------------------------------------ */
static inline uint32_t byte_of(uint32_t w, int i) { return (w >> (8 * i)) & 0xFFu; }

uint32_t orc_vcmpeq4(uint32_t a, uint32_t b) {
    uint32_t r = 0;
    for (int i = 0; i < 4; i++)
        if (byte_of(a, i) == byte_of(b, i)) r |= 0xFFu << (8 * i);
    return r;
}
uint32_t orc_vcmpltu4(uint32_t a, uint32_t b) {
    uint32_t r = 0;
    for (int i = 0; i < 4; i++)
        if (byte_of(a, i) < byte_of(b, i)) r |= 0xFFu << (8 * i);
    return r;
}
uint32_t orc_vcmpgeu4(uint32_t a, uint32_t b) {
    uint32_t r = 0;
    for (int i = 0; i < 4; i++)
        if (byte_of(a, i) >= byte_of(b, i)) r |= 0xFFu << (8 * i);
    return r;
}
uint32_t orc_vcmpgtu4(uint32_t a, uint32_t b) {
    uint32_t r = 0;
    for (int i = 0; i < 4; i++)
        if (byte_of(a, i) > byte_of(b, i)) r |= 0xFFu << (8 * i);
    return r;
}
uint32_t orc_vsubus4(uint32_t a, uint32_t b) {
    uint32_t r = 0;
    for (int i = 0; i < 4; i++) {
        uint32_t x = byte_of(a, i), y = byte_of(b, i);
        r |= (x > y ? x - y : 0u) << (8 * i);
    }
    return r;
}
uint32_t orc_vsubss4(uint32_t a, uint32_t b) {
    uint32_t r = 0;
    for (int i = 0; i < 4; i++) {
        int x = (int8_t)byte_of(a, i), y = (int8_t)byte_of(b, i);
        int d = x - y;
        if (d > 127) d = 127;
        if (d < -128) d = -128;
        r |= ((uint32_t)(uint8_t)(int8_t)d) << (8 * i);
    }
    return r;
}

static inline int popc32(uint32_t x) { return __builtin_popcount(x); }
static inline int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }

/* Little-endian 32-bit word j of the unit, zero-padded past n (PSJ:1608-1619). */
static inline uint32_t word_at(const uint8_t *d, size_t n, size_t j) {
    uint32_t w = 0;
    for (int i = 0; i < 4; i++) {
        size_t p = 4 * j + (size_t)i;
        if (p < n) w |= (uint32_t)d[p] << (8 * i);
    }
    return w;
}

/* ------------------------------------------------------------------------ */
/* Phase 1: UTF-8                                                             */
/* ------------------------------------------------------------------------ */

/* vectorizedClassification, PSJ:9-148.  Only the TWO_CONTS test survives
 * the reassignment at PSJ:122-123; every other flag is computed and dropped. */
static uint32_t classify(uint32_t block, uint32_t prev1) {
    const uint32_t TOO_SHORT = 0x01010101u, TOO_LONG = 0x02020202u, OVERLONG_2 = 0x20202020u,
                   OVERLONG_3 = 0x04040404u, OVERLONG_4 = 0x40404040u, SURROGATE = 0x10101010u,
                   TWO_CONTS = 0x80808080u, TOO_LARGE = 0x08080808u, TOO_LARGE_1000 = 0x40404040u;
    uint32_t byte_1 = (orc_vcmpltu4(prev1, 0x80808080u) & TOO_LONG) |
                      (orc_vcmpgeu4(prev1, 0xC0C0C0C0u) & TOO_SHORT) |
                      ((orc_vcmpeq4(prev1, 0xC0C0C0C0u) | orc_vcmpeq4(prev1, 0xC1C1C1C1u)) &
                       OVERLONG_2) |
                      (orc_vcmpeq4(prev1, 0xEDEDEDEDu) & SURROGATE) |
                      (orc_vcmpeq4(prev1, 0xE0E0E0E0u) & OVERLONG_3) |
                      (orc_vcmpeq4(prev1, 0xF0F0F0F0u) & OVERLONG_4) |
                      (orc_vcmpgtu4(prev1, 0xF4F4F4F4u) & TOO_LARGE_1000) |
                      (orc_vcmpgtu4(prev1, 0xF3F3F3F3u) & TOO_LARGE);
    byte_1 = orc_vcmpeq4(byte_1, 0u) & TWO_CONTS; /* PSJ:122-123 */
    uint32_t hi = (block >> 4) & 0x0F0F0F0Fu;      /* PSJ:129 */
    uint32_t lt12 = orc_vcmpltu4(hi, 0x0C0C0C0Cu);
    uint32_t byte_2_high =
        ((orc_vcmpltu4(hi, 0x08080808u) | orc_vcmpgtu4(hi, 0x0B0B0B0Bu)) & TOO_SHORT) |
        (lt12 & orc_vcmpgeu4(hi, 0x08080808u) & (TOO_LONG | OVERLONG_2 | TWO_CONTS)) |
        (lt12 & orc_vcmpgtu4(hi, 0x08080808u) & TOO_LARGE) |
        (orc_vcmpeq4(hi, 0x08080808u) & (TOO_LARGE_1000 | OVERLONG_4)) |
        (orc_vcmpgtu4(hi, 0x09090909u) & lt12 & SURROGATE);
    return byte_1 & byte_2_high; /* PSJ:145 */
}

/* continuationBytes, PSJ:152-184. */
static uint32_t continuation(uint32_t prev2, uint32_t prev3, uint32_t sc) {
    const uint32_t third = 0xDFDFDFDFu, fourth = 0xEFEFEFEFu;
    uint32_t is_third = orc_vsubus4(prev2, third);
    uint32_t is_fourth = orc_vsubus4(prev3, fourth);
    uint32_t gt = orc_vsubss4(is_third | is_fourth, 0u);
    uint32_t must32 = orc_vcmpgtu4(gt, 0u);
    return (must32 & 0x80808080u) ^ sc;
}

/* checkUTF8 for one word j (PSJ:237-257).  The reference writes every word's
 * result with atomicExch, so the surviving value is whichever thread wrote
 * last (PSJ:242, 256); the defined semantics here is the OR over words. */
static uint32_t utf8_word_error(uint32_t current, uint32_t previous) {
    const uint32_t max_val = (0xBFu << 24) | (0xDFu << 16) | (0xEFu << 8) | 0xFFu; /* PSJ:219-223 */
    if ((current & 0x80808080u) == 0) return orc_vsubus4(previous, max_val);
    uint64_t dist = ((uint64_t)current << 32) | (uint64_t)previous;
    uint32_t prev1 = (uint32_t)(dist >> 24);
    uint32_t prev2 = (uint32_t)(dist >> 16);
    uint32_t prev3 = (uint32_t)(dist >> 8);
    uint32_t sc = classify(current, prev1);
    return continuation(prev2, prev3, sc);
}

int orc_utf8_valid_strict(const uint8_t *s, size_t n) {
    size_t i = 0;
    while (i < n) {
        uint8_t c = s[i];
        if (c < 0x80) { i++; continue; }
        int len;
        uint32_t cp;
        if (c >= 0xC2 && c <= 0xDF) { len = 2; cp = c & 0x1F; }
        else if (c >= 0xE0 && c <= 0xEF) { len = 3; cp = c & 0x0F; }
        else if (c >= 0xF0 && c <= 0xF4) { len = 4; cp = c & 0x07; }
        else return 0;
        if (i + (size_t)len > n) return 0;
        for (int k = 1; k < len; k++) {
            uint8_t t = s[i + k];
            if ((t & 0xC0) != 0x80) return 0;
            cp = (cp << 6) | (t & 0x3F);
        }
        if (len == 3 && (cp < 0x800 || (cp >= 0xD800 && cp <= 0xDFFF))) return 0;
        if (len == 4 && (cp < 0x10000 || cp > 0x10FFFF)) return 0;
        i += (size_t)len;
    }
    return 1;
}

/* UTF8Validation, PSJ:264-327: ASCII early-out (checkAscii PSJ:186-208), then
 * the per-word check over W4 = ceil(n/4) words, prev = 0 for word 0. */
int orc_utf8_check(const uint8_t *d, size_t n, unsigned flags) {
    if (flags & ORC_FLAG_UTF8_STRICT) return !orc_utf8_valid_strict(d, n);
    size_t W4 = (n + 3) / 4;
    int has_high = 0;
    for (size_t j = 0; j < W4; j++)
        if (word_at(d, n, j) & 0x80808080u) { has_high = 1; break; }
    if (!has_high) return 0; /* PSJ:290-295 */
    uint32_t err = 0, prev = 0;
    for (size_t j = 0; j < W4; j++) {
        uint32_t cur = word_at(d, n, j);
        err |= utf8_word_error(cur, prev);
        prev = cur;
    }
    return err != 0;
}

/* ------------------------------------------------------------------------ */
/* Phase 2: tokenization                                                      */
/* ------------------------------------------------------------------------ */

/* bitMapCreatorSimd, PSJ:446-571 (JSONL adds '\n' to op, PJL:453-456):
 * bitmap word k, bit j <-> byte 32k+j; bits at or past n are zero. */
static void build_char_bitmaps(const uint8_t *d, size_t n, unsigned flags, size_t W32,
                               uint32_t *bs, uint32_t *q, uint32_t *op, uint32_t *oc) {
    memset(bs, 0, W32 * 4);
    memset(q, 0, W32 * 4);
    memset(op, 0, W32 * 4);
    memset(oc, 0, W32 * 4);
    for (size_t p = 0; p < n; p++) {
        uint8_t c = d[p];
        uint32_t bit = 1u << (p & 31);
        size_t k = p >> 5;
        if (c == 0x5C) bs[k] |= bit;
        if (c == 0x22) q[k] |= bit;
        int is_oc = (c == 0x5B || c == 0x5D || c == 0x7B || c == 0x7D);
        int is_op = is_oc || c == 0x3A || c == 0x2C || ((flags & ORC_FLAG_JSONL) && c == 0x0A);
        if (is_oc) oc[k] |= bit;
        if (is_op) op[k] |= bit;
    }
}

/* findEscapedQuoteMerge_NEW, PSJ:766-834: overflow = parity of the maximal
 * backslash run ending just before word k, walking back over all-backslash
 * words (PSJ:810-816).  Walking off the unit start gives 0 (the reference
 * reads backslashes_GPU[-1] there). */
static uint32_t escape_overflow(const uint32_t *bs, size_t k) {
    if (k == 0) return 0;
    for (size_t j = k; j-- > 0;) {
        int run = clz32(~bs[j]);
        if (run != 32) return (uint32_t)(run & 1);
    }
    return 0;
}

static uint32_t unescaped_quotes(const uint32_t *bs_all, const uint32_t *q, size_t k) {
    const uint32_t evenBits = 0x55555555u, oddBits = ~evenBits;
    uint32_t quote = q[k], backslashes = bs_all[k];
    if ((quote & ((backslashes << 1) | 1u)) == 0) return quote; /* PSJ:802-808 */
    uint32_t overflow = escape_overflow(bs_all, k);
    backslashes = backslashes & ~overflow;                         /* PSJ:820 */
    uint32_t applyEscapedChar = (backslashes << 1) | overflow;      /* PSJ:821 */
    uint32_t oddSequence = backslashes & oddBits & ~applyEscapedChar;
    uint32_t sequenceStartatEven = oddSequence + backslashes;
    uint32_t invert_mask = sequenceStartatEven << 1;
    uint32_t escaped = (evenBits ^ invert_mask) & applyEscapedChar; /* PSJ:827 */
    return ~escaped & quote;                                        /* PSJ:830 */
}

/* prefix_xor, UTL:141-149 (inclusive). */
static inline uint32_t prefix_xor(uint32_t x) {
    x ^= x << 1;
    x ^= x << 2;
    x ^= x << 4;
    x ^= x << 8;
    x ^= x << 16;
    return x;
}

void orc_bitmaps(const uint8_t *d, size_t n, unsigned flags, uint32_t *o_bs, uint32_t *o_q,
                 uint32_t *o_op, uint32_t *o_oc, uint32_t *o_rq, uint32_t *o_instr,
                 uint32_t *o_st, uint32_t *o_ocm) {
    size_t W32 = (n + 31) / 32;
    uint32_t *bs = (uint32_t *)malloc(W32 * 4 + 4), *q = (uint32_t *)malloc(W32 * 4 + 4);
    uint32_t *op = (uint32_t *)malloc(W32 * 4 + 4), *oc = (uint32_t *)malloc(W32 * 4 + 4);
    build_char_bitmaps(d, n, flags, W32, bs, q, op, oc);
    uint32_t parity = 0; /* thrust::exclusive_scan of popcounts, PSJ:1163; only bit 0 is used */
    for (size_t k = 0; k < W32; k++) {
        uint32_t rq = unescaped_quotes(bs, q, k);
        uint32_t instr = prefix_xor(rq);               /* PSJ:858 */
        if (parity & 1) instr = ~instr;                /* PSJ:860 */
        parity += (uint32_t)popc32(rq);
        if (o_bs) o_bs[k] = bs[k];
        if (o_q) o_q[k] = q[k];
        if (o_op) o_op[k] = op[k];
        if (o_oc) o_oc[k] = oc[k];
        if (o_rq) o_rq[k] = rq;
        if (o_instr) o_instr[k] = instr;
        if (o_st) o_st[k] = ~instr & op[k];            /* PSJ:929 */
        if (o_ocm) o_ocm[k] = ~instr & oc[k];          /* PSJ:930 */
    }
    free(bs); free(q); free(op); free(oc);
}

/* ------------------------------------------------------------------------ */
/* Phase 3: structure recognition                                             */
/* ------------------------------------------------------------------------ */

int orc_pair(const uint8_t *oc_char, const uint32_t *oc_idx, uint64_t M, unsigned flags,
             uint64_t struct_base, int32_t *pair, uint32_t *keys_out, uint32_t *perm_out) {
    const int u16 = (flags & ORC_FLAG_DEPTH_U16) != 0;
    const uint32_t kmask = u16 ? 0xFFFFu : 0xFFu;
    const size_t nkeys = (size_t)kmask + 1;
    if (M == 0) return 0;
    uint32_t *key = (uint32_t *)malloc(M * 4);
    /* depth_init_MathAPI PSJ:1345-1358: opener -> +1, closer -> -1 (0xFF);
     * thrust::inclusive_scan on uint8, PSJ:1537 (mod 256);
     * transform_if(decrease, is_opening), PSJ:1540 / UTL:104-125. */
    uint32_t depth = 0;
    for (uint64_t m = 0; m < M; m++) {
        uint8_t c = oc_char[m];
        int open = (c == 0x5B || c == 0x7B);
        int close = (c == 0x5D || c == 0x7D);
        depth = (depth + (open ? 1u : (close ? kmask : 0u))) & kmask;
        key[m] = open ? ((depth - 1u) & kmask) : depth;
    }
    if (keys_out) memcpy(keys_out, key, M * 4);
    /* thrust::stable_sort_by_key, PSJ:1546: stable counting sort by key. */
    uint64_t *cnt = (uint64_t *)calloc(nkeys + 1, sizeof(uint64_t));
    for (uint64_t m = 0; m < M; m++) cnt[key[m] + 1]++;
    for (size_t k = 0; k < nkeys; k++) cnt[k + 1] += cnt[k];
    uint32_t *perm = (uint32_t *)malloc(M * 4);
    for (uint64_t m = 0; m < M; m++) perm[cnt[key[m]]++] = (uint32_t)m;
    free(cnt);
    if (perm_out) memcpy(perm_out, perm, M * 4);
    /* validate_expand_MathAPI_new2, PSJ:1431-1514: groups of four sorted
     * elements; pairs (k,k+1) and (k+2,k+3); the char array is signed char. */
    int err = 0;
    uint64_t groups = (M + 3) / 4;
    for (uint64_t i = 0; i < groups; i++) {
        uint64_t k = 4 * i;
#define C(x) ((int)(signed char)oc_char[perm[(x)]])
#define IDX(x) (oc_idx[perm[(x)]])
        if (i == groups - 1) {
            if (k + 1 >= M) {
                err = 1;                                                /* PSJ:1447-1449 */
            } else if (k + 2 >= M) {                                    /* PSJ:1451-1465 */
                if ((C(k) ^ C(k + 1)) != 6) err = 1;
                pair[IDX(k)] = (int32_t)(uint32_t)(IDX(k + 1) + struct_base + 1);
            } else if (k + 3 >= M) {
                err = 1;                                                /* PSJ:1467-1469 */
            } else {                                                    /* PSJ:1471-1486 */
                if ((C(k) ^ C(k + 1)) != 6 || (C(k + 2) ^ C(k + 3)) != 6) err = 1;
                pair[IDX(k)] = (int32_t)(uint32_t)(IDX(k + 1) + struct_base + 1);
                pair[IDX(k + 2)] = (int32_t)(uint32_t)(IDX(k + 3) + struct_base + 1);
            }
        } else {                                                        /* PSJ:1487-1509 */
            if (C(k) > C(k + 1) || C(k + 2) > C(k + 3)) err = 1;
            if ((C(k) ^ C(k + 1)) != 6 || (C(k + 2) ^ C(k + 3)) != 6) err = 1;
            pair[IDX(k)] = (int32_t)(uint32_t)(IDX(k + 1) + struct_base + 1);
            pair[IDX(k + 2)] = (int32_t)(uint32_t)(IDX(k + 3) + struct_base + 1);
        }
#undef C
#undef IDX
    }
    free(perm);
    free(key);
    return err;
}

int orc_parse_unit(const uint8_t *d, size_t n, unsigned flags, uint64_t pos_base,
                   uint64_t struct_base, uint32_t *out_struct, int32_t *out_pair,
                   uint64_t *out_N, uint64_t *out_M, int *out_utf8_err) {
    if (out_N) *out_N = 0;
    if (out_M) *out_M = 0;
    if (out_utf8_err) *out_utf8_err = 0;
    if (d == NULL || n == 0) return ORC_ERR_INPUT; /* PSJ:1597-1600 */

    int utf8_err = orc_utf8_check(d, n, flags);
    if (out_utf8_err) *out_utf8_err = utf8_err;

    size_t W32 = (n + 31) / 32;
    uint32_t *st = (uint32_t *)malloc(W32 * 4 + 4), *ocm = (uint32_t *)malloc(W32 * 4 + 4);
    orc_bitmaps(d, n, flags, NULL, NULL, NULL, NULL, NULL, NULL, st, ocm);

    /* removeCopy, PSJ:953-1021: positions are p + 1 + lastChunkIndex stored
     * as uint32 (PSJ:1007); brackets keep their char and structural index. */
    uint8_t *oc_char = (uint8_t *)malloc(n + 4);
    uint32_t *oc_idx = (uint32_t *)malloc((n + 1) * 4);
    uint64_t t = 0, m = 0;
    for (size_t k = 0; k < W32; k++) {
        uint32_t s = st[k];
        while (s) {
            int j = __builtin_ctz(s);
            s &= s - 1;
            size_t p = 32 * k + (size_t)j;
            if (p >= n) break;
            out_struct[t] = (uint32_t)(p + 1 + pos_base);
            if ((ocm[k] >> j) & 1u) {
                oc_char[m] = d[p];
                oc_idx[m] = (uint32_t)t;
                m++;
            }
            t++;
        }
    }
    free(st);
    free(ocm);
    for (uint64_t i = 0; i < t; i++) out_pair[i] = -1;
    int pair_err = orc_pair(oc_char, oc_idx, m, flags, struct_base, out_pair, NULL, NULL);
    free(oc_char);
    free(oc_idx);
    if (out_N) *out_N = t;
    if (out_M) *out_M = m;
    if (utf8_err) return ORC_ERR_UTF8;
    if (pair_err) return ORC_ERR_STRUCTURE;
    return ORC_OK;
}
