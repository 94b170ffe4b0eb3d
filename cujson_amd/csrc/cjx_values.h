// cjx_values.h - decoding of JSON scalars (numbers, strings, literals), for
// the typed value table (cjx_values.hip).  Host and device code, so that the
// host exports used by the tests run exactly what the kernel runs.
//
// Numbers (RFC 8259 grammar) become int64 when they are integers in range,
// otherwise binary64, correctly rounded (round half to even):
//   1. w = the first 19 significant digits, x = (w + f) * 10^q with f in
//      [0, 1) standing for digits beyond the 19th.
//   2. Clinger's fast path: w <= 2^53, |q| <= 22, no dropped digits: one
//      IEEE multiplication or division by an exact power of ten.
//   3. Otherwise w (shifted to bit 63) times the 128-bit truncated mantissa of
//      5^q (cjx_pow5.h) gives a 192-bit P' with the true product in
//      [P', P' + width): width < 2^64 from the truncation of 5^q, larger when
//      digits were dropped.  The rounding is decided from P' unless that
//      interval holds a halfway point between two doubles (or crosses the
//      top bit); then
//   4. an exact comparison of the decimal value with the halfway point, in
//      big integers, for at most 800 significant digits (exact_round_up).  Values
//      with more digits that land there keep the nearest candidate of step 3
//      and the flag CJX_VALUE_APPROX.
// Strings: escapes \" \\ \/ \b \f \n \r \t and \uXXXX (surrogate pairs to
// 4-byte UTF-8; an unpaired surrogate is written as its 3-byte form and
// flagged); a raw byte below 0x20 or an unknown escape makes the text
// invalid.
#pragma once
#include <stdint.h>

#include "cjx_pow5.h"

#define CJX_HD __host__ __device__ __forceinline__

namespace cjx {

// Value types and flags (the C ABI's CJX_VALUE_*).
enum : uint8_t {
    kValNone = 0,
    kValKey = 1,
    kValString = 2,
    kValInt = 3,
    kValFloat = 4,
    kValTrue = 5,
    kValFalse = 6,
    kValNull = 7,
    kValInvalid = 8
};
enum : uint8_t {
    kValEscaped = 1,   // the string held escapes
    kValBigInt = 2,    // an integer outside int64, given as binary64
    kValUnpaired = 4,  // an unpaired surrogate escape (3-byte form written)
    kValApprox = 8,    // float whose rounding was not decided exactly
    kValExact = 16     // internal: step 4 is needed
};

CJX_HD bool is_digit(uint8_t c) { return (uint8_t)(c - '0') < 10u; }
CJX_HD bool is_ws(uint8_t c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; }

CJX_HD void mul_64x64(uint64_t a, uint64_t b, uint64_t &hi, uint64_t &lo) {
#if defined(__HIP_DEVICE_COMPILE__)
    hi = __umul64hi(a, b);
    lo = a * b;
#else
    const unsigned __int128 p = (unsigned __int128)a * b;
    hi = (uint64_t)(p >> 64);
    lo = (uint64_t)p;
#endif
}

CJX_HD int clz64(uint64_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __clzll((long long)x);
#else
    return __builtin_clzll(x);
#endif
}

CJX_HD double bits_to_double(uint64_t b) {
    union {
        uint64_t u;
        double d;
    } v;
    v.u = b;
    return v.d;
}

// Parsed number text: sign, the first 19 significant digits w, their
// decimal exponent q (x = (w + f) 10^q), whether nonzero digits were dropped,
// the integer form, and where the digits lie (for step 4).
struct NumText {
    bool neg, integer, dropped;
    uint64_t w;
    int64_t q;
    int nd;                    // significant digits in all
    uint32_t d0, dot, dend;    // digits span [d0, dend) with the '.' at dot (or dend)
    uint32_t exp_end;
    int64_t exp;               // the written exponent (saturated)
};

// RFC 8259 number grammar over s[0..len); false if the text is no number.
CJX_HD bool scan_number(const uint8_t *s, uint32_t len, NumText &t) {
    uint32_t i = 0;
    t.neg = false;
    if (i < len && s[i] == '-') {
        t.neg = true;
        i++;
    }
    if (i >= len) return false;
    t.d0 = i;
    if (s[i] == '0') {
        i++;
    } else if (s[i] >= '1' && s[i] <= '9') {
        while (i < len && is_digit(s[i])) i++;
    } else {
        return false;
    }
    t.dot = i;
    t.integer = true;
    if (i < len && s[i] == '.') {
        t.integer = false;
        i++;
        if (i >= len || !is_digit(s[i])) return false;
        while (i < len && is_digit(s[i])) i++;
    }
    t.dend = i;
    t.exp = 0;
    if (i < len && (s[i] == 'e' || s[i] == 'E')) {
        t.integer = false;
        i++;
        bool eneg = false;
        if (i < len && (s[i] == '+' || s[i] == '-')) {
            eneg = s[i] == '-';
            i++;
        }
        if (i >= len || !is_digit(s[i])) return false;
        int64_t e = 0;
        while (i < len && is_digit(s[i])) {
            if (e < 100000000) e = e * 10 + (s[i] - '0');
            i++;
        }
        t.exp = eneg ? -e : e;
    }
    t.exp_end = i;
    if (i != len) return false;
    // significant digits
    t.w = 0;
    t.nd = 0;
    t.dropped = false;
    int frac_digits = 0, kept_after = 0;
    bool started = false;
    for (uint32_t j = t.d0; j < t.dend; j++) {
        if (j == t.dot) continue;  // the '.'
        const uint8_t c = s[j];
        const bool frac = j > t.dot;
        if (frac) frac_digits++;
        if (!started && c == '0') continue;
        started = true;
        t.nd++;
        if (t.nd <= 19) {
            t.w = t.w * 10 + (uint64_t)(c - '0');
        } else {
            kept_after++;  // a digit beyond the 19th
            if (c != '0') t.dropped = true;
        }
    }
    t.q = t.exp - frac_digits + kept_after;
    return true;
}

// ---------------------------------------------------------------------------
// 192-bit helpers: limbs [0] low .. [2] high.
CJX_HD void add192(uint64_t (&a)[3], const uint64_t (&b)[3]) {
    uint64_t c = 0;
    for (int k = 0; k < 3; k++) {
        const uint64_t s = a[k] + b[k];
        const uint64_t c1 = s < a[k] ? 1u : 0u;
        const uint64_t s2 = s + c;
        const uint64_t c2 = s2 < s ? 1u : 0u;
        a[k] = s2;
        c = c1 | c2;
    }
}
CJX_HD int cmp192(const uint64_t (&a)[3], const uint64_t (&b)[3]) {
    for (int k = 2; k >= 0; k--)
        if (a[k] != b[k]) return a[k] < b[k] ? -1 : 1;
    return 0;
}
// a mod 2^bits (bits in [0, 192])
CJX_HD void low_bits192(const uint64_t (&a)[3], int bits, uint64_t (&r)[3]) {
    for (int k = 0; k < 3; k++) {
        const int lo = 64 * k;
        if (bits >= lo + 64) r[k] = a[k];
        else if (bits <= lo) r[k] = 0;
        else r[k] = a[k] & ((1ull << (bits - lo)) - 1u);
    }
}
CJX_HD void pow2_192(int bit, uint64_t (&r)[3]) {  // 2^bit, bit in [0, 191]
    r[0] = r[1] = r[2] = 0;
    r[bit >> 6] = 1ull << (bit & 63);
}
// a >> s (s in [0, 191]) as 64 bits (the caller knows it fits)
CJX_HD uint64_t shr192_64(const uint64_t (&a)[3], int s) {
    const int k = s >> 6, r = s & 63;
    uint64_t lo = a[k], hi = k + 1 < 3 ? a[k + 1] : 0;
    return r ? (lo >> r) | (hi << (64 - r)) : lo;
}

// Steps 2-3.  Returns the binary64 bits of (w + f) 10^q, sets kValExact when
// step 4 must decide, and returns in m_lo / e_ulp the lower candidate
// (m_lo * 2^e_ulp) for it.
CJX_HD uint64_t approx_double(uint64_t w, int64_t q, bool dropped, uint8_t &flags, uint64_t &m_lo,
                              int64_t &e_ulp) {
    if (w == 0) return 0;
    if (q < kPow5Min) return 0;                     // below 10^-324: rounds to 0
    if (q > kPow5Max) return 0x7FF0000000000000ull;  // above 10^309: infinity
    if (!dropped && w <= (1ull << 53) && q >= -22 && q <= 22) {
        constexpr double p10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,
                                    1e8,  1e9,  1e10, 1e11, 1e12, 1e13, 1e14, 1e15,
                                    1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
        double d = (double)w;
        d = q < 0 ? d / p10[-q] : d * p10[q];
        union {
            double d;
            uint64_t u;
        } v;
        v.d = d;
        return v.u;
    }
    const int lz = clz64(w);
    const uint64_t wn = w << lz;
    const uint64_t th = kPow5[q - kPow5Min][0], tl = kPow5[q - kPow5Min][1];
    uint64_t h1, l1, h2, l2;
    mul_64x64(wn, th, h1, l1);
    mul_64x64(wn, tl, h2, l2);
    uint64_t P[3] = {l2, l1 + h2, h1 + ((l1 + h2) < l1 ? 1u : 0u)};
    // width of the interval holding the true product
    const bool t_exact = q >= 0 && q <= 55;
    uint64_t W[3] = {0, 0, 0};
    if (dropped) {  // (wn + 2^lz)(T + 1) - wn T = wn + 2^lz T + 2^lz
        uint64_t a[3] = {wn, 0, 0};
        uint64_t b[3] = {tl << lz, lz ? (th << lz) | (tl >> (64 - lz)) : th, lz ? th >> (64 - lz) : 0};
        uint64_t c[3] = {1ull << lz, 0, 0};
        add192(W, a);
        add192(W, b);
        add192(W, c);
    } else if (!t_exact) {
        W[0] = wn;
    }
    const int upper = (int)(P[2] >> 63);
    const int t = 190 + upper;  // top bit of P'
    // exponent of a unit of P' in the value: x = P 2^(e_unit)
    // 5^q = T 2^(-k_q), k_q = 127 - floor(q log2 5); floor(q log2 10) = (217706 q) >> 16
    const int64_t fl10 = (217706 * q) >> 16;
    const int64_t e_unit = fl10 - 127 - lz;  // = floor(q log2 5) + q - 127 - lz
    // normal: round bit at t - 53; biased exponent of the result
    int64_t be = e_unit + t + 1023;
    int cut = t - 53;
    if (be <= 0) cut += (int)(1 - be);  // subnormal: fewer mantissa bits
    if (cut > 191) {  // below half the smallest subnormal, unless W carries it there
        if (W[0] | W[1] | W[2]) flags |= kValExact;
        m_lo = 0;  // candidates: 0 and the smallest subnormal 2^-1074
        e_ulp = -1074;
        return 0;
    }
    uint64_t u[3], B[3];
    low_bits192(P, cut + 1, u);
    pow2_192(cut, B);
    const int c = cmp192(u, B);
    bool up;
    if (W[0] | W[1] | W[2]) {
        // ambiguous if [P', P' + W) holds B (mod 2B) or the top bit moves
        uint64_t uw[3] = {u[0], u[1], u[2]};
        add192(uw, W);
        bool amb = c <= 0 && cmp192(uw, B) > 0;
        if (!upper) {
            uint64_t pw[3] = {P[0], P[1], P[2]};
            add192(pw, W);
            if (pw[2] >> 63) amb = true;
        }
        if (amb) flags |= kValExact;
        up = c > 0;
    } else {
        up = c > 0 || (c == 0 && cut + 1 < 192 && ((shr192_64(P, cut + 1) & 1u) != 0));  // tie: to even
    }
    uint64_t m = cut + 1 < 192 ? shr192_64(P, cut + 1) : 0u;
    m_lo = m;
    e_ulp = e_unit + cut + 1;
    if (up) m++;
    if (be <= 0) {  // subnormal (m < 2^52) or rounded up to the smallest normal (m = 2^52)
        return m;
    }
    if (m >> 53) {
        m >>= 1;
        be++;
    }
    if (be >= 2047) return 0x7FF0000000000000ull;
    return ((uint64_t)be << 52) | (m & ((1ull << 52) - 1u));
}

// ---------------------------------------------------------------------------
// Step 4: exact comparison with the halfway point, in big integers.
constexpr int kExactDigits = 800;
struct Big {
    static constexpr int kLimbs = 128;  // 4096 bits
    uint32_t v[kLimbs];
    int n;  // limbs in use
    bool overflow;
};
CJX_HD void big_set(Big &b, uint64_t x) {
    b.v[0] = (uint32_t)x;
    b.v[1] = (uint32_t)(x >> 32);
    b.n = b.v[1] ? 2 : (b.v[0] ? 1 : 0);
    b.overflow = false;
}
CJX_HD void big_mul_add(Big &b, uint32_t m, uint32_t a) {
    uint64_t carry = a;
    for (int k = 0; k < b.n; k++) {
        const uint64_t p = (uint64_t)b.v[k] * m + carry;
        b.v[k] = (uint32_t)p;
        carry = p >> 32;
    }
    if (carry) {
        if (b.n == Big::kLimbs) b.overflow = true;
        else b.v[b.n++] = (uint32_t)carry;
    }
}
CJX_HD void big_mul_pow5(Big &b, int64_t k) {
    while (k >= 13) {
        big_mul_add(b, 1220703125u, 0);  // 5^13
        k -= 13;
    }
    uint32_t m = 1;
    while (k-- > 0) m *= 5;
    if (m != 1) big_mul_add(b, m, 0);
}
CJX_HD void big_shl(Big &b, int64_t s) {
    if (b.n == 0 || s <= 0) return;
    const int64_t limbs = s >> 5;
    const int r = (int)(s & 31);
    if (b.n + limbs + 1 > Big::kLimbs) {
        b.overflow = true;
        return;
    }
    int n2 = b.n + (int)limbs + 1;
    for (int k = n2 - 1; k >= 0; k--) {
        const int64_t src = k - limbs;
        const uint32_t hi = (src >= 0 && src < b.n) ? b.v[src] : 0u;
        const uint32_t lo = (src - 1 >= 0 && src - 1 < b.n) ? b.v[src - 1] : 0u;
        b.v[k] = r ? (hi << r) | (lo >> (32 - r)) : hi;
    }
    while (n2 > 0 && b.v[n2 - 1] == 0) n2--;
    b.n = n2;
}
CJX_HD int big_cmp(const Big &a, const Big &b) {
    if (a.n != b.n) return a.n < b.n ? -1 : 1;
    for (int k = a.n - 1; k >= 0; k--)
        if (a.v[k] != b.v[k]) return a.v[k] < b.v[k] ? -1 : 1;
    return 0;
}

// Decides between m_lo 2^e_ulp and (m_lo + 1) 2^e_ulp for the number text
// s[0..len): compares the exact value with (2 m_lo + 1) 2^(e_ulp - 1).
// Returns false when it cannot decide (more than kExactDigits significant
// digits or an overflow of the big integers).
CJX_HD bool exact_round_up(const uint8_t *s, const NumText &t, uint64_t m_lo, int64_t e_ulp,
                           bool &up) {
    if (t.nd > kExactDigits) return false;
    Big a, b;
    big_set(a, 0);
    bool started = false;
    int frac_digits = 0;
    for (uint32_t j = t.d0; j < t.dend; j++) {
        if (j == t.dot) continue;
        if (j > t.dot) frac_digits++;
        const uint8_t c = s[j];
        if (!started && c == '0') continue;
        started = true;
        if (a.n == 0) big_set(a, (uint64_t)(c - '0'));
        else big_mul_add(a, 10u, (uint32_t)(c - '0'));
    }
    const int64_t q = t.exp - frac_digits;  // value = a 10^q
    big_set(b, 2 * m_lo + 1);               // halfway = b 2^(e_ulp - 1)
    if (q >= 0) big_mul_pow5(a, q);
    else big_mul_pow5(b, -q);
    const int64_t z = q - (e_ulp - 1);  // powers of two: a 2^q vs b 2^(e_ulp - 1)
    if (z >= 0) big_shl(a, z);
    else big_shl(b, -z);
    if (a.overflow || b.overflow) return false;
    const int c = big_cmp(a, b);
    up = c > 0 || (c == 0 && (m_lo & 1u));
    return true;
}

// Encodes m * 2^e_ulp where m is the chosen candidate (the same exponent
// bookkeeping as approx_double).
CJX_HD uint64_t encode_double(uint64_t m, int64_t e_ulp) {
    // value = m 2^e_ulp; normalise m to 53 bits
    if (m == 0) return 0;
    int64_t e = e_ulp;
    while (m >> 53) {  // only a carry out of the rounding: m = 2^53
        m >>= 1;
        e++;
    }
    if (m < (1ull << 52)) {  // subnormal candidates have e_ulp = -1074
        return m;            // the exponent field stays 0 (or 1 when m = 2^52)
    }
    const int64_t be = e + 1075;  // m 2^e with m in [2^52, 2^53): unbiased exponent e + 52
    if (be >= 2047) return 0x7FF0000000000000ull;
    return ((uint64_t)be << 52) | (m & ((1ull << 52) - 1u));
}

// Parses the number text s[0..len): returns kValInt (iv), kValFloat (dv
// bits in iv) or kValInvalid; flags as above (kValExact when step 4 is
// still due, with m_lo / e_ulp for it).
CJX_HD uint8_t parse_number(const uint8_t *s, uint32_t len, int64_t &iv, uint8_t &flags,
                            NumText &t, uint64_t &m_lo, int64_t &e_ulp) {
    if (!scan_number(s, len, t)) return kValInvalid;
    if (t.integer && t.nd <= 19) {
        const uint64_t lim = t.neg ? (1ull << 63) : (1ull << 63) - 1u;
        if (t.w <= lim) {
            iv = t.neg ? (int64_t)(0 - t.w) : (int64_t)t.w;
            return kValInt;
        }
    }
    if (t.integer) flags |= kValBigInt;
    uint64_t bits = approx_double(t.w, t.q, t.dropped, flags, m_lo, e_ulp);
    if (t.neg) bits |= 1ull << 63;
    iv = (int64_t)bits;
    return kValFloat;
}

// Completes step 4 for a number whose flags hold kValExact.
CJX_HD void finish_exact(const uint8_t *s, const NumText &t, uint64_t m_lo, int64_t e_ulp,
                         int64_t &iv, uint8_t &flags) {
    bool up = false;
    flags &= (uint8_t)~kValExact;
    if (!exact_round_up(s, t, m_lo, e_ulp, up)) {
        flags |= kValApprox;
        return;
    }
    uint64_t bits = encode_double(m_lo + (up ? 1u : 0u), e_ulp);
    if (t.neg) bits |= 1ull << 63;
    iv = (int64_t)bits;
}

// ---------------------------------------------------------------------------
CJX_HD int hex4(const uint8_t *s) {
    int v = 0;
    for (int k = 0; k < 4; k++) {
        const uint8_t c = s[k];
        int d;
        if (c >= '0' && c <= '9') d = c - '0';
        else if (c >= 'a' && c <= 'f') d = c - 'a' + 10;
        else if (c >= 'A' && c <= 'F') d = c - 'A' + 10;
        else return -1;
        v = v * 16 + d;
    }
    return v;
}
CJX_HD uint32_t put_utf8(uint8_t *out, uint32_t cp) {
    if (cp < 0x80) {
        out[0] = (uint8_t)cp;
        return 1;
    }
    if (cp < 0x800) {
        out[0] = (uint8_t)(0xC0 | (cp >> 6));
        out[1] = (uint8_t)(0x80 | (cp & 0x3F));
        return 2;
    }
    if (cp < 0x10000) {
        out[0] = (uint8_t)(0xE0 | (cp >> 12));
        out[1] = (uint8_t)(0x80 | ((cp >> 6) & 0x3F));
        out[2] = (uint8_t)(0x80 | (cp & 0x3F));
        return 3;
    }
    out[0] = (uint8_t)(0xF0 | (cp >> 18));
    out[1] = (uint8_t)(0x80 | ((cp >> 12) & 0x3F));
    out[2] = (uint8_t)(0x80 | ((cp >> 6) & 0x3F));
    out[3] = (uint8_t)(0x80 | (cp & 0x3F));
    return 4;
}

// Unescapes the string body s[0..len) (between the quotes) into out (may be
// null: lengths only); returns false for an invalid body (a raw byte below
// 0x20, an unknown escape, an unescaped quote).  The output is never longer
// than the body.
typedef uint32_t __attribute__((may_alias)) word_alias;

// True if a byte of w is below 0x20, a quote or a backslash (exact as a
// whole: each test is the zero-byte test on w, w ^ 0x22.. or w ^ 0x5C..,
// and the below-0x20 test of hasless, all exact for any set byte).
CJX_HD bool word_special(uint32_t w) {
    const uint32_t q = w ^ 0x22222222u, b = w ^ 0x5C5C5C5Cu;
    const uint32_t lo = (w - 0x20202020u) & ~w;
    const uint32_t zq = (q - 0x01010101u) & ~q, zb = (b - 0x01010101u) & ~b;
    return ((lo | zq | zb) & 0x80808080u) != 0;
}

CJX_HD bool unescape(const uint8_t *s, uint32_t len, uint8_t *out, uint32_t &out_len,
                     uint8_t &flags) {
    uint32_t o = 0, i0 = 0;
    // Up to the first escape the output offset equals the input offset, so
    // when both pointers share their alignment mod 4, whole 4-byte words
    // without a special byte are copied at once.
    if (out && ((uintptr_t)s & 3u) == ((uintptr_t)out & 3u)) {
        uint32_t i = 0;
        while (i < len && ((uintptr_t)(s + i) & 3u)) {
            const uint8_t c = s[i];
            if (c < 0x20 || c == '"' || c == '\\') break;
            out[i] = c;
            i++;
        }
        if (i < len && !((uintptr_t)(s + i) & 3u)) {
            for (; i + 4 <= len; i += 4) {
                const uint32_t w = *reinterpret_cast<const word_alias *>(s + i);
                if (word_special(w)) break;
                *reinterpret_cast<word_alias *>(out + i) = w;
            }
        }
        i0 = o = i;
    }
    for (uint32_t i = i0; i < len;) {
        const uint8_t c = s[i];
        if (c < 0x20 || c == '"') return false;
        if (c != '\\') {
            if (out) out[o] = c;
            o++;
            i++;
            continue;
        }
        flags |= kValEscaped;
        if (i + 1 >= len) return false;
        const uint8_t e = s[i + 1];
        uint8_t r;
        switch (e) {
            case '"': r = '"'; break;
            case '\\': r = '\\'; break;
            case '/': r = '/'; break;
            case 'b': r = '\b'; break;
            case 'f': r = '\f'; break;
            case 'n': r = '\n'; break;
            case 'r': r = '\r'; break;
            case 't': r = '\t'; break;
            case 'u': r = 0; break;
            default: return false;
        }
        if (e != 'u') {
            if (out) out[o] = r;
            o++;
            i += 2;
            continue;
        }
        if (i + 6 > len) return false;
        int cp = hex4(s + i + 2);
        if (cp < 0) return false;
        i += 6;
        if (cp >= 0xD800 && cp <= 0xDBFF && i + 6 <= len && s[i] == '\\' && s[i + 1] == 'u') {
            const int lo = hex4(s + i + 2);
            if (lo >= 0xDC00 && lo <= 0xDFFF) {
                cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                i += 6;
            }
        }
        if (cp >= 0xD800 && cp <= 0xDFFF) flags |= kValUnpaired;
        uint8_t tmp[4];
        const uint32_t k = put_utf8(tmp, (uint32_t)cp);
        if (out)
            for (uint32_t j = 0; j < k; j++) out[o + j] = tmp[j];
        o += k;
    }
    out_len = o;
    return true;
}

}  // namespace cjx
