/*
 * cjx_gen.c - deterministic synthetic JSON for the benchmark configurations.
 *
 * Byte-reproducible on any host: xoshiro256** seeded through splitmix64.
 * The post-like object has 12 keys:
 *   id, id2     int64 uniform in [0, 2^62)
 *   score       float64 N(0, 1e3) printed with 17 significant digits
 *   flag        boolean, nil null, lang a 2-letter string
 *   s1..s4      strings, geometric lengths with mean 24; 2% of the characters
 *               escaped (\" \\ \n \uXXXX), 5% valid UTF-8 of 2/3/4 bytes
 *               in a 60/30/10 split, the rest printable ASCII (structural
 *               characters included, to exercise string masking)
 *   user        nested object of depth 3
 *   arr         0-8 int64
 * Standard documents are a top-level array of such objects; JSON Lines are
 * one object per line with a size target; fault injection plants exactly
 * one error of a chosen kind into a valid document.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint64_t s[4];
} rng_t;

static uint64_t splitmix64(uint64_t *x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static void rng_seed(rng_t *r, uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
}
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static uint64_t rng_next(rng_t *r) { /* xoshiro256** */
    uint64_t *s = r->s;
    uint64_t res = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return res;
}
static double rng_unit(rng_t *r) { return (rng_next(r) >> 11) * (1.0 / 9007199254740992.0); }
static uint64_t rng_below(rng_t *r, uint64_t n) { return n ? rng_next(r) % n : 0; }
static double rng_normal(rng_t *r) {
    double u1 = rng_unit(r), u2 = rng_unit(r);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

typedef struct {
    uint8_t *p;
    size_t n, cap;
    int overflow;
} out_t;

static void put(out_t *o, const void *s, size_t k) {
    if (o->n + k > o->cap) {
        o->overflow = 1;
        return;
    }
    memcpy(o->p + o->n, s, k);
    o->n += k;
}
static void puts_(out_t *o, const char *s) { put(o, s, strlen(s)); }
static void putc_(out_t *o, uint8_t c) { put(o, &c, 1); }

static void put_utf8(out_t *o, uint32_t cp) {
    uint8_t b[4];
    int k;
    if (cp < 0x80) { b[0] = (uint8_t)cp; k = 1; }
    else if (cp < 0x800) { b[0] = 0xC0 | (cp >> 6); b[1] = 0x80 | (cp & 0x3F); k = 2; }
    else if (cp < 0x10000) {
        b[0] = 0xE0 | (cp >> 12); b[1] = 0x80 | ((cp >> 6) & 0x3F); b[2] = 0x80 | (cp & 0x3F);
        k = 3;
    } else {
        b[0] = 0xF0 | (cp >> 18); b[1] = 0x80 | ((cp >> 12) & 0x3F);
        b[2] = 0x80 | ((cp >> 6) & 0x3F); b[3] = 0x80 | (cp & 0x3F);
        k = 4;
    }
    put(o, b, (size_t)k);
}

/* One string character following the configured mix. */
static void gen_char(rng_t *r, out_t *o) {
    double u = rng_unit(r);
    if (u < 0.02) { /* escape */
        switch (rng_below(r, 4)) {
            case 0: puts_(o, "\\\""); break;
            case 1: puts_(o, "\\\\"); break;
            case 2: puts_(o, "\\n"); break;
            default: {
                char buf[8];
                snprintf(buf, sizeof buf, "\\u%04x", (unsigned)(rng_below(r, 0xD000) + 0x20));
                puts_(o, buf);
            }
        }
    } else if (u < 0.07) { /* multi-byte UTF-8: 60/30/10 */
        double v = rng_unit(r);
        if (v < 0.6) put_utf8(o, 0x80 + (uint32_t)rng_below(r, 0x800 - 0x80));
        else if (v < 0.9) {
            uint32_t cp = 0x800 + (uint32_t)rng_below(r, 0x10000 - 0x800 - 0x800);
            if (cp >= 0xD800) cp += 0x800; /* skip surrogates */
            put_utf8(o, cp);
        } else put_utf8(o, 0x10000 + (uint32_t)rng_below(r, 0x110000 - 0x10000));
    } else {
        uint8_t c;
        do c = (uint8_t)(0x20 + rng_below(r, 0x5F)); while (c == '"' || c == '\\');
        putc_(o, c);
    }
}

static size_t geometric(rng_t *r, double mean) {
    double p = 1.0 / (mean + 1.0);
    double u = rng_unit(r);
    if (u < 1e-300) u = 1e-300;
    return (size_t)floor(log(u) / log(1.0 - p));
}

static void gen_string(rng_t *r, out_t *o, size_t len) {
    putc_(o, '"');
    for (size_t i = 0; i < len; i++) gen_char(r, o);
    putc_(o, '"');
}

static void put_u64(out_t *o, uint64_t v) {
    char buf[32];
    snprintf(buf, sizeof buf, "%llu", (unsigned long long)v);
    puts_(o, buf);
}

/* The post-like object.  extra_depth adds nested {"x": ...} wrappers inside
 * "user" (JSON Lines depth 1-6); pad_len adds a padding string key to reach a
 * size target; force_str_len > 0 makes s1 exactly that long (chunk-boundary
 * forcing). */
static void gen_object(rng_t *r, out_t *o, int extra_depth, size_t pad_len, size_t force_str_len) {
    char buf[64];
    puts_(o, "{\"id\":");
    put_u64(o, rng_next(r) >> 2);
    puts_(o, ",\"id2\":");
    put_u64(o, rng_next(r) >> 2);
    snprintf(buf, sizeof buf, ",\"score\":%.17g", rng_normal(r) * 1e3);
    puts_(o, buf);
    puts_(o, rng_next(r) & 1 ? ",\"flag\":true" : ",\"flag\":false");
    puts_(o, ",\"nil\":null,\"lang\":\"");
    putc_(o, (uint8_t)('a' + rng_below(r, 26)));
    putc_(o, (uint8_t)('a' + rng_below(r, 26)));
    puts_(o, "\"");
    for (int k = 1; k <= 4; k++) {
        snprintf(buf, sizeof buf, ",\"s%d\":", k);
        puts_(o, buf);
        gen_string(r, o, (k == 1 && force_str_len) ? force_str_len : geometric(r, 24.0));
    }
    puts_(o, ",\"user\":{\"name\":");
    gen_string(r, o, geometric(r, 24.0));
    puts_(o, ",\"lang\":\"en\",\"meta\":{\"level\":");
    put_u64(o, rng_below(r, 100));
    puts_(o, ",\"tags\":{\"k\":");
    gen_string(r, o, geometric(r, 8.0));
    for (int d = 0; d < extra_depth; d++) puts_(o, ",\"x\":{\"v\":1");
    for (int d = 0; d < extra_depth; d++) putc_(o, '}');
    puts_(o, "}}}");
    puts_(o, ",\"arr\":[");
    int na = (int)rng_below(r, 9);
    for (int i = 0; i < na; i++) {
        if (i) putc_(o, ',');
        put_u64(o, rng_next(r) >> 2);
    }
    putc_(o, ']');
    if (pad_len) {
        puts_(o, ",\"pad\":\"");
        for (size_t i = 0; i < pad_len; i++) putc_(o, (uint8_t)('a' + (i % 26)));
        putc_(o, '"');
    }
    putc_(o, '}');
}

/* Alternating nested arrays/objects to the given depth. */
static void gen_deep(out_t *o, int depth) {
    for (int d = 0; d < depth; d++) puts_(o, (d & 1) ? "{\"d\":" : "[");
    puts_(o, "0");
    for (int d = depth - 1; d >= 0; d--) putc_(o, (d & 1) ? '}' : ']');
}

/*
 * Standard document: a top-level array of post-like objects.
 *   n_objects       stop after this many objects (0 = no limit)
 *   target_bytes    stop once the document would exceed this size (0 = no limit)
 *   deep_depth      if > 0, one deep alternating branch is inserted after the first object
 *   boundary_every  if > 0, every boundary_every-th multiple of chunk_bytes is forced to fall
 *                   inside a string
 * Returns the document size, or 0 if cap is too small.
 */
size_t cjx_gen_standard(uint64_t seed, uint64_t n_objects, uint64_t target_bytes, int deep_depth,
                        uint64_t chunk_bytes, int boundary_every, uint8_t *buf, size_t cap) {
    rng_t r;
    rng_seed(&r, seed);
    out_t o = {buf, 0, cap, 0};
    putc_(&o, '[');
    uint64_t count = 0, next_forced = 0;
    if (boundary_every > 0 && chunk_bytes) next_forced = chunk_bytes * (uint64_t)boundary_every;
    const size_t reserve = 4096;
    for (;;) {
        if (n_objects && count >= n_objects) break;
        if (target_bytes && o.n + reserve >= target_bytes) break;
        if (count) putc_(&o, ',');
        size_t force = 0;
        if (next_forced && o.n + 2048 >= next_forced) {
            /* s1 starts roughly 150-250 bytes into the object; make it long
             * enough that the boundary lands well inside it */
            size_t dist = next_forced > o.n ? (size_t)(next_forced - o.n) : 0;
            force = dist + 256;
            next_forced += chunk_bytes * (uint64_t)boundary_every;
        }
        gen_object(&r, &o, 0, 0, force);
        count++;
        if (count == 1 && deep_depth > 0) {
            putc_(&o, ',');
            gen_deep(&o, deep_depth);
        }
        if (o.overflow) return 0;
    }
    putc_(&o, ']');
    return o.overflow ? 0 : o.n;
}

/*
 * One fixed-size segment of a large standard document, so segments can be
 * generated in parallel and concatenated: "[" + seg_0 + seg_1 + ... + "]".
 * Segment k holds whole objects ("," before each object except the very
 * first of the document) and is padded with spaces to exactly seg_bytes.
 * seg_start is the segment's offset in the final document, which places the
 * forced in-string chunk boundaries.  Returns seg_bytes, or 0 if too small.
 */
size_t cjx_gen_standard_segment(uint64_t seed, uint64_t seg_index, uint64_t seg_start,
                                uint64_t seg_bytes, int deep_depth, uint64_t chunk_bytes,
                                int boundary_every, uint8_t *buf) {
    rng_t r;
    rng_seed(&r, seed ^ (0x9E3779B97F4A7C15ull * (seg_index + 1)));
    out_t o = {buf, 0, seg_bytes, 0};
    uint64_t step = (boundary_every > 0 && chunk_bytes) ? chunk_bytes * (uint64_t)boundary_every : 0;
    uint64_t next_forced = 0;
    if (step) next_forced = ((seg_start / step) + 1) * step;
    const size_t reserve = 8192;
    uint64_t count = 0;
    for (;;) {
        if (o.n + reserve >= seg_bytes) break;
        if (seg_index > 0 || count > 0) putc_(&o, ',');
        size_t force = 0;
        uint64_t here = seg_start + o.n;
        if (next_forced && here + 2048 >= next_forced) {
            force = (size_t)(next_forced > here ? next_forced - here : 0) + 256;
            next_forced += step;
        }
        gen_object(&r, &o, 0, 0, force);
        count++;
        if (seg_index == 0 && count == 1 && deep_depth > 0) {
            putc_(&o, ',');
            gen_deep(&o, deep_depth);
        }
        if (o.overflow) return 0;
    }
    while (o.n < seg_bytes) putc_(&o, ' ');
    return o.overflow ? 0 : o.n;
}

/* One JSON Lines record of roughly target bytes (object + padding), extra
 * nesting depth in [0, 5] (record depth 1-6 counting the user branch), '\n'. */
size_t cjx_gen_record(uint64_t seed, uint64_t target, int extra_depth, uint8_t *buf, size_t cap) {
    rng_t r;
    rng_seed(&r, seed);
    out_t o = {buf, 0, cap, 0};
    gen_object(&r, &o, extra_depth, 0, 0);
    if (!o.overflow && o.n + 16 < target) {
        /* re-generate with padding to land on the target size */
        size_t base = o.n;
        rng_seed(&r, seed);
        o.n = 0;
        size_t pad = target - base - 10;
        gen_object(&r, &o, extra_depth, pad, 0);
    }
    putc_(&o, '\n');
    return o.overflow ? 0 : o.n;
}

/* JSON Lines buffer: n_records records, sizes log-normal(median, sigma)
 * clipped to [min_size, max_size].  Returns total bytes (0 on overflow). */
size_t cjx_gen_lines(uint64_t seed, uint64_t n_records, double median, double sigma,
                     uint64_t min_size, uint64_t max_size, uint8_t *buf, size_t cap) {
    rng_t r;
    rng_seed(&r, seed);
    size_t n = 0;
    for (uint64_t i = 0; i < n_records; i++) {
        double z = rng_normal(&r);
        double sz = median * exp(sigma * z);
        if (sz < (double)min_size) sz = (double)min_size;
        if (sz > (double)max_size) sz = (double)max_size;
        int depth = (int)rng_below(&r, 6);
        size_t k = cjx_gen_record(rng_next(&r), (uint64_t)sz, depth, buf + n, cap - n);
        if (!k) return 0;
        n += k;
    }
    return n;
}

/*
 * Plant exactly one fault into a valid document in place.
 *   kind 0 overlong UTF-8 (C0 AF), 1 surrogate (ED A0 80), 2 truncated
 *   sequence (E2 82 followed by ASCII), 3 unbalanced bracket (drop a closer),
 *   4 unterminated string (drop a closing quote).
 * The planted bytes replace bytes inside a string so the length is kept
 * (kinds 0-2), or one byte is overwritten with a space (kinds 3-4).
 * Returns 1 if a fault was planted.
 */
int cjx_plant_fault(uint64_t seed, int kind, uint8_t *buf, size_t n) {
    rng_t r;
    rng_seed(&r, seed);
    if (n < 16) return 0;
    if (kind <= 2) {
        /* find a run of 3 plain ASCII letters inside a string after a random offset */
        size_t start = (size_t)rng_below(&r, n / 2);
        for (size_t p = start; p + 3 < n; p++) {
            if (buf[p] >= 'a' && buf[p] <= 'z' && buf[p + 1] >= 'a' && buf[p + 1] <= 'z' &&
                buf[p + 2] >= 'a' && buf[p + 2] <= 'z') {
                if (kind == 0) { buf[p] = 0xC0; buf[p + 1] = 0xAF; }
                else if (kind == 1) { buf[p] = 0xED; buf[p + 1] = 0xA0; buf[p + 2] = 0x80; }
                else { buf[p] = 0xE2; buf[p + 1] = 0x82; }
                return 1;
            }
        }
        return 0;
    }
    if (kind == 3) {
        for (size_t p = n; p-- > 0;)
            if (buf[p] == '}' || buf[p] == ']') {
                buf[p] = ' ';
                return 1;
            }
        return 0;
    }
    /* kind 4: remove the closing quote of the last string value */
    for (size_t p = n; p-- > 1;)
        if (buf[p] == '"' && buf[p - 1] != '\\') {
            buf[p] = ' ';
            return 1;
        }
    return 0;
}
