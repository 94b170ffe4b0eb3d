/*
 * Runs the oracle (oracle/oracle.c) over the cases of a file; built with
 * AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_sanitizers.py,
 * which compares every printed line with the same call through the oracle's
 * ordinary build.
 *
 *   oracle_san <cases file>
 *
 * A case is: u32 kind, u32 flags, u64 n, then n bytes.  One line per case:
 *   kind 0  orc_parse_unit          "parse <code> <N> <M> <utf8> <hash struct> <hash pair>"
 *   kind 1  orc_utf8_check          "utf8 <result>"
 *   kind 2  orc_utf8_valid_strict   "strict <result>"
 *   kind 3  orc_bitmaps             "bitmaps <hash of the eight bitmaps>"
 *   kind 4  orc_pair over the brackets that orc_parse_unit finds
 *                                   "pair <result> <hash pair> <hash keys> <hash perm>"
 * A hash is sum over i of (i + 1) * v_i, v_i taken as uint32, mod 2^64.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/oracle.h"

static uint64_t hash32(const void *p, size_t count) {
    const uint32_t *v = (const uint32_t *)p;
    uint64_t h = 0;
    for (size_t i = 0; i < count; i++) h += (uint64_t)(i + 1) * (uint64_t)v[i];
    return h;
}

static int read_exact(FILE *f, void *p, size_t n) { return n == 0 || fread(p, 1, n, f) == n; }

static int is_bracket(uint8_t c) { return c == '[' || c == ']' || c == '{' || c == '}'; }

static void run_case(uint32_t kind, uint32_t flags, const uint8_t *data, size_t n) {
    const size_t cap = n ? n : 1;
    if (kind == 1) {
        printf("utf8 %d\n", orc_utf8_check(data, n, flags));
        return;
    }
    if (kind == 2) {
        printf("strict %d\n", orc_utf8_valid_strict(data, n));
        return;
    }
    if (kind == 3) {
        const size_t w = (n + 31) / 32;
        uint32_t *m = (uint32_t *)calloc(8 * (w ? w : 1), 4);
        orc_bitmaps(data, n, flags, m, m + w, m + 2 * w, m + 3 * w, m + 4 * w, m + 5 * w,
                    m + 6 * w, m + 7 * w);
        printf("bitmaps %llu\n", (unsigned long long)hash32(m, 8 * w));
        free(m);
        return;
    }
    uint32_t *st = (uint32_t *)malloc(cap * 4);
    int32_t *pr = (int32_t *)malloc(cap * 4);
    uint64_t N = 0, M = 0;
    int u8 = 0;
    const int code = orc_parse_unit(data, n, flags, 0, 0, st, pr, &N, &M, &u8);
    if (kind == 0) {
        printf("parse %d %llu %llu %d %llu %llu\n", code, (unsigned long long)N,
               (unsigned long long)M, u8, (unsigned long long)hash32(st, N),
               (unsigned long long)hash32(pr, N));
    } else {
        /* the brackets among the structurals: their bytes and rows */
        uint8_t *oc = (uint8_t *)malloc(N ? N : 1);
        uint32_t *idx = (uint32_t *)malloc((N ? N : 1) * 4);
        uint64_t m = 0;
        for (uint64_t i = 0; i < N; i++) {
            const uint64_t at = (uint64_t)st[i] - 1;
            if (at < n && is_bracket(data[at])) {
                oc[m] = data[at];
                idx[m] = (uint32_t)i;
                m++;
            }
        }
        int32_t *pair = (int32_t *)malloc((N ? N : 1) * 4);
        for (uint64_t i = 0; i < N; i++) pair[i] = -1;
        uint32_t *keys = (uint32_t *)malloc((m ? m : 1) * 4);
        uint32_t *perm = (uint32_t *)malloc((m ? m : 1) * 4);
        const int r = orc_pair(oc, idx, m, flags, 0, pair, keys, perm);
        printf("pair %d %llu %llu %llu\n", r, (unsigned long long)hash32(pair, N),
               (unsigned long long)hash32(keys, m), (unsigned long long)hash32(perm, m));
        free(oc);
        free(idx);
        free(pair);
        free(keys);
        free(perm);
    }
    free(st);
    free(pr);
}

int main(int argc, char **argv) {
    if (argc != 2) return 2;
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 3;
    for (;;) {
        uint32_t hdr[2];
        uint64_t n;
        if (fread(hdr, 4, 2, f) != 2) break;
        if (!read_exact(f, &n, 8)) return 4;
        /* exactly n bytes, so reads past the end are caught */
        uint8_t *data = (uint8_t *)malloc(n ? n : 1);
        if (!read_exact(f, data, n)) return 4;
        run_case(hdr[0], hdr[1], n ? data : NULL, n);
        free(data);
    }
    fclose(f);
    return 0;
}
