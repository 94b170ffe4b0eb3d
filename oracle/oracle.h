/*
 * oracle.h - single-threaded CPU restatement of cuJSON's three-phase parse.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation: it may be loaded by tests/, by __graft_entry__.smoke() and
 * by bench.py's cpu_baseline leg, and by nothing else.  The product library
 * (libcujson_amd.so) never links or calls it.
 *
 * Every function restates one step of the reference algorithm and cites the
 * reference file:line it follows (PSJ = cujson/parse_standard_json.cu,
 * PJL = cujson/parse_json_lines.cu, UTL = cujson/utils.cu).
 */
#ifndef CJX_ORACLE_H
#define CJX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes: identical numbering to the product ABI (include/cujson_amd.h). */
enum { ORC_OK = 0, ORC_ERR_INPUT = 1, ORC_ERR_UTF8 = 2, ORC_ERR_STRUCTURE = 3 };

/* Flags: identical bit values to the product ABI. */
#define ORC_FLAG_UTF8_STRICT 0x1u /* RFC 3629 validation instead of the reference's weak check */
#define ORC_FLAG_DEPTH_U16   0x2u /* 16-bit depth keys instead of the reference's uint8 keys   */
#define ORC_FLAG_JSONL       0x4u /* '\n' is a structural character (JSON Lines)              */

/*
 * Parse one unit (a standard-JSON document, or one JSON Lines chunk).
 *
 *   data, n        unit bytes (no padding needed; bytes past n read as zero)
 *   pos_base       added to every 1-based byte position (PJL lastChunkIndex)
 *   struct_base    added to every pair value (PJL lastStructuralIndex)
 *   out_struct     capacity >= n; receives N positions (uint32, wraps like PSJ:1007)
 *   out_pair       capacity >= n; receives N pair cells, -1 where the reference
 *                  leaves the cell unwritten (PSJ:1266 never initialises them)
 *   out_N, out_M   structural count and bracket count
 *   out_utf8_err   nonzero if the UTF-8 phase flags the unit
 *
 * Returns ORC_OK, ORC_ERR_INPUT (n == 0 or data == NULL), ORC_ERR_UTF8 or
 * ORC_ERR_STRUCTURE, with the reference's precedence (PSJ:1597, :1625, :1567).
 * All outputs are produced for every non-empty unit, errors included.
 */
int orc_parse_unit(const uint8_t *data, size_t n, unsigned flags, uint64_t pos_base,
                   uint64_t struct_base, uint32_t *out_struct, int32_t *out_pair,
                   uint64_t *out_N, uint64_t *out_M, int *out_utf8_err);

/* Phase 1 alone (PSJ:264-327): 1 if the unit fails the UTF-8 check. */
int orc_utf8_check(const uint8_t *data, size_t n, unsigned flags);

/* RFC 3629 validity of a byte string (strict mode). 1 = valid. */
int orc_utf8_valid_strict(const uint8_t *data, size_t n);

/*
 * Phase 2 bitmaps for inspection: W32 = ceil(n/32) words each.
 * Any pointer may be NULL.  bs/quote/op/oc are the raw character bitmaps
 * (PSJ:446-571), rq the unescaped-quote bitmap (PSJ:766-834), instr the
 * in-string mask (PSJ:850-862), st/ocm the selected structural/bracket
 * bitmaps (PSJ:915-937).
 */
void orc_bitmaps(const uint8_t *data, size_t n, unsigned flags, uint32_t *bs, uint32_t *quote,
                 uint32_t *op, uint32_t *oc, uint32_t *rq, uint32_t *instr, uint32_t *st,
                 uint32_t *ocm);

/*
 * Phase 3 alone over a compact bracket array (PSJ:1519-1582):
 *   oc_char[M] bracket bytes, oc_idx[M] their structural indices.
 *   keys_out[M] (optional) depth keys before sorting, perm_out[M] (optional)
 *   the stable-sorted order as indices into oc_char.
 *   pair[] (capacity N) receives the pair writes.
 * Returns 1 on a pairing error.
 */
int orc_pair(const uint8_t *oc_char, const uint32_t *oc_idx, uint64_t M, unsigned flags,
             uint64_t struct_base, int32_t *pair, uint32_t *keys_out, uint32_t *perm_out);

/* Emulations of the CUDA SIMD-within-a-word intrinsics the reference uses. */
uint32_t orc_vcmpeq4(uint32_t a, uint32_t b);
uint32_t orc_vcmpltu4(uint32_t a, uint32_t b);
uint32_t orc_vcmpgeu4(uint32_t a, uint32_t b);
uint32_t orc_vcmpgtu4(uint32_t a, uint32_t b);
uint32_t orc_vsubus4(uint32_t a, uint32_t b);
uint32_t orc_vsubss4(uint32_t a, uint32_t b);

#ifdef __cplusplus
}
#endif
#endif
