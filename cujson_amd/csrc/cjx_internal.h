// cjx_internal.h - device-side data structures shared by the kernels and the
// host driver of the MI355X JSON parse.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cjx {

// Tokenizer tile geometry: a tile is one workgroup of CJX_TOK_WAVES wave64s;
// each lane owns 32 contiguous bytes per round, so a wave covers 2 KiB per
// round and runs CJX_TOK_ROUNDS rounds.  4 x 8 (64 KiB tiles) measured
// fastest on a 2 GB document: four workgroups fit a CU (4 waves per SIMD),
// so one waiting in its look-back leaves three to compute (8 x 8: 2-5 %
// slower, 4 x 6 or 4 x 12: 7-14 % and 3 % slower, 2 x 8: 50 % slower).
#ifndef CJX_TOK_ROUNDS
#define CJX_TOK_ROUNDS 8
#endif
#ifndef CJX_TOK_WAVES
#define CJX_TOK_WAVES 4
#endif
constexpr int kTokThreads = 64 * CJX_TOK_WAVES;
constexpr int kWaves = kTokThreads / 64;
constexpr int kLaneBytes = 32;
constexpr int kTokRounds = CJX_TOK_ROUNDS;
constexpr int kRoundBytes = 64 * kLaneBytes;              // 2 KiB per wave round
constexpr int kWaveBytes = kRoundBytes * kTokRounds;
constexpr int kTileBytes = kWaveBytes * kWaves;
#ifndef CJX_STAGE_ROWS
#define CJX_STAGE_ROWS 512
#endif
constexpr int kStageRows = CJX_STAGE_ROWS;  // LDS rows per wave for coalesced structural stores

// Phase-3 geometry.
constexpr int kP3Threads = 256;
constexpr int kPairChunk = 4096;  // structural indices per pair-writer chunk (LDS staging)

// Bracket matching of well-nested units (cjx_match.hip): brackets are cut
// into chunks of kMatchChunk; a chunk keeps one slot per depth key for its
// unmatched openers and closers.
constexpr int kMatchChunk = 4096;
constexpr int kMatchThreads = 256;
constexpr int kMatchKeys = 256;
constexpr int kMatchMaxDepth = 256;  // deeper units take the sort path
constexpr int kSmallBrackets = 1024;  // units with at most this many brackets: one wave matches them

// Grid of the sort-path kernels of one unit (phase3_blocks: depth walk,
// keys, histogram, scatter over 64 Ki brackets per block) and of its pair
// writer (pair_blocks: kPairChunk rows per block), from the unit's row bound.
inline uint32_t phase3_blocks(uint64_t max_m) {
    const uint64_t g = (max_m + 65535) / 65536;
    return (uint32_t)(g < 1 ? 1 : (g > 512 ? 512 : g));
}
inline uint32_t pair_blocks(uint64_t max_n) {
    const uint64_t g = (max_n + kPairChunk - 1) / kPairChunk;
    return (uint32_t)(g < 1 ? 1 : (g > 2048 ? 2048 : g));
}

// One parse unit inside a launch: a standard document, a JSON Lines chunk or
// one document of a batch.  Escape/string state, depth and counts restart at
// every unit.  The unit's outputs (structural row, pair row and bracket
// arrays) live in its own region [out_off, out_off + out_cap) of the output
// buffers; out_cap >= n always suffices since M <= N <= n.
struct UnitDesc {
    uint64_t in_off;    // byte offset of the unit in the device input (16-byte aligned)
    uint64_t n;         // unit length in bytes
    uint64_t pos_base;  // added to the 1-based byte positions (lastChunkIndex)
    uint32_t tile0;     // first tokenizer tile of the unit
    uint32_t ntiles;    // tiles of the unit
    uint32_t out_off;   // first row index of the unit's output region
    uint32_t out_cap;   // rows available in the region
};

// Per-unit results written on the device.
struct UnitOut {
    uint32_t n_base, m_base;  // region start (row index) of the unit's rows
    uint32_t n_end, m_end;    // one past the unit's last structural / bracket
    uint32_t utf8_err;        // phase-1 flag
    uint32_t struct_err;      // phase-3 flag
    uint32_t pair_base;       // added to pair values (lastStructuralIndex), 0 if local
    uint32_t fast;            // well nested: paired by cjx_match.hip, not by the sort path
    uint32_t chunk0;          // first matching chunk of the unit
    uint32_t nchunks;         // matching chunks (at least one when the unit has structurals)
    uint32_t grp0;            // first 64-chunk group (unit-local groups)
    uint32_t sgrp0;           // first 4096-chunk group
    uint32_t row_base;        // structural count of the units before (packed rows)
    uint32_t pad[3];
};

// Depth walk of a run of brackets: net change, and the minimum and maximum
// of the running depth, the start (0) included.
struct DepthAgg {
    int32_t d, mn, mx;
};

// Per matching chunk: its depth walk and the depth where it starts.
struct ChunkStat {
    int32_t d, mn, mx;
    int32_t d0;
};

// Per-launch counters, cleared by k_reset before every parse.
struct LaunchState {
    uint32_t tile_counter;  // dynamic tile ticket
    uint32_t overflow;      // an output exceeded its capacity
    uint32_t n_total, m_total;
    uint32_t n_chunks;  // matching chunks of all units
    uint32_t n_big;     // units with more than 64 matching chunks (listed in big_units)
    uint32_t n_records;     // JSON Lines records (cjx_records.hip)
    uint32_t rec_overflow;  // more records than the caller's table holds
    uint32_t chunk_ticket;  // dynamic matching-chunk ticket (k_match_local)
    uint32_t n_slow;        // units that are not well nested (listed in slow_units)
    uint32_t hist_ticket;   // blocks done with their histogram (sort path, one unit)
};

// One JSON Lines record (cjx_records.hip); same layout as cjx_record.
struct RecordOut {
    int64_t byte_begin, byte_end;
    int32_t struct_first, struct_last;
    int32_t chunk, error;
};

// Look-back status encoding (two 64-bit words per tile, see cjx_tokenize.hip).
constexpr uint32_t kStInvalid = 0, kStAgg = 1, kStInc = 2;
// Word A of an inclusive status: 2 << 62 | exit state << 60 | the unit's
// structural count up to the tile's end (the row where the next tile starts).
constexpr unsigned long long kStCountMask = (1ull << 60) - 1;

}  // namespace cjx

namespace cjx {

// Device buffers one launch uses (owned by the context).
struct LaunchBufs {
    UnitDesc *units;
    uint32_t *tile_unit;   // unit of every tokenizer tile (host-built)
    uint32_t *small_units; // units of at most kWaveBytes bytes (host-built)
    uint32_t *chunk_unit;  // unit of every matching chunk (device-built)
    uint32_t *big_units;   // units with more than 64 matching chunks
    uint32_t *slow_units;  // units found not well nested, in no order (the sort path's list)
    uint32_t *unit_part;   // per 256 units: 4 block sums of the unit-base scan
    UnitOut *uout;
    LaunchState *ls;
    unsigned long long *tile_st;  // 2 words per tile
    uint8_t *oc_char;
    uint32_t *oc_idx;
    uint16_t *keys;
    uint32_t *sorted;
    uint32_t *tmp;
    uint32_t *spos;
    uint32_t *hist;
    ChunkStat *ck;    // per matching chunk
    int32_t *gmin;    // per chunk: lowest depth reached (d0 + mn)
    int32_t *gmin1;   // per 64 chunks of a unit
    int32_t *gmin2;   // per 4096 chunks of a unit
    DepthAgg *gagg;   // per 64 chunks: their depth walk
    int32_t *gd0;     // per 64 chunks: start depth
    uint32_t *lows;   // per chunk x key: first closer reaching that depth (bracket index)
    uint32_t *highs;  // per chunk x key: last opener left open
    uint8_t *row_nl;  // JSON Lines with a record table: 1 per structural row that is a '\n'
                      // (written by the tokenizer), else null
};

hipError_t launch_reset(hipStream_t s, const LaunchBufs &b, uint32_t ntiles, uint32_t n_units);
// launch_reset for one unit, whose table (descriptor, tile map, small-unit
// entry) the kernel writes from its argument.
hipError_t launch_reset_one(hipStream_t s, const LaunchBufs &b, const UnitDesc &u, uint32_t ntiles,
                            uint32_t n_small);
hipError_t launch_rows_out(hipStream_t s, const uint32_t *src, uint32_t *host_dst, uint64_t count);
// Launch shape of the tokenizer, chosen at run time from the built
// instantiations.  A 64 KiB tile is tile_waves waves of 64 KiB / tile_waves
// bytes each (2 x 32 KiB, 4 x 16 KiB, 8 x 8 KiB or 16 x 4 KiB: 1024, 512, 256
// or 128 bytes per lane); small units take one wave of 16 KiB each,
// small_waves of them per workgroup (1, 2, 4 or 8).
struct TokShape {
    uint32_t tile_waves = 4;
    uint32_t small_waves = 4;
};
inline bool tok_shape_valid(uint32_t tile_waves, uint32_t small_waves) {
    const bool t = tile_waves == 2 || tile_waves == 4 || tile_waves == 8 || tile_waves == 16;
    const bool sm = small_waves == 1 || small_waves == 2 || small_waves == 4 || small_waves == 8;
    return t && sm;
}
// Units of at most kWaveBytes bytes are "small": no tiles, one wave each
// (listed in small_units, the first n_tiny of them of at most kRoundBytes
// bytes); the others are cut into kTileBytes tiles.
hipError_t launch_tokenize(hipStream_t s, const LaunchBufs &b, const uint8_t *d_in,
                           uint32_t ntiles, uint32_t n_small, bool jsonl, bool weak_utf8,
                           uint32_t *out_struct, TokShape shape = TokShape{},
                           uint32_t n_tiny = 0);
// `count` more big-tile workgroups (tickets continue from the launches before
// on the same stream), no small units: the tokenizer of a launch split into
// parts that each wait for their input bytes.
hipError_t launch_tokenize_tiles(hipStream_t s, const LaunchBufs &b, const uint8_t *d_in,
                                 uint32_t count, bool jsonl, bool weak_utf8, uint32_t *out_struct,
                                 TokShape shape = TokShape{});
hipError_t launch_utf8_strict(hipStream_t s, const LaunchBufs &b, const uint8_t *d_in,
                              uint32_t ntiles, uint32_t n_small);
hipError_t launch_unit_bases(hipStream_t s, const LaunchBufs &b, uint32_t n_units, bool concat);
hipError_t launch_match(hipStream_t s, const LaunchBufs &b, uint32_t n_units, uint32_t chunk_bound,
                        uint32_t group_bound, int32_t *pair_row);
// launch_match in two parts: the chunk walks, the small units' wave matcher
// and the chunk scans (which decide which units are well nested), then the
// chunk matcher and the cross-chunk pairs.
hipError_t launch_match_front(hipStream_t s, const LaunchBufs &b, uint32_t n_units,
                              uint32_t chunk_bound, int32_t *pair_row);
hipError_t launch_match_local(hipStream_t s, const LaunchBufs &b, uint32_t chunk_bound,
                              int32_t *pair_row);
// Packs the units' row regions densely: structural row i of unit u goes to
// dst_struct[row_base_u + i], its pair cell to dst_pair[row_base_u + i]; with
// host_layout, dst_struct = dst + 1 and dst_pair = dst + N + 2 of one 2(N+2)
// buffer whose sentinels are written too.
hipError_t launch_assemble(hipStream_t s, const LaunchBufs &b, uint32_t n_units,
                           const uint32_t *d_struct, const int32_t *d_pair, int32_t *dst_struct,
                           int32_t *dst_pair, uint32_t N, bool host_layout);
// Per-document host layouts of a batch: document u's 2(n_u + 2) cells start
// at cell 2 row_base(u) + 4u of dst (2N + 4 n_units cells in all).
hipError_t launch_assemble_docs(hipStream_t s, const LaunchBufs &b, uint32_t n_units,
                                const uint32_t *d_struct, const int32_t *d_pair, int32_t *dst,
                                uint32_t N);
// Kernel copy of two byte ranges (src/dst 16-byte aligned; either side may be
// pinned host memory); n1 = 0 for one range.
hipError_t launch_copy(hipStream_t s, const void *src0, void *dst0, uint64_t n0, const void *src1,
                       void *dst1, uint64_t n1);
// Sort-path phase 3 for every unit that is not well nested (listed in
// slow_units by the matching kernels), one block per unit at a time
// (launches with several units).
hipError_t launch_sort_units(hipStream_t s, const LaunchBufs &b, uint32_t n_units, bool depth_u16,
                             int32_t *pair_row);
// Per-record table of JSON Lines units: write = false counts the records of
// each unit and numbers them (rec_count becomes the bases, ls->n_records the
// total); write = true fills rec.
// Newline scan of device text (cjx_lines.hip): segment counts, bases and
// status {newline count, last byte}; the newline offsets; line starts at
// given line indices.
uint32_t nl_segments(uint64_t n);
hipError_t launch_nl_count(hipStream_t s, const uint8_t *d_in, uint64_t n, uint32_t *seg_count,
                           uint64_t *seg_base, uint64_t *stat);
hipError_t launch_nl_write(hipStream_t s, const uint8_t *d_in, uint64_t n,
                           const uint64_t *seg_base, uint64_t *out, uint64_t cap);
hipError_t launch_place_units(hipStream_t s, const uint8_t *src, uint64_t src_limit,
                              const uint64_t *src_off, const UnitDesc *units, uint32_t n_units,
                              uint8_t *dst, uint64_t dst_bytes);
hipError_t launch_nl_gather(hipStream_t s, const uint64_t *nl, uint64_t n_nl, uint64_t n,
                            const uint64_t *line_idx, uint64_t count, uint64_t *out);
// One entry of the typed value table (the C ABI's cjx_value).
struct ValueOut {
    int64_t v;  // INT value, FLOAT bits, KEY / STRING unescaped length
    uint32_t row, text_off, text_len;
    uint8_t type, flags;
    uint16_t reserved;
};
hipError_t launch_values(hipStream_t s, const uint8_t *d_in, uint64_t n, const int32_t *st,
                         uint64_t N, ValueOut *out, uint8_t *strings, uint32_t *exact_list,
                         uint32_t *exact_count);
hipError_t launch_records(hipStream_t s, const LaunchBufs &b, uint32_t n_units,
                          const uint32_t *d_struct, uint32_t *rec_count, RecordOut *rec,
                          uint64_t max_records, bool write);
hipError_t launch_phase3(hipStream_t s, const LaunchBufs &b, uint32_t u, uint32_t G,
                         uint32_t pair_grid, bool depth_u16, int32_t *pair_row);

}  // namespace cjx
