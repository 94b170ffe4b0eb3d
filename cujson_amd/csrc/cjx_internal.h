// cjx_internal.h - device-side data structures shared by the kernels and the
// host driver of the MI355X JSON parse.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cjx {

// Tokenizer tile geometry: a tile is one 256-thread workgroup (4 wave64s);
// each lane owns 4 coalesced 16-byte segments, one per "round", so a wave
// covers 4 KiB and a tile 16 KiB of input.
constexpr int kTokThreads = 256;
constexpr int kWaves = kTokThreads / 64;
constexpr int kRounds = 4;
constexpr int kWaveBytes = 64 * 16 * kRounds;       // 4096
constexpr int kTileBytes = kTokThreads * 16 * kRounds;  // 16384

// Phase-3 geometry.
constexpr int kP3Threads = 256;
constexpr int kPairChunk = 4096;  // structural indices per pair-writer chunk (LDS staging)

// One parse unit inside a launch: a standard document, a JSON Lines chunk or
// one document of a batch.  Units are concatenated: structural and bracket
// indices run on across units (JSON Lines lastStructuralIndex semantics).
struct UnitDesc {
    uint64_t in_off;    // byte offset of the unit in the device input (16-byte aligned)
    uint64_t n;         // unit length in bytes
    uint64_t pos_base;  // added to the 1-based byte positions (lastChunkIndex)
    uint32_t tile0;     // first tokenizer tile of the unit
    uint32_t ntiles;    // tiles of the unit
};

// Per-unit results written on the device.
struct UnitOut {
    uint32_t n_base, m_base;  // first structural / bracket index of the unit
    uint32_t n_end, m_end;    // one past the last
    uint32_t utf8_err;        // phase-1 flag
    uint32_t struct_err;      // phase-3 flag
    uint32_t pad0, pad1;
};

// Per-launch counters, cleared by k_reset before every parse.
struct LaunchState {
    uint32_t tile_counter;  // dynamic tile ticket
    uint32_t overflow;      // an output exceeded its capacity
    uint32_t n_total, m_total;
};

// Look-back status encoding.
constexpr uint32_t kStInvalid = 0, kStAgg = 1, kStInc = 2;

}  // namespace cjx

namespace cjx {

// Device buffers one launch uses (owned by the context).
struct LaunchBufs {
    UnitDesc *units;
    UnitOut *uout;
    LaunchState *ls;
    uint32_t *carry_st;
    unsigned long long *count_st;
    uint8_t *oc_char;
    uint32_t *oc_idx;
    uint16_t *keys;
    uint32_t *sorted;
    uint32_t *tmp;
    uint32_t *spos;
    int32_t *block_delta;
    uint32_t *hist;
};

hipError_t launch_reset(hipStream_t s, const LaunchBufs &b, uint32_t ntiles, uint32_t n_units);
hipError_t launch_tokenize(hipStream_t s, const LaunchBufs &b, const uint8_t *d_in,
                           uint32_t ntiles, uint32_t n_units, bool jsonl, uint32_t *out_struct,
                           uint32_t n_cap, uint32_t m_cap);
hipError_t launch_phase3(hipStream_t s, const LaunchBufs &b, uint32_t u, uint32_t G,
                         uint32_t pair_grid, bool depth_u16, int32_t *pair_row, bool local_pairs);

}  // namespace cjx
