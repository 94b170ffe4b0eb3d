// cjx_noalloc.hip - cjx_parse_device_noalloc: the device-resident parse of one
// standard-JSON document (cjx_parse_device's kernels) with every working
// buffer carved out of one file-scope device array, so that a call allocates
// nothing: no hipMalloc / hipFree, no pinned upload buffer, no events.  For
// callers that cannot allocate (code under HIP graph capture, tight serving
// loops, harnesses that forbid allocation).
//
// The array is `scratch` below.  Its host-side name is only a stand-in for the
// device symbol, so the host never passes it to a kernel: the first call on a
// device (with that device current) resolves the device address with
// hipGetSymbolAddress and keeps it per device; every later use goes through
// that address.  Each device holds its own copy of the array for the life of
// the process.  Every call on a device shares that device's array: calls that
// overlap on one device (on different streams, or from different threads)
// race and must be serialized by the caller.  The setup kernel clears what
// must start at zero (look-back status, unit results, launch counters) and
// writes the one-unit table, so nothing left by an earlier call is read.
#include <mutex>

#include "../../include/cujson_amd.h"
#include "cjx_device.h"

namespace cjx {
namespace {

// Largest input of one call; every buffer below is sized for it.
constexpr uint64_t kNaMaxBytes = 1ull << 20;
constexpr uint64_t kNaRows = kNaMaxBytes;  // structural rows never exceed the input bytes
constexpr uint64_t kNaTiles = (kNaMaxBytes + kTileBytes - 1) / kTileBytes;
// matching chunks and chunk groups of one unit of kNaRows rows, as the
// context path bounds them (rows / kMatchChunk rounded up, plus one)
constexpr uint64_t kNaChunks = (kNaRows + kMatchChunk - 1) / kMatchChunk + 1;
constexpr uint64_t kNaGroups = kNaChunks / 64 + 1;

constexpr uint64_t align256(uint64_t x) { return (x + 255) & ~255ull; }

struct NaLayout {
    uint64_t units, tile_unit, uout, ls, tile_st, oc_char, oc_idx, keys, sorted, tmp, spos,
        hist, ck, gmin, gmin1, gmin2, gagg, gd0, lows, highs, chunk_unit,
        big_units, slow_units, unit_part, end;
};

constexpr NaLayout make_layout() {
    NaLayout L{};
    uint64_t o = 0;
    auto take = [&o](uint64_t bytes) {
        const uint64_t at = o;
        o = align256(o + bytes);
        return at;
    };
    L.units = take(sizeof(UnitDesc));
    L.tile_unit = take((kNaTiles + 1) * 4);  // the tile -> unit map, then the small-unit list
    L.uout = take(sizeof(UnitOut));
    L.ls = take(sizeof(LaunchState));
    L.tile_st = take(kNaTiles * 16);
    L.oc_char = take(kNaRows + 16);
    L.oc_idx = take(kNaRows * 4 + 16);
    L.keys = take(kNaRows * 2 + 16);
    L.sorted = take(kNaRows * 4 + 16);
    L.tmp = take(kNaRows * 4 + 16);
    L.spos = take(kNaRows * 4 + 16);
    L.hist = take(512 * 256 * 4);
    L.ck = take(kNaChunks * sizeof(ChunkStat) + 16);
    L.gmin = take(kNaChunks * 4 + 16);
    L.gmin1 = take(kNaGroups * 4 + 16);
    L.gmin2 = take(kNaGroups * 4 + 16);
    L.gagg = take(kNaGroups * sizeof(DepthAgg) + 16);
    L.gd0 = take(kNaGroups * 4 + 16);
    L.lows = take(kNaChunks * kMatchKeys * 4 + 16);
    L.highs = take(kNaChunks * kMatchKeys * 4 + 16);
    L.chunk_unit = take(kNaChunks * 4 + 16);
    L.big_units = take(4 + 16);
    L.slow_units = take(4 + 16);  // one unit
    L.unit_part = take(2 * 32);
    L.end = o;
    return L;
}
constexpr NaLayout kNa = make_layout();

}  // namespace

// About 20.1 MiB per device.
static __device__ __align__(256) uint8_t scratch[kNa.end];

namespace {

// The one-unit table and the state that must start at zero.  One block.
__global__ __launch_bounds__(256) void k_na_setup(UnitDesc *units, uint32_t *tile_unit,
                                                  UnitOut *uo, LaunchState *ls,
                                                  unsigned long long *tile_st, uint64_t n,
                                                  uint32_t ntiles, uint32_t out_cap) {
    const uint32_t t = threadIdx.x;
    if (t == 0) {
        UnitDesc u;
        u.in_off = 0;
        u.n = n;
        u.pos_base = 0;
        u.tile0 = 0;
        u.ntiles = ntiles;
        u.out_off = 0;
        u.out_cap = out_cap;
        units[0] = u;
        *uo = UnitOut{};
        *ls = LaunchState{};
    }
    // tile -> unit map (all unit 0), then the small-unit list {0}
    for (uint32_t i = t; i <= ntiles; i += blockDim.x) tile_unit[i] = 0;
    for (uint32_t i = t; i < 2 * ntiles; i += blockDim.x) tile_st[i] = 0;
}

// Status of the parse into the caller's cjx_device_status (what
// cjx_device_wait reports for cjx_parse_device).
__global__ void k_na_status(const UnitOut *uo, const LaunchState *ls, cjx_device_status *st) {
    if (threadIdx.x != 0) return;
    int32_t code = CJX_OK;
    if (uo->utf8_err) code = CJX_ERR_UTF8;
    else if (uo->struct_err) code = CJX_ERR_STRUCTURE;
    if (ls->overflow) code = CJX_ERR_CAPACITY;
    cjx_device_status s;
    s.n_structural = ls->n_total;
    s.n_brackets = ls->m_total;
    s.error = code;
    s.capacity_exceeded = ls->overflow ? 1 : 0;
    s.n_records = 0;
    *st = s;
}

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mu;
uint8_t *g_na_ptr[kNaMaxDevices];  // device address of `scratch` per device

// The device address of `scratch` on the current device, resolved on the
// first call there.
int na_scratch(uint8_t **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNaMaxDevices)
        return CJX_ERR_DEVICE;
    std::lock_guard<std::mutex> lock(g_na_mu);
    if (!g_na_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
            return CJX_ERR_DEVICE;
        g_na_ptr[dev] = static_cast<uint8_t *>(p);
    }
    *out = g_na_ptr[dev];
    return CJX_OK;
}

}  // namespace
}  // namespace cjx

extern "C" {

size_t cjx_noalloc_max_bytes(void) { return (size_t)cjx::kNaMaxBytes; }

int cjx_parse_device_noalloc(const uint8_t *d_input, size_t n, unsigned flags,
                             int32_t *d_structural, int32_t *d_pair, size_t cap,
                             cjx_device_status *d_status, void *stream) {
    using namespace cjx;
    if (!d_input || n == 0 || !d_structural || !d_pair || !d_status) return CJX_ERR_INPUT;
    if ((reinterpret_cast<uintptr_t>(d_input) & 15u) ||
        (flags & ~(CJX_FLAG_UTF8_STRICT | CJX_FLAG_DEPTH_U16)))
        return CJX_ERR_INPUT;
    if (n > kNaMaxBytes) return CJX_ERR_TOO_LARGE;
    uint8_t *base = nullptr;
    if (int rc = na_scratch(&base)) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const bool strict = (flags & CJX_FLAG_UTF8_STRICT) != 0;
    const bool u16 = (flags & CJX_FLAG_DEPTH_U16) != 0;

    // one unit: tiles unless it fits one wave (then the small-unit path)
    const uint32_t ntiles = n <= (uint64_t)kWaveBytes ? 0u
                                                      : (uint32_t)((n + kTileBytes - 1) / kTileBytes);
    const uint32_t n_small = ntiles ? 0u : 1u;
    const uint32_t rows = (uint32_t)(cap < n ? cap : n);
    const uint32_t chunks = (rows + kMatchChunk - 1) / kMatchChunk + 1;
    const uint32_t groups = chunks / 64 + 1;

    LaunchBufs b;
    b.units = reinterpret_cast<UnitDesc *>(base + kNa.units);
    b.tile_unit = reinterpret_cast<uint32_t *>(base + kNa.tile_unit);
    b.small_units = b.tile_unit + ntiles;
    b.chunk_unit = reinterpret_cast<uint32_t *>(base + kNa.chunk_unit);
    b.big_units = reinterpret_cast<uint32_t *>(base + kNa.big_units);
    b.slow_units = reinterpret_cast<uint32_t *>(base + kNa.slow_units);
    b.unit_part = reinterpret_cast<uint32_t *>(base + kNa.unit_part);
    b.uout = reinterpret_cast<UnitOut *>(base + kNa.uout);
    b.ls = reinterpret_cast<LaunchState *>(base + kNa.ls);
    b.tile_st = reinterpret_cast<unsigned long long *>(base + kNa.tile_st);
    b.oc_char = base + kNa.oc_char;
    b.oc_idx = reinterpret_cast<uint32_t *>(base + kNa.oc_idx);
    b.keys = reinterpret_cast<uint16_t *>(base + kNa.keys);
    b.sorted = reinterpret_cast<uint32_t *>(base + kNa.sorted);
    b.tmp = reinterpret_cast<uint32_t *>(base + kNa.tmp);
    b.spos = reinterpret_cast<uint32_t *>(base + kNa.spos);
    b.hist = reinterpret_cast<uint32_t *>(base + kNa.hist);
    b.ck = reinterpret_cast<ChunkStat *>(base + kNa.ck);
    b.gmin = reinterpret_cast<int32_t *>(base + kNa.gmin);
    b.gmin1 = reinterpret_cast<int32_t *>(base + kNa.gmin1);
    b.gmin2 = reinterpret_cast<int32_t *>(base + kNa.gmin2);
    b.gagg = reinterpret_cast<DepthAgg *>(base + kNa.gagg);
    b.gd0 = reinterpret_cast<int32_t *>(base + kNa.gd0);
    b.lows = reinterpret_cast<uint32_t *>(base + kNa.lows);
    b.highs = reinterpret_cast<uint32_t *>(base + kNa.highs);
    b.row_nl = nullptr;

    uint32_t *d_st = reinterpret_cast<uint32_t *>(d_structural);
    hipLaunchKernelGGL(k_na_setup, dim3(1), dim3(256), 0, s, b.units, b.tile_unit, b.uout, b.ls,
                       b.tile_st, (uint64_t)n, ntiles, rows);
    if (hipGetLastError() != hipSuccess) return CJX_ERR_DEVICE;
    if (strict && launch_utf8_strict(s, b, d_input, ntiles, n_small) != hipSuccess)
        return CJX_ERR_DEVICE;
    const uint32_t n_tiny = (n_small && n <= (uint64_t)kRoundBytes) ? 1u : 0u;
    if (launch_tokenize(s, b, d_input, ntiles, n_small, false, !strict, d_st, TokShape{},
                        n_tiny) != hipSuccess ||
        launch_unit_bases(s, b, 1, false) != hipSuccess ||
        launch_match(s, b, 1, chunks, groups, d_pair) != hipSuccess ||
        launch_phase3(s, b, 0, phase3_blocks(rows), pair_blocks(rows), u16, d_pair) != hipSuccess)
        return CJX_ERR_DEVICE;
    hipLaunchKernelGGL(k_na_status, dim3(1), dim3(64), 0, s, b.uout, b.ls, d_status);
    return hipGetLastError() == hipSuccess ? CJX_OK : CJX_ERR_DEVICE;
}

}  // extern "C"
