// cjx_records.hip - per-record table of a JSON Lines parse (CJX_FLAG_RECORDS).
//
// The reference has no record table: a record is implicit as the structural
// entries up to a '\n' entry (parse_json_lines.cu:453-456), and only chunk
// sizes are reported (resultSizes / resultSizesPrefix, :1218-1219).  Here a
// record is the run of structural rows of one chunk that ends with a '\n'
// row, plus, when the chunk's last rows follow its last '\n', one
// unterminated record.  For record r:
//   byte_begin  absolute offset of its first byte: the chunk start, or the
//               byte after the previous '\n' (that row's 1-based position)
//   byte_end    the byte after its '\n' (exclusive), or the chunk end
//   struct_first / struct_last  its first and last rows as host-layout
//               structural indices (1-based, the '\n' row included)
//   chunk, error  the chunk and that chunk's status code
// One block per chunk.  The tokenizer leaves a flag per structural row
// (row_nl, 1 for a '\n' row), so no pass re-reads the rows or the input:
// k_rec_count counts a chunk's flags, k_rec_scan numbers the records, and
// k_rec_write finds the '\n' rows in order with block scans and reads the
// structural values of those rows only.
#include "cjx_device.h"

namespace cjx {

namespace {

constexpr int kRecThreads = 256;

__device__ __forceinline__ uint32_t block_sum_u32(uint32_t x, uint32_t *lds) {
    for (int d = 32; d >= 1; d >>= 1) x += (uint32_t)__shfl_xor((int)x, d, 64);
    __syncthreads();
    if (lane_id() == 0) lds[threadIdx.x >> 6] = x;
    __syncthreads();
    uint32_t r = 0;
    for (int k = 0; k < kRecThreads / 64; k++) r += lds[k];
    return r;
}

// Sum of the four bytes of a word (each 0 or 1).
__device__ __forceinline__ uint32_t byte_sum(uint32_t v) {
    return (uint32_t)__popc(v & 0x01010101u);
}

}  // namespace

__global__ __launch_bounds__(kRecThreads) void k_rec_count(const UnitOut *__restrict__ uo,
                                                           const uint8_t *__restrict__ row_nl,
                                                           uint32_t *__restrict__ rec_count,
                                                           const LaunchState *__restrict__ ls) {
    __shared__ uint32_t red[kRecThreads / 64];
    if (ls->overflow) return;
    const uint32_t u = blockIdx.x;
    const uint32_t nb = uo[u].n_base, ne = uo[u].n_end;
    uint32_t c = 0;
    // head bytes up to a 16-byte boundary, then 16 flags per load
    const uint32_t a = min(ne, (nb + 15u) & ~15u);
    for (uint32_t t = nb + threadIdx.x; t < a; t += kRecThreads) c += row_nl[t];
    const uint32_t b = a + ((ne - a) & ~15u);
    for (uint32_t t = a + 16u * threadIdx.x; t < b; t += 16u * kRecThreads) {
        const uint4 v = *reinterpret_cast<const uint4 *>(row_nl + t);
        c += byte_sum(v.x) + byte_sum(v.y) + byte_sum(v.z) + byte_sum(v.w);
    }
    for (uint32_t t = b + threadIdx.x; t < ne; t += kRecThreads) c += row_nl[t];
    c = block_sum_u32(c, red);
    if (threadIdx.x == 0) {
        // an unterminated tail: rows after the last '\n'
        const bool tail = ne > nb && !row_nl[ne - 1];
        rec_count[u] = c + (tail ? 1u : 0u);
    }
}

// One block: exclusive prefix of the record counts (in place) and the total.
__global__ __launch_bounds__(kRecThreads) void k_rec_scan(uint32_t *__restrict__ rec_count,
                                                          uint32_t n_units,
                                                          LaunchState *__restrict__ ls) {
    __shared__ uint32_t wsum[kRecThreads / 64];
    __shared__ uint32_t carry;
    if (ls->overflow) {  // k_rec_count wrote no counts: no records
        if (threadIdx.x == 0) ls->n_records = 0;
        return;
    }
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < n_units; base += kRecThreads) {
        const uint32_t u = base + threadIdx.x;
        const uint32_t v = u < n_units ? rec_count[u] : 0u;
        const uint32_t in = wave_incl_scan(v);
        if (lane_id() == 63) wsum[threadIdx.x >> 6] = in;
        __syncthreads();
        uint32_t pre = carry + in - v, tot = 0;
        for (uint32_t k = 0; k < kRecThreads / 64; k++) {
            if (k < (threadIdx.x >> 6)) pre += wsum[k];
            tot += wsum[k];
        }
        if (u < n_units) rec_count[u] = pre;
        __syncthreads();
        if (threadIdx.x == 0) carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) ls->n_records = carry;
}

// Thread t of a block owns 16 consecutive rows per step (one 16-byte load of
// flags, packed to 16 bits); block scans give each '\n' row its record index
// and the last '\n' row before it.
__global__ __launch_bounds__(kRecThreads) void k_rec_write(const UnitDesc *__restrict__ units,
                                                           const UnitOut *__restrict__ uo,
                                                           const uint32_t *__restrict__ st,
                                                           const uint8_t *__restrict__ row_nl,
                                                           const uint32_t *__restrict__ rec_base,
                                                           RecordOut *__restrict__ rec,
                                                           uint64_t max_records,
                                                           LaunchState *__restrict__ ls) {
    constexpr uint32_t kRows = 16;
    __shared__ uint32_t wcnt[kRecThreads / 64];
    __shared__ int32_t wmax[kRecThreads / 64];
    __shared__ uint32_t s_count;
    __shared__ int32_t s_last;
    if (ls->overflow) return;
    const uint32_t u = blockIdx.x;
    const UnitDesc U = units[u];
    const uint32_t nb = uo[u].n_base, ne = uo[u].n_end;
    const uint32_t hb = uo[u].row_base + 1u - nb;  // region row -> host-layout index
    const int32_t code = uo[u].utf8_err ? 2 : (uo[u].struct_err ? 3 : 0);
    const uint32_t w = threadIdx.x >> 6, lane = lane_id();
    const uint32_t k = rec_base[u];
    if (threadIdx.x == 0) {
        s_count = 0;
        s_last = (int32_t)nb - 1;  // the row before the chunk's first row
    }
    __syncthreads();
    // Steps start at a 16-row boundary; flags outside [nb, ne) are masked off
    // (the flag buffer has 16 bytes of slack past the last row).
    for (uint32_t base = nb & ~(kRows - 1); base < ne; base += kRows * kRecThreads) {
        const uint32_t t0 = base + kRows * threadIdx.x;
        uint32_t bits = 0;
        if (t0 < ne) {
            const uint4 v = *reinterpret_cast<const uint4 *>(row_nl + t0);
            // bytes -> one bit each: bit 0 of every byte only (rows outside the
            // unit hold other data), then the four products land in distinct bits
            const uint32_t m = 0x01010101u, f = 0x01020408u;
            bits = ((((v.x & m) * f) >> 24) & 0xFu) | (((((v.y & m) * f) >> 24) & 0xFu) << 4) |
                   (((((v.z & m) * f) >> 24) & 0xFu) << 8) |
                   (((((v.w & m) * f) >> 24) & 0xFu) << 12);
            const uint32_t lo = t0 < nb ? nb - t0 : 0u, hi = min(kRows, ne - t0);
            bits &= ((1u << hi) - 1u) & ~((1u << lo) - 1u);
        }
        const uint32_t cnt = (uint32_t)__popc(bits);
        const int32_t last = bits ? (int32_t)(t0 + 31u - (uint32_t)__clz(bits)) : -1;
        const uint32_t ic = wave_incl_scan(cnt);
        int32_t mx = last;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int32_t o = __shfl_up(mx, off, 64);
            if (lane >= (uint32_t)off) mx = max(mx, o);
        }
        if (lane == 63) {
            wcnt[w] = ic;
            wmax[w] = mx;
        }
        __syncthreads();
        uint32_t before = s_count + ic - cnt;
        int32_t prev = s_last;
        for (uint32_t j = 0; j < w; j++) {
            before += wcnt[j];
            prev = max(prev, wmax[j]);
        }
        const int32_t up = __shfl_up(mx, 1, 64);
        if (lane > 0) prev = max(prev, up);
        while (bits) {
            const uint32_t r = (uint32_t)__builtin_ctz(bits);
            bits &= bits - 1u;
            const uint32_t t = t0 + r;
            RecordOut o;
            o.byte_begin = prev >= (int32_t)nb ? (int64_t)st[prev] : (int64_t)U.pos_base;
            o.byte_end = (int64_t)st[t];
            o.struct_first = (int32_t)((uint32_t)(prev + 1) + hb);
            o.struct_last = (int32_t)(t + hb);
            o.chunk = (int32_t)u;
            o.error = code;
            if (k + before < max_records) rec[k + before] = o;
            else atomicOr(&ls->rec_overflow, 1u);
            prev = (int32_t)t;
            before++;
        }
        __syncthreads();
        if (threadIdx.x == kRecThreads - 1) {
            uint32_t c = s_count;
            int32_t l = s_last;
            for (uint32_t j = 0; j < kRecThreads / 64; j++) {
                c += wcnt[j];
                l = max(l, wmax[j]);
            }
            s_count = c;
            s_last = l;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && ne > nb && s_last < (int32_t)ne - 1) {  // unterminated tail
        RecordOut r;
        r.byte_begin = s_last >= (int32_t)nb ? (int64_t)st[s_last] : (int64_t)U.pos_base;
        r.byte_end = (int64_t)(U.pos_base + U.n);
        r.struct_first = (int32_t)((uint32_t)(s_last + 1) + hb);
        r.struct_last = (int32_t)((ne - 1) + hb);
        r.chunk = (int32_t)u;
        r.error = code;
        if (k + s_count < max_records) rec[k + s_count] = r;
        else atomicOr(&ls->rec_overflow, 1u);
    }
}

hipError_t launch_records(hipStream_t s, const LaunchBufs &b, uint32_t n_units,
                          const uint32_t *d_struct, uint32_t *rec_count, RecordOut *rec,
                          uint64_t max_records, bool write) {
    if (!b.row_nl) return hipErrorInvalidValue;
    if (!write) {
        hipLaunchKernelGGL(k_rec_count, dim3(n_units), dim3(kRecThreads), 0, s, b.uout, b.row_nl,
                           rec_count, b.ls);
        hipLaunchKernelGGL(k_rec_scan, dim3(1), dim3(kRecThreads), 0, s, rec_count, n_units,
                           b.ls);
    } else {
        hipLaunchKernelGGL(k_rec_write, dim3(n_units), dim3(kRecThreads), 0, s, b.units, b.uout,
                           d_struct, b.row_nl, rec_count, rec, max_records, b.ls);
    }
    return hipGetLastError();
}

}  // namespace cjx
