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
// One block per chunk: k_rec_count counts '\n' rows (the byte at a row's
// position is read from the input), k_rec_scan numbers the records, and
// k_rec_write fills the table with block scans over the rows in order.
#include "cjx_device.h"

namespace cjx {

namespace {

constexpr int kRecThreads = 256;

__device__ __forceinline__ bool row_is_newline(const uint8_t *__restrict__ in,
                                               const UnitDesc &U, uint32_t value) {
    return in[U.in_off + (value - 1u - (uint32_t)U.pos_base)] == '\n';
}

__device__ __forceinline__ uint32_t block_sum_u32(uint32_t x, uint32_t *lds) {
    for (int d = 32; d >= 1; d >>= 1) x += (uint32_t)__shfl_xor((int)x, d, 64);
    __syncthreads();
    if (lane_id() == 0) lds[threadIdx.x >> 6] = x;
    __syncthreads();
    uint32_t r = 0;
    for (int k = 0; k < kRecThreads / 64; k++) r += lds[k];
    return r;
}

}  // namespace

__global__ __launch_bounds__(kRecThreads) void k_rec_count(const uint8_t *__restrict__ in,
                                                           const UnitDesc *__restrict__ units,
                                                           const UnitOut *__restrict__ uo,
                                                           const uint32_t *__restrict__ st,
                                                           uint32_t *__restrict__ rec_count,
                                                           const LaunchState *__restrict__ ls) {
    __shared__ uint32_t red[kRecThreads / 64];
    if (ls->overflow) return;
    const uint32_t u = blockIdx.x;
    const UnitDesc U = units[u];
    const uint32_t nb = uo[u].n_base, ne = uo[u].n_end;
    uint32_t c = 0;
    for (uint32_t t = nb + threadIdx.x; t < ne; t += kRecThreads) c += row_is_newline(in, U, st[t]);
    c = block_sum_u32(c, red);
    if (threadIdx.x == 0) {
        // an unterminated tail: rows after the last '\n'
        const bool tail = ne > nb && !row_is_newline(in, U, st[ne - 1]);
        rec_count[u] = c + (tail ? 1u : 0u);
    }
}

// One block: exclusive prefix of the record counts (in place) and the total.
__global__ __launch_bounds__(kRecThreads) void k_rec_scan(uint32_t *__restrict__ rec_count,
                                                          uint32_t n_units,
                                                          LaunchState *__restrict__ ls) {
    __shared__ uint32_t wsum[kRecThreads / 64];
    __shared__ uint32_t carry;
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

__global__ __launch_bounds__(kRecThreads) void k_rec_write(const uint8_t *__restrict__ in,
                                                           const UnitDesc *__restrict__ units,
                                                           const UnitOut *__restrict__ uo,
                                                           const uint32_t *__restrict__ st,
                                                           const uint32_t *__restrict__ rec_base,
                                                           RecordOut *__restrict__ rec,
                                                           uint64_t max_records,
                                                           LaunchState *__restrict__ ls) {
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
    uint32_t k = rec_base[u];
    if (threadIdx.x == 0) {
        s_count = 0;
        s_last = (int32_t)nb - 1;  // the row before the chunk's first row
    }
    __syncthreads();
    for (uint32_t base = nb; base < ne; base += kRecThreads) {
        const uint32_t t = base + threadIdx.x;
        const bool valid = t < ne;
        const uint32_t val = valid ? st[t] : 0u;
        const bool nl = valid && row_is_newline(in, U, val);
        // records before this row in the tile, and the last '\n' row before it
        const uint32_t ic = wave_incl_scan(nl ? 1u : 0u);
        int32_t mx = nl ? (int32_t)t : -1;
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
        uint32_t before = s_count + ic - (nl ? 1u : 0u);
        int32_t prev = s_last;
        for (uint32_t j = 0; j < w; j++) {
            before += wcnt[j];
            prev = max(prev, wmax[j]);
        }
        const int32_t up = __shfl_up(mx, 1, 64);
        if (lane > 0) prev = max(prev, up);
        if (nl) {
            RecordOut r;
            r.byte_begin = prev >= (int32_t)nb ? (int64_t)st[prev] : (int64_t)U.pos_base;
            r.byte_end = (int64_t)val;
            r.struct_first = (int32_t)((uint32_t)(prev + 1) + hb);
            r.struct_last = (int32_t)(t + hb);
            r.chunk = (int32_t)u;
            r.error = code;
            if (k + before < max_records) rec[k + before] = r;
            else atomicOr(&ls->rec_overflow, 1u);
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

hipError_t launch_records(hipStream_t s, const LaunchBufs &b, const uint8_t *d_in,
                          uint32_t n_units, const uint32_t *d_struct, uint32_t *rec_count,
                          RecordOut *rec, uint64_t max_records, bool write) {
    if (!write) {
        hipLaunchKernelGGL(k_rec_count, dim3(n_units), dim3(kRecThreads), 0, s, d_in, b.units,
                           b.uout, d_struct, rec_count, b.ls);
        hipLaunchKernelGGL(k_rec_scan, dim3(1), dim3(kRecThreads), 0, s, rec_count, n_units,
                           b.ls);
    } else {
        hipLaunchKernelGGL(k_rec_write, dim3(n_units), dim3(kRecThreads), 0, s, d_in, b.units,
                           b.uout, d_struct, rec_count, rec, max_records, b.ls);
    }
    return hipGetLastError();
}

}  // namespace cjx
