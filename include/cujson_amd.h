/*
 * cujson_amd.h - C ABI of the MI355X-native three-phase JSON parse.
 *
 * This is the drop-in boundary for cuJSON's parse path.  Every entry point
 * replaces one reference interface (cited per declaration); the C++ headers
 * under include/cujson/ rebuild the reference's own C++ API
 * (cuJSONInput / cuJSONLinesInput / cuJSONResult, loadJSON,
 * parse_standard_json, parse_json_lines) on top of these functions, so the
 * reference's query iterator consumes the results unchanged.
 *
 * Output layout (identical to the reference, cujson/parse_standard_json.cu:1659-1680):
 *   one int32 buffer of 2(N+2) cells
 *   structural = buf;        structural[0] = 0, structural[1..N] = 1-based byte
 *                            positions of { } [ ] : , (and '\n' for JSON Lines)
 *                            outside strings, structural[N+1] = N+1
 *   pair_pos   = buf + N + 1 pair_pos[h] for an opener at structural index h is
 *                            the structural index of its closer; cells the
 *                            reference leaves unwritten hold -1.
 *
 * Threading: a cjx_context owns one HIP stream and its device buffers; use
 * one context per host thread.  Calls on one context are synchronous with
 * respect to the caller unless stated otherwise.
 */
#ifndef CUJSON_AMD_H
#define CUJSON_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  CJX_ERR_UTF8 / CJX_ERR_STRUCTURE are where the reference
 * calls exit(0) (parse_standard_json.cu:1626, :1569; parse_json_lines.cu:1172,
 * :1058); CJX_ERR_INPUT is where it returns an empty cuJSONResult
 * (parse_standard_json.cu:1597-1600; parse_json_lines.cu:1103-1150). */
enum {
    CJX_OK = 0,
    CJX_ERR_INPUT = 1,
    CJX_ERR_UTF8 = 2,
    CJX_ERR_STRUCTURE = 3,
    CJX_ERR_ALLOC = 4,
    CJX_ERR_DEVICE = 5,
    CJX_ERR_CAPACITY = 6,
    CJX_ERR_TOO_LARGE = 7 /* input above an entry point's stated maximum */
};

/* Flags. */
#define CJX_FLAG_UTF8_STRICT 0x1u /* RFC 3629 validation (off: the reference's check)  */
#define CJX_FLAG_DEPTH_U16   0x2u /* 16-bit depth keys (off: the reference's uint8)     */
#define CJX_FLAG_JSONL       0x4u /* '\n' is structural (set by cjx_parse_lines)        */
#define CJX_FLAG_RECORDS     0x10u /* cjx_parse_lines: also return the per-record table   */
#define CJX_FLAG_GRAPH       0x8u /* device-resident calls: replay the launch sequence
                                     as a HIP graph captured on the first call of the
                                     same shape (no counterpart in the reference)      */

typedef struct cjx_context cjx_context;

/* One JSON Lines record (CJX_FLAG_RECORDS).  The reference has no record
 * table; a record is implicit as the structural entries up to a '\n' entry
 * (parse_json_lines.cu:453-456).  Here a record is the run of structural
 * entries of one chunk that ends with a '\n' entry, plus an unterminated one
 * when entries follow the chunk's last '\n'. */
typedef struct cjx_record {
    int64_t byte_begin;   /* offset of its first byte in the input            */
    int64_t byte_end;     /* one past its '\n' (or the chunk end)             */
    int32_t struct_first; /* first structural index (host layout, 1-based)    */
    int32_t struct_last;  /* last one (its '\n' entry when terminated)        */
    int32_t chunk;        /* the chunk it belongs to                          */
    int32_t error;        /* that chunk's status code                         */
} cjx_record;

/* Host-side result; field meanings follow cuJSONResult (cujson/cujson_types.h:22-33). */
typedef struct cjx_result {
    int32_t *buf;                 /* pinned host buffer of 2*(n_structural+2) int32  */
    int32_t *structural;          /* = buf                                          */
    int32_t *pair_pos;            /* = buf + n_structural + 1                       */
    int64_t n_structural;         /* N                                              */
    int64_t n_brackets;           /* M                                              */
    int32_t chunk_count;          /* cuJSONResult::chunkCount                       */
    int32_t *result_sizes;        /* per chunk N   (cuJSONResult::resultSizes)       */
    int32_t *result_sizes_prefix; /* inclusive sum (cuJSONResult::resultSizesPrefix) */
    int32_t total_result_size;    /* N + 2                                          */
    int32_t file_size;            /* N + 2 (the reference's fileSize)               */
    int32_t error;                /* status of the parse                            */
    int32_t error_chunk;          /* first failing chunk, -1 if none                */
    double device_ms;             /* device time of the phases (hipEvent)           */
    int64_t n_records;            /* JSON Lines with CJX_FLAG_RECORDS: record count */
    cjx_record *records;          /* and the records (pinned host memory)           */
} cjx_result;

/* Pinned host input (cuJSONInput, cujson/cujson_types.h:9-12). */
typedef struct cjx_input {
    uint8_t *data;
    size_t size;
} cjx_input;

/* Create / destroy a context on a HIP device (the reference uses device 0
 * implicitly, parse_standard_json.cu:1617). */
int cjx_context_create(int device, cjx_context **out);
void cjx_context_destroy(cjx_context *ctx);

/* Replaces cuJSONInput loadJSON(const std::string&), cujson/load_file.h:9 /
 * load_file.cu:7-36: read a file into a pinned host buffer. */
int cjx_load(const char *path, cjx_input *out);
void cjx_free_input(cjx_input *in);

/* Replaces the line scan of loadJSONLines_chunkCount / _chunkSizeBytes /
 * _chunkSizeMegaBytes, cujson/load_file.h:10-12 / load_file.cu:71-103,
 * :137-176: split a JSON Lines buffer into line-aligned chunks.
 *   mode 0: chunk_param = chunk count (1 .. 2^32); mode 1: chunk_param = max
 *   bytes per chunk.
 * Writes up to cap (offset, size) pairs; *n_chunks receives the count. */
int cjx_chunk_lines(const uint8_t *data, size_t n, int mode, size_t chunk_param,
                    size_t *chunk_off, size_t *chunk_size, size_t cap, size_t *n_chunks);

/* Replaces cuJSONResult parse_standard_json(cuJSONInput),
 * cujson/parse_standard_json.h:9 / parse_standard_json.cu:1584-1687.
 * data is host memory; the result buffer is pinned host memory owned by the
 * caller (release with cjx_free_result). */
int cjx_parse_standard(cjx_context *ctx, const uint8_t *data, size_t n, unsigned flags,
                       cjx_result *out);

/* Replaces cuJSONResult parse_json_lines(cuJSONLinesInput),
 * cujson/parse_json_lines.h:9 / parse_json_lines.cu:1092-1264.
 * Chunks (chunk_off[i], chunk_size[i]) are consecutive line-aligned slices of
 * data; positions count from data[0] as in the reference (lastChunkIndex). */
int cjx_parse_lines(cjx_context *ctx, const uint8_t *data, size_t n, const size_t *chunk_off,
                    const size_t *chunk_size, size_t n_chunks, unsigned flags,
                    cjx_result *out);

/* Releases a result (the reference's cuJSONIterator::freeJson calls
 * cudaFreeHost(structural), query_iterator_standard_json.cpp:143-146). */
void cjx_free_result(cjx_result *r);

const char *cjx_strerror(int code);

/* ---------------------------------------------------------------------- */
/* Device-resident form (no counterpart in the reference, which always    */
/* copies through the host): the input is already in HBM and the output   */
/* rows stay in HBM.  Used by pipelines that keep data on the GPU and by  */
/* the throughput benchmark.                                              */
/* ---------------------------------------------------------------------- */

/* Counts and codes of a device-resident parse, read back after sync. */
typedef struct cjx_device_status {
    int64_t n_structural;
    int64_t n_brackets;
    int32_t error;
    int32_t capacity_exceeded;
    int64_t n_records; /* JSON Lines with CJX_FLAG_RECORDS */
} cjx_device_status;

/* Parse a device-resident document.
 *   d_input      device pointer, 16-byte aligned, readable up to n rounded up to 16
 *   d_structural device int32[cap]: structural row (1-based positions, no sentinels)
 *   d_pair       device int32[cap]: pair row (structural index of the closer, 1-based
 *                in the host layout sense, -1 elsewhere)
 *   cap          capacity of both rows; N > cap reports CJX_ERR_CAPACITY
 * Launches on the context's stream and returns without synchronizing;
 * cjx_device_wait() synchronizes and fills the status. */
int cjx_parse_device(cjx_context *ctx, const uint8_t *d_input, size_t n, unsigned flags,
                     int32_t *d_structural, int32_t *d_pair, size_t cap);
int cjx_device_wait(cjx_context *ctx, cjx_device_status *st);

/* cjx_parse_device without allocation (no counterpart in the reference).
 * For callers that cannot allocate: code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  Same kernels, inputs,
 * outputs and results as cjx_parse_device, but no context: every working
 * buffer lives in one device array of the library (about 20 MiB per device,
 * held until the process ends), so a call makes no hipMalloc / hipFree, no
 * pinned or host allocation and creates no event.
 *   n            at most cjx_noalloc_max_bytes() (1 MiB); larger inputs return
 *                CJX_ERR_TOO_LARGE before anything is launched
 *   flags        CJX_FLAG_UTF8_STRICT, CJX_FLAG_DEPTH_U16 (others: CJX_ERR_INPUT)
 *   d_status     device-accessible cjx_device_status, written by the last kernel
 *                of the call (what cjx_device_wait reports for cjx_parse_device)
 *   stream       hipStream_t to launch on (NULL: the null stream)
 * Returns after enqueueing; the caller synchronizes the stream before
 * reading d_status or the rows.  Calls on one device share that device's
 * array (the last call's contents stay in it): calls that overlap on one
 * device, from any streams or threads, race and must be serialized by the
 * caller.  The first call on a device resolves the array's device address
 * (with that device current). */
size_t cjx_noalloc_max_bytes(void);
int cjx_parse_device_noalloc(const uint8_t *d_input, size_t n, unsigned flags,
                             int32_t *d_structural, int32_t *d_pair, size_t cap,
                             cjx_device_status *d_status, void *stream);

/* Per-phase device times of recent parses on a context, from hipEvents
 * recorded on the context stream around each phase while timing is enabled
 * (off by default; enabling clears the history).  cjx_context_timings fills
 * up to `max` entries for the most recent parses (at most 256), oldest
 * first, waiting for them to finish, and returns the count (-1 on error).
 * No counterpart in the reference, whose per-phase timers live only in its
 * paper drivers (paper_reproduced/src/cuJSON-standardjson.cu:2134-2249). */
typedef struct cjx_timings {
    float tokenize_ms; /* phases 1+2: the fused tokenizer kernel            */
    float phase3_ms;   /* phase 3: depth scan, sort, pairing (all units)     */
    float total_ms;    /* from the reset kernel to the last phase-3 kernel   */
} cjx_timings;
int cjx_context_set_timing(cjx_context *ctx, int enable);
int cjx_context_timings(cjx_context *ctx, cjx_timings *out, int max);

/* Tokenizer launch shape of the context's later parses, chosen at run time
 * from the library's built instantiations (the autotuner's knob).  A 64 KiB
 * tile is one workgroup of tile_waves wave64s (2, 4, 8 or 16: 1024, 512, 256
 * or 128 input bytes per lane); documents of at most 16 KiB take one wave
 * each, small_waves per workgroup (1, 2, 4 or 8).  Default 4 and 4.  Results
 * do not depend on the shape.  CJX_ERR_INPUT for other values.  No
 * counterpart in the reference, whose block size and per-thread work are
 * compile-time constants (BLOCKSIZE, cujson/utils.cu:41; WORDS,
 * cujson/parse_standard_json.cu:273, :1071, :1524). */
int cjx_context_set_tokenizer_shape(cjx_context *ctx, int tile_waves, int small_waves);
int cjx_context_tokenizer_shape(cjx_context *ctx, int *tile_waves, int *small_waves);

/* HIP stream of the context (hipStream_t), for callers that order their own
 * work against the parse. */
void *cjx_context_stream(cjx_context *ctx);

/* Library build identification. */
const char *cjx_version(void);

/* ---------------------------------------------------------------------- */
/* Batches of independent standard-JSON documents.  The reference parses  */
/* one document per parse_standard_json call (main.cu:26); a batch runs   */
/* all of them in one launch sequence.  Each document is its own parse    */
/* unit: positions are 1-based within the document and pair values are    */
/* host indices of the document's own layout, so rows [struct_begin,      */
/* struct_end) with the sentinels 0 / N_d+1 around them form exactly the  */
/* cuJSONResult that parse_standard_json returns for that document.       */
/* ---------------------------------------------------------------------- */

typedef struct cjx_doc_result {
    int64_t byte_base;    /* offset of the document in the batch input          */
    int64_t struct_begin; /* first row of the document                          */
    int64_t struct_end;   /* one past its last row                              */
    int32_t error;        /* CJX_OK, CJX_ERR_UTF8 or CJX_ERR_STRUCTURE           */
    int32_t reserved;
} cjx_doc_result;

typedef struct cjx_batch_result {
    int32_t *structural;  /* pinned host: the rows of all documents, packed     */
    int32_t *pair_pos;    /* pinned host: the pair cells of the same rows       */
    int64_t n_structural; /* total rows                                         */
    int64_t n_docs;
    cjx_doc_result *docs; /* n_docs entries                                     */
    int32_t error;        /* first document error, or the call's failure code   */
    int32_t reserved;
    double device_ms;
} cjx_batch_result;

/* Host documents (data + doc_off[i], doc_size[i] bytes each).  Per-document
 * codes are in out->docs; the return value is CJX_OK unless the call as a
 * whole failed (bad arguments, allocation, device).  Release with
 * cjx_free_batch_result.  Replaces a loop of parse_standard_json calls
 * (cujson/parse_standard_json.cu:1584-1687) over the documents. */
int cjx_parse_batch(cjx_context *ctx, const uint8_t *data, const size_t *doc_off,
                    const size_t *doc_size, size_t n_docs, unsigned flags,
                    cjx_batch_result *out);
void cjx_free_batch_result(cjx_batch_result *r);

/* Device-resident batch (no counterpart in the reference): documents at
 * offsets of d_input (16-byte aligned: read in place; otherwise all
 * documents are first copied into an aligned layout in the context's input
 * buffer; d_input is readable up to each document's end rounded up to 16);
 * rows stay in d_structural / d_pair
 * (cap rows each), document i in its own region.  Returns after enqueueing;
 * cjx_batch_wait synchronizes and fills docs[0..n_docs) (struct_begin /
 * struct_end index the caller's row buffers).  If the documents' bytes sum
 * to at most cap, regions are exact and never overflow; otherwise they are
 * sized n/3 + 256 and an overflow reports CJX_ERR_CAPACITY. */
int cjx_parse_batch_device(cjx_context *ctx, const uint8_t *d_input, const size_t *doc_off,
                           const size_t *doc_size, size_t n_docs, unsigned flags,
                           int32_t *d_structural, int32_t *d_pair, size_t cap);
int cjx_batch_wait(cjx_context *ctx, cjx_doc_result *docs, size_t n_docs,
                   cjx_device_status *st);

/* Device-resident JSON Lines (no counterpart in the reference, whose
 * parse_json_lines copies each chunk in, parse_json_lines.cu:1160-1230):
 * line-aligned chunks at offsets of d_input (as for cjx_parse_batch_device:
 * aligned chunks are read in place, others copied into an aligned layout
 * first, so the chunks of cjx_chunk_lines_device can be passed as they are),
 * parsed in one launch sequence with the reference's carries (positions count from
 * the first chunk as if the chunks were contiguous, pair values are numbered
 * across chunks).  Rows stay in per-chunk regions (cjx_batch_wait gives each
 * chunk's range; row_base ordering follows chunk order).  With
 * CJX_FLAG_RECORDS and d_records, the per-record table (host-layout
 * structural indices) is written to d_records[0..max_records). */
int cjx_parse_lines_device(cjx_context *ctx, const uint8_t *d_input, const size_t *chunk_off,
                           const size_t *chunk_size, size_t n_chunks, unsigned flags,
                           int32_t *d_structural, int32_t *d_pair, size_t cap,
                           cjx_record *d_records, size_t max_records);

/* Replaces the host line scan of loadJSONLines_* (cujson/load_file.cu:72-81)
 * for text already in device memory: the byte offset of every '\n' of
 * d_input[0..n) (16-byte aligned), in order, i.e. the record ends, written
 * to d_newlines[0..cap).  *n_newlines receives the count; a count above cap
 * returns CJX_ERR_CAPACITY (nothing written).  Synchronous. */
int cjx_scan_lines_device(cjx_context *ctx, const uint8_t *d_input, size_t n,
                          uint64_t *d_newlines, size_t cap, size_t *n_newlines);

/* Replaces the chunking of loadJSONLines_chunkCount / _chunkSizeBytes
 * (cujson/load_file.cu:83-103, :153-176) for text in device memory: the
 * chunks cjx_chunk_lines gives for the same bytes (same modes, same results,
 * host arrays), found from the device newline scan.  Mode 0 gathers the
 * chunk starts on the device; mode 1 (greedy by size) walks the line starts
 * on the host after one copy of the newline offsets.  Synchronous. */
int cjx_chunk_lines_device(cjx_context *ctx, const uint8_t *d_input, size_t n, int mode,
                           size_t chunk_param, size_t *chunk_off, size_t *chunk_size, size_t cap,
                           size_t *n_chunks);

/* Device counterpart of loadJSON / loadJSONLines' file read
 * (cujson/load_file.cu:7-36, :38-70): reads the file at path into
 * d_buf[0..cap) through pinned staging buffers, reading each part from the
 * file while earlier parts are being copied.  *n receives the file size; a
 * file larger than cap returns CJX_ERR_CAPACITY (nothing read).  Work queued
 * earlier on the context stream finishes before the buffer is written.
 * Synchronous. */
int cjx_load_device(cjx_context *ctx, const char *path, uint8_t *d_buf, size_t cap, size_t *n);

/* ------------------------------------------------------------------------
 * Typed value table (no counterpart in the reference, whose iterator returns
 * values as raw substrings: query_iterator_standard_json.cpp:475-567).
 * Entry h (h = 0 .. N) describes the text after structural h (host-layout
 * index; 0 = before the first) up to structural h + 1 (the end of the text
 * after the last), trimmed of JSON whitespace.
 *   KEY / STRING  text_off / text_len: the body between the quotes; the
 *                 unescaped UTF-8 (length v.i) is in strings[text_off ..)
 *                 when a strings buffer is given (an unpaired surrogate
 *                 escape is written as its 3-byte form, flag UNPAIRED)
 *   INT           v.i: an integer literal that fits int64
 *   FLOAT         v.d: any other number, binary64 correctly rounded (round
 *                 half to even); BIG_INT marks an integer literal outside
 *                 int64; APPROX one of more than 800 significant digits
 *                 whose rounding was not decided exactly (nearest candidate)
 *   TRUE / FALSE / NULL, NONE (empty), INVALID (no JSON scalar, e.g. 01,
 *                 NaN, a bad escape, a raw control character in a string)
 * ------------------------------------------------------------------------ */
enum {
    CJX_VALUE_NONE = 0,
    CJX_VALUE_KEY = 1,
    CJX_VALUE_STRING = 2,
    CJX_VALUE_INT = 3,
    CJX_VALUE_FLOAT = 4,
    CJX_VALUE_TRUE = 5,
    CJX_VALUE_FALSE = 6,
    CJX_VALUE_NULL = 7,
    CJX_VALUE_INVALID = 8
};
#define CJX_VALUE_ESCAPED  0x1u /* the string held escapes                      */
#define CJX_VALUE_BIG_INT  0x2u /* integer literal outside int64, given as v.d  */
#define CJX_VALUE_UNPAIRED 0x4u /* unpaired surrogate escape in the string      */
#define CJX_VALUE_APPROX   0x8u /* v.d not decided exactly (over 800 digits)    */

typedef struct cjx_value {
    union {
        int64_t i;
        double d;
    } v;
    uint32_t row;      /* h */
    uint32_t text_off; /* byte offset of the text (strings: of the body)    */
    uint32_t text_len; /* its length                                         */
    uint8_t type;      /* CJX_VALUE_*                                         */
    uint8_t flags;     /* CJX_VALUE_ESCAPED | BIG_INT | UNPAIRED | APPROX     */
    uint16_t reserved;
} cjx_value;

/* Device form: d_input[0..n) (n < 2^32) and the structural row of a parse of
 * it (d_structural[0..n_structural): 1-based positions, as cjx_parse_device
 * writes them); fills d_values[0..n_structural] and, if d_strings is not
 * null, the unescaped strings in d_strings[0..n).  *n_exact (optional)
 * receives how many numbers needed the exact comparison.  Synchronous. */
int cjx_values_device(cjx_context *ctx, const uint8_t *d_input, size_t n,
                      const int32_t *d_structural, size_t n_structural, cjx_value *d_values,
                      uint8_t *d_strings, size_t *n_exact);

/* Host form for a cjx_result of cjx_parse_standard / cjx_parse_lines over
 * data[0..n): fills values[0..N] and, if strings is not null, strings[0..n). */
int cjx_values(cjx_context *ctx, const uint8_t *data, size_t n, const cjx_result *r,
               cjx_value *values, uint8_t *strings);

/* ------------------------------------------------------------------------
 * Request server (no counterpart in the reference, whose parse calls are
 * synchronous one-document calls: parse_standard_json.cu:1590-1700).  For
 * serving callers with many concurrent requests: submit from any thread, a
 * worker thread batches what is pending onto n_slots streams (continuous
 * batching: a batch starts as soon as a slot is free), and completions come
 * back per request in the reference's host layout.
 *
 * device_budget bounds the estimated device bytes of the batches in flight
 * (each slot's batch stays within device_budget / n_slots; a larger single
 * request runs alone); slot buffers stay resident between batches.  A batch
 * holds at most max_batch requests and 64 MiB of input.  n_slots / 2 of the
 * slots are express slots that take only requests below 4 MiB, in batches
 * of at most 16 MiB, so small requests never queue behind the transfers of
 * very large ones.  The
 * document bytes passed to cjx_server_submit must stay valid until its
 * completion is returned.  flags: CJX_FLAG_UTF8_STRICT, CJX_FLAG_DEPTH_U16
 * (requests with different flags are never batched together).
 * ---------------------------------------------------------------------- */
typedef struct cjx_server cjx_server;

typedef struct cjx_completion {
    uint64_t ticket;      /* from cjx_server_submit                              */
    int32_t *buf;         /* 2*(n_structural+2) int32, the cjx_result layout;
                             valid until cjx_server_release                      */
    void *owner;          /* pinned region holding buf (server-internal)         */
    int64_t n_structural; /* N                                                   */
    int32_t error;        /* CJX_OK, CJX_ERR_UTF8, CJX_ERR_STRUCTURE, ...        */
    int32_t reserved;
    double latency_ms;    /* submit -> completion ready                          */
    double queue_ms;      /* submit -> batch launched                            */
} cjx_completion;

typedef struct cjx_server_stats {
    uint64_t batches, documents, bytes; /* launched batches, completed docs, admitted bytes */
    uint64_t max_inflight_bytes;        /* peak estimated device bytes in flight          */
    uint64_t relaunches;                /* batches re-run with exact output regions       */
    uint64_t pending;                   /* requests waiting for a slot                    */
    uint64_t region_bytes;              /* pinned result regions held                     */
} cjx_server_stats;

int cjx_server_create(int device, size_t device_budget, int n_slots, size_t max_batch,
                      cjx_server **out);
/* Before the first request: grows every slot's device buffers and the
 * pinned result pool for the largest batches and for single requests of up
 * to max_request_bytes, so that serving allocates nothing (device frees
 * synchronize the whole device).  CJX_ERR_INPUT once requests were submitted. */
int cjx_server_reserve(cjx_server *srv, size_t max_request_bytes);
int cjx_server_submit(cjx_server *srv, const uint8_t *data, size_t n, unsigned flags,
                      uint64_t *ticket);
/* Up to cap completions (any order); waits up to timeout_ms when none is ready. */
int cjx_server_wait(cjx_server *srv, cjx_completion *out, size_t cap, size_t *n_out,
                    int timeout_ms);
void cjx_server_release(cjx_server *srv, cjx_completion *c);
int cjx_server_get_stats(cjx_server *srv, cjx_server_stats *st);
/* Finishes every submitted request, then frees the server (completions not
 * yet released become invalid). */
void cjx_server_destroy(cjx_server *srv);

#ifdef __cplusplus
}
#endif
#endif
