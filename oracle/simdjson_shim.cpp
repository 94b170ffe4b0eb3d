// simdjson_shim.cpp - C entry points over the reference's vendored simdjson 3.2.1
// (paper_reproduced/related_works/simdjson/simdjson.{h,cpp}), built into
// oracle/_ref/libsimdjson_ref.so by oracle/Makefile.
//
// TEST INFRASTRUCTURE ONLY: the structural-index cross-check in tests/ and the
// CPU baseline leg of bench.py.  The product library never loads it.
//
// sjr_stage1 follows simdjson's own stage 1 (dom_parser_implementation::stage1,
// simdjson.h:3022, members at :3115-3117) and returns every structural index it
// finds; the caller filters them to the bytes { } [ ] : , that cuJSON records.
// sjr_iterate_ms times ondemand::parser::iterate, the call the reference's
// harness times (quickstart-original-iterate.cpp:105-117).
#include <chrono>
#include <cstdint>
#include <cstring>

#include "simdjson.h"

using namespace simdjson;

extern "C" {

// Returns 0 on success, otherwise the simdjson error code.  *n_out receives
// the index count (capacity must be >= len + 2).
int sjr_stage1(const uint8_t *buf, size_t len, uint32_t *idx_out, size_t cap, size_t *n_out) {
    padded_string ps(reinterpret_cast<const char *>(buf), len);
    dom::parser parser;
    error_code e = parser.allocate(len);
    if (e) return int(e);
    e = parser.implementation->stage1(reinterpret_cast<const uint8_t *>(ps.data()), len,
                                      stage1_mode::regular);
    size_t n = parser.implementation->n_structural_indexes;
    if (n > cap) n = cap;
    std::memcpy(idx_out, parser.implementation->structural_indexes.get(), n * sizeof(uint32_t));
    *n_out = n;
    return int(e);
}

// RFC 3629 validation (simdjson::validate_utf8, simdjson.h:3212).  1 = valid.
int sjr_validate_utf8(const uint8_t *buf, size_t len) {
    return validate_utf8(reinterpret_cast<const char *>(buf), len) ? 1 : 0;
}

// Median-free simple timer: best of `reps` runs of ondemand iterate (stage 1)
// over one document already resident in host memory, in milliseconds.
// *err_out receives the first error code seen (0 = none).
double sjr_iterate_ms(const uint8_t *buf, size_t len, int reps, int *err_out) {
    padded_string ps(reinterpret_cast<const char *>(buf), len);
    ondemand::parser parser;
    double best = 1e30;
    int err = 0;
    for (int r = 0; r < reps; r++) {
        auto t0 = std::chrono::high_resolution_clock::now();
        auto doc = parser.iterate(ps);
        auto t1 = std::chrono::high_resolution_clock::now();
        if (doc.error() && !err) err = int(doc.error());
        double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        if (ms < best) best = ms;
    }
    if (err_out) *err_out = err;
    return best;
}

// JSON Lines: time parse_many over the whole buffer (quickstart-original-parse-many.cpp),
// returning milliseconds for one full pass and the document count.
double sjr_parse_many_ms(const uint8_t *buf, size_t len, size_t batch, size_t *ndocs_out) {
    padded_string ps(reinterpret_cast<const char *>(buf), len);
    ondemand::parser parser;
    auto t0 = std::chrono::high_resolution_clock::now();
    size_t nd = 0;
    ondemand::document_stream stream;
    if (!parser.iterate_many(ps, batch).get(stream)) {
        for (auto doc : stream) {
            (void)doc;
            nd++;
        }
    }
    auto t1 = std::chrono::high_resolution_clock::now();
    if (ndocs_out) *ndocs_out = nd;
    return std::chrono::duration<double, std::milli>(t1 - t0).count();
}

// Many independent documents (data + off[i], size[i] bytes each): ondemand
// iterate of every document in turn with one parser, the reference's
// one-call-per-document form (main.cu:26) on the CPU; the documents are
// copied with simdjson's padding before the timed loop.  Best of `reps`
// passes in milliseconds; *err_docs receives how many documents failed.
double sjr_iterate_docs_ms(const uint8_t *data, const size_t *off, const size_t *size,
                           size_t count, int reps, size_t *err_docs) {
    size_t total = 0;
    for (size_t i = 0; i < count; i++) total += size[i] + SIMDJSON_PADDING;
    char *buf = new char[total + 64];
    size_t *pos = new size_t[count];
    size_t at = 0;
    for (size_t i = 0; i < count; i++) {
        pos[i] = at;
        std::memcpy(buf + at, data + off[i], size[i]);
        std::memset(buf + at + size[i], 0, SIMDJSON_PADDING);
        at += size[i] + SIMDJSON_PADDING;
    }
    ondemand::parser parser;
    double best = 1e30;
    size_t errs = 0;
    for (int r = 0; r < reps; r++) {
        size_t e = 0;
        auto t0 = std::chrono::high_resolution_clock::now();
        for (size_t i = 0; i < count; i++) {
            auto doc = parser.iterate(padded_string_view(buf + pos[i], size[i],
                                                         size[i] + SIMDJSON_PADDING));
            if (doc.error()) e++;
        }
        auto t1 = std::chrono::high_resolution_clock::now();
        errs = e;
        double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        if (ms < best) best = ms;
    }
    delete[] pos;
    delete[] buf;
    if (err_docs) *err_docs = errs;
    return best;
}

}  // extern "C"
