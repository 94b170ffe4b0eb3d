// cujson/cujson_types.h - the reference's public types
// (cujson/cujson_types.h:9-38), kept field for field so that code written
// against them (the query iterator, the example mains) compiles unchanged.
#ifndef CUJSON_AMD_TYPES_H
#define CUJSON_AMD_TYPES_H

#include <cstddef>
#include <cstdint>
#include <vector>

struct cuJSONInput {
    uint8_t *data;  // pinned host buffer (cjx_load)
    size_t size;
};

struct cuJSONLinesInput {
    uint8_t *data;
    size_t chunkCount;
    size_t size;
    std::vector<uint8_t *> chunks;   // start of each line-aligned chunk
    std::vector<size_t> chunksSize;  // its size in bytes
};

struct cuJSONResult {
    uint8_t *inputJSON;  // never set by the reference either
    int chunkCount;
    int bufferSize;      // never set by the reference either
    std::vector<int> resultSizes;
    std::vector<int> resultSizesPrefix;
    int32_t *structural;  // 2(N+2) int32 pinned buffer, see cujson_amd.h
    int32_t *pair_pos;    // structural + N + 1
    int depth;            // never set by the reference either
    int totalResultSize;  // N + 2
    int fileSize;         // N + 2
};

enum tokens_type_enum { OBJECT, ARRAY, KEYVALUE, VALUE, CLOSING };
typedef tokens_type_enum token_type;

enum primitive_type_enum { NUMBER, TRUE, FALSE, NULL_TYPE, STRING };
typedef primitive_type_enum primitive_type;

#endif  // CUJSON_AMD_TYPES_H
