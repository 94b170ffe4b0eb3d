// cujson/cujsonlines.h - drop-in for the reference's JSON Lines header
// (cujson/cujsonlines.h): loadJSONLines_* + parse_json_lines on the C ABI.
#ifndef CUJSON_AMD_CUJSONLINES_H
#define CUJSON_AMD_CUJSONLINES_H

#include <vector>

#include "cjx_bridge.h"

// parse_json_lines (cujson/parse_json_lines.h:9, parse_json_lines.cu:1092-1264).
// As in the reference (parse_json_lines.cu:1261), the input buffer is released.
inline cuJSONResult parse_json_lines(cuJSONLinesInput input) {
    if (input.data == nullptr || input.size == 0 || input.chunkCount == 0 ||
        input.chunks.size() != input.chunkCount || input.chunksSize.size() != input.chunkCount) {
        std::cerr << "\033[1;31m Error: Invalid JSON Lines input. \033[0m\n";
        return cuJSONResult{};
    }
    cjx_context *ctx = cujson_amd_detail::context();
    if (!ctx) {  // no device: the input is released all the same
        cjx_input in{input.data, input.size};
        cjx_free_input(&in);
        return cuJSONResult{};
    }
    std::vector<size_t> off(input.chunkCount);
    for (size_t i = 0; i < input.chunkCount; i++) off[i] = (size_t)(input.chunks[i] - input.data);
    cjx_result r{};
    const int rc = cjx_parse_lines(ctx, input.data, input.size, off.data(),
                                   input.chunksSize.data(), input.chunkCount, 0, &r);
    cjx_input in{input.data, input.size};
    cjx_free_input(&in);
    return cujson_amd_detail::to_result(r, rc);
}

#endif  // CUJSON_AMD_CUJSONLINES_H
