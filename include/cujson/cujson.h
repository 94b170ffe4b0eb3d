// cujson/cujson.h - drop-in for the reference's standard-JSON header
// (cujson/cujson.h): loadJSON + parse_standard_json on the MI355X C ABI.
#ifndef CUJSON_AMD_CUJSON_H
#define CUJSON_AMD_CUJSON_H

#include "cjx_bridge.h"

// parse_standard_json (cujson/parse_standard_json.h:9, parse_standard_json.cu:1584-1687).
// As in the reference, the input buffer is released by the parse.
inline cuJSONResult parse_standard_json(cuJSONInput input) {
    if (input.data == nullptr || input.size == 0) {
        std::cerr << "\033[1;31m Error: Invalid JSON content or input.size. \033[0m\n";
        return cuJSONResult{};
    }
    cjx_context *ctx = cujson_amd_detail::context();
    if (!ctx) {  // no device: the input is released all the same
        cjx_input in{input.data, input.size};
        cjx_free_input(&in);
        return cuJSONResult{};
    }
    cjx_result r{};
    const int rc = cjx_parse_standard(ctx, input.data, input.size, 0, &r);
    cjx_input in{input.data, input.size};
    cjx_free_input(&in);
    return cujson_amd_detail::to_result(r, rc);
}

#endif  // CUJSON_AMD_CUJSON_H
