// Exercises the reference-shaped C++ API (include/cujson/*.h) the way the
// reference's example mains do (main.cu:17-32, main_jsonlines.cu:16-36), and
// dumps what it got for the Python tests to compare with the oracle.
//
//   wrapper_check std   <json file>  <out.bin>   parse_standard_json
//   wrapper_check lines <jsonl file> <chunks> <out.bin>   parse_json_lines
//   wrapper_check load  <jsonl file> <chunks>    loaders only (no device)
//
// out.bin: int64 N, int64 chunkCount, int32 resultSizes[chunkCount],
// int32 resultSizesPrefix[chunkCount], int32 buffer[2(N+2)].
#define CUJSON_AMD_NO_EXIT
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "cujson/cujson.h"
#include "cujson/cujsonlines.h"

static int dump(const cuJSONResult &r, const char *path) {
    FILE *f = std::fopen(path, "wb");
    if (!f) return 1;
    const int64_t N = r.totalResultSize - 2, cc = r.chunkCount;
    std::fwrite(&N, 8, 1, f);
    std::fwrite(&cc, 8, 1, f);
    std::fwrite(r.resultSizes.data(), 4, r.resultSizes.size(), f);
    std::fwrite(r.resultSizesPrefix.data(), 4, r.resultSizesPrefix.size(), f);
    std::fwrite(r.structural, 4, 2 * (size_t)(N + 2), f);
    std::fclose(f);
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 3) return 2;
    const std::string mode = argv[1];
    if (mode == "load") {
        cuJSONLinesInput in = loadJSONLines_chunkCount(argv[2], std::strtoul(argv[3], nullptr, 10));
        std::printf("{\"size\": %zu, \"chunkCount\": %zu, \"sizes\": [", in.size, in.chunkCount);
        for (size_t i = 0; i < in.chunkCount; i++)
            std::printf("%s%zu", i ? ", " : "", in.chunksSize[i]);
        std::printf("], \"offsets\": [");
        for (size_t i = 0; i < in.chunkCount; i++)
            std::printf("%s%zu", i ? ", " : "", (size_t)(in.chunks[i] - in.data));
        std::printf("]}\n");
        cjx_input raw{in.data, in.size};
        cjx_free_input(&raw);
        return 0;
    }
    cuJSONResult r;
    const char *out;
    if (mode == "std") {
        cuJSONInput in = loadJSON(argv[2]);
        r = parse_standard_json(in);
        out = argv[3];
    } else if (mode == "lines") {
        cuJSONLinesInput in = loadJSONLines_chunkCount(argv[2], std::strtoul(argv[3], nullptr, 10));
        r = parse_json_lines(in);
        out = argv[4];
    } else {
        return 2;
    }
    if (r.structural == nullptr) {  // the reference's failure test (main.cu:32)
        std::printf("{\"error\": %d}\n", cujson_amd_last_error());
        return 3;
    }
    const int rc = dump(r, out);
    std::printf("{\"error\": 0, \"N\": %d}\n", r.totalResultSize - 2);
    cujson_free_result(r);
    return rc;
}
