// Runs queries through include/cujson/query_iterator.h the way the
// reference's mains do (main.cu:46-58, main_jsonlines.cu:60-75,
// real_world_use_case/main.cu:66-87): a parse, an iterator over the result
// and the file, then per query reset() and a sequence of steps, and
// getValue() at the end.
//
//   query_check <json file> <queries file>
//   query_check --lines <chunk count> <json lines file> <queries file>
//   query_check --result std|lines <result file> <json file> <queries file>
// The third form takes the parse result from a file of int32 cells in the
// host layout instead of parsing (no device needed).
//
// A queries file has one query per line, steps separated by spaces:
//   i:<n>  gotoArrayIndex(n)      k:<key>  gotoKey(key)
//   s:<n>  gotoNextSibling(n)     (JSON Lines only)
// Output: one line per query, "ok <hex of getValue()>" or "miss <step>".
//
// A JSON Lines query line may instead be a scan over the records, the loop
// of real_world_use_case/main.cu:
//   scan <key> <value> <key2> <records>
// gotoArrayIndex(0), then per record: if checkKeyValue(key, value),
// gotoKey(key2) and take getValue(); gotoNextSibling(1).  Output:
// "scan <count> <hex of the values, each followed by a newline>".
#define CUJSON_AMD_NO_EXIT
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <type_traits>
#include <vector>

#include "cujson/cujson.h"
#include "cujson/cujsonlines.h"
#include "cujson/query_iterator.h"

static void print_hex(const std::string &v) {
    for (unsigned char c : v) std::printf("%02x", c);
}

template <class Iterator>
static bool run_steps(Iterator &itr, std::istringstream &steps);

template <>
bool run_steps(cuJSONIterator &itr, std::istringstream &steps) {
    std::string step;
    for (int k = 0; steps >> step; k++) {
        const int at = step[0] == 'i' ? itr.gotoArrayIndex(std::stoi(step.substr(2)))
                                      : itr.gotoKey(step.substr(2));
        // gotoArrayIndex(0) from the root stays at node 0 legitimately
        if (at == 0 && !(step == "i:0" && k == 0)) {
            std::printf("miss %d\n", k);
            return false;
        }
    }
    return true;
}

template <>
bool run_steps(cuJSONLinesIterator &itr, std::istringstream &steps) {
    std::string step;
    for (int k = 0; steps >> step; k++) {
        int at;
        if (step[0] == 'i')
            at = itr.gotoArrayIndex(std::stoi(step.substr(2)));
        else if (step[0] == 's')
            at = itr.gotoNextSibling(std::stoi(step.substr(2)));
        else
            at = itr.gotoKey(step.substr(2));
        if (at == 0 && !(step == "i:0" && k == 0)) {
            std::printf("miss %d\n", k);
            return false;
        }
    }
    return true;
}

static void run_scan(cuJSONLinesIterator &itr, std::istringstream &args) {
    std::string key, value, key2;
    long records = 0;
    args >> key >> value >> key2 >> records;
    itr.reset();
    itr.gotoArrayIndex(0);
    long count = 0;
    std::string found;
    for (long i = 0; i < records; i++) {
        if (itr.checkKeyValue(key, value)) {
            itr.gotoKey(key2);
            found += itr.getValue();
            found += '\n';
            count++;
        }
        itr.gotoNextSibling(1);
    }
    std::printf("scan %ld ", count);
    print_hex(found);
    std::printf("\n");
}

template <class Iterator>
static void run_queries(Iterator &itr, const char *queries) {
    std::ifstream q(queries);
    std::string line;
    while (std::getline(q, line)) {
        if constexpr (std::is_same_v<Iterator, cuJSONLinesIterator>) {
            std::istringstream words(line);
            std::string first;
            words >> first;
            if (first == "scan") {
                run_scan(itr, words);
                continue;
            }
        }
        itr.reset();
        std::istringstream steps(line);
        if (!run_steps(itr, steps)) continue;
        std::printf("ok ");
        print_hex(itr.getValue());
        std::printf("\n");
    }
}

// A result read from a file of int32 cells in the host layout (2(N+2)
// cells: structural rows then pair rows), with the fields a parse sets.
static bool read_result(const char *path, std::vector<int32_t> &cells, cuJSONResult &r) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) return false;
    const std::streamsize bytes = f.tellg();
    if (bytes < 16 || bytes % 8) return false;
    cells.resize((size_t)bytes / 4);
    f.seekg(0);
    f.read(reinterpret_cast<char *>(cells.data()), bytes);
    const int N = (int)(cells.size() / 2) - 2;
    r = cuJSONResult{};
    r.structural = cells.data();
    r.pair_pos = cells.data() + N + 1;
    r.totalResultSize = r.fileSize = N + 2;
    r.chunkCount = 1;
    r.resultSizes = {N};
    r.resultSizesPrefix = {0};
    return true;
}

static bool read_text(const char *path, std::vector<uint8_t> &text) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) return false;
    text.resize((size_t)f.tellg() + 1);
    f.seekg(0);
    f.read(reinterpret_cast<char *>(text.data()), (std::streamsize)text.size() - 1);
    return true;
}

int main(int argc, char **argv) {
    if (argc == 3) {
        cuJSONResult r = parse_standard_json(loadJSON(argv[1]));
        if (r.structural == nullptr) return 3;
        cuJSONIterator itr(&r, argv[1]);
        run_queries(itr, argv[2]);
        itr.freeJson();
        return 0;
    }
    if (argc == 5 && std::strcmp(argv[1], "--lines") == 0) {
        cuJSONResult r =
            parse_json_lines(loadJSONLines_chunkCount(argv[3], (size_t)std::stoul(argv[2])));
        if (r.structural == nullptr) return 3;
        cuJSONLinesIterator itr(&r, argv[3]);
        run_queries(itr, argv[4]);
        itr.freeJson();
        return 0;
    }
    if (argc == 6 && std::strcmp(argv[1], "--result") == 0) {
        std::vector<int32_t> cells;
        std::vector<uint8_t> text;
        cuJSONResult r;
        if (!read_result(argv[3], cells, r) || !read_text(argv[4], text)) return 3;
        if (std::strcmp(argv[2], "lines") == 0) {
            cuJSONLinesIterator itr(&r, text.data(), text.size() - 1);
            run_queries(itr, argv[5]);
        } else {
            cuJSONIterator itr(&r, text.data(), text.size() - 1);
            run_queries(itr, argv[5]);
        }
        return 0;
    }
    return 2;
}
