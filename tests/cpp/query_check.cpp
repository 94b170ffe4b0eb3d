// Runs queries through include/cujson/query_iterator.h the way the
// reference's mains do (main.cu:46-58, main_jsonlines_chunksize.cu:66-71):
// parse_standard_json, a cuJSONIterator over the result and the file, then
// per query reset() and a sequence of steps, and getValue() at the end.
//
//   query_check <json file> <queries file>
//
// A queries file has one query per line, steps separated by spaces:
//   i:<n>  gotoArrayIndex(n)      k:<key>  gotoKey(key)
// Output: one line per query, "ok <hex of getValue()>" or "miss <step>".
#define CUJSON_AMD_NO_EXIT
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>

#include "cujson/cujson.h"
#include "cujson/query_iterator.h"

int main(int argc, char **argv) {
    if (argc != 3) return 2;
    cuJSONResult r = parse_standard_json(loadJSON(argv[1]));
    if (r.structural == nullptr) return 3;
    cuJSONIterator itr(&r, argv[1]);
    std::ifstream q(argv[2]);
    std::string line;
    while (std::getline(q, line)) {
        itr.reset();
        std::istringstream steps(line);
        std::string step;
        int k = 0;
        bool ok = true;
        while (steps >> step) {
            const int at = step[0] == 'i' ? itr.gotoArrayIndex(std::stoi(step.substr(2)))
                                          : itr.gotoKey(step.substr(2));
            // gotoArrayIndex(0) from the root stays at node 0 legitimately
            if (at == 0 && !(step == "i:0" && k == 0)) {
                std::printf("miss %d\n", k);
                ok = false;
                break;
            }
            k++;
        }
        if (!ok) continue;
        const std::string v = itr.getValue();
        std::printf("ok ");
        for (unsigned char c : v) std::printf("%02x", c);
        std::printf("\n");
    }
    itr.freeJson();
    return 0;
}
