// cujson/query_iterator.h - navigation over a parse result, with the public
// interface of the reference's cuJSONIterator
// (cujson/query/query_iterator_standard_json.cpp:22-133): the same member
// names, argument meanings and return values, over this build's cuJSONResult
// (the reference's layout).  Written for this build; freeJson() releases the
// result through cujson_free_result instead of cudaFreeHost.
//
// Index space: structural indices 0 .. totalResultSize-1, where 0 is a virtual
// '[' paired with the virtual ']' at totalResultSize-1 and 1..N are the
// structural characters (a '\n' of JSON Lines reads as ',').  "node" is the
// current index; key and value strings are cut from the input bytes between
// structural positions.
#ifndef CUJSON_AMD_QUERY_ITERATOR_H
#define CUJSON_AMD_QUERY_ITERATOR_H

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "cjx_bridge.h"

class cuJSONIterator {
  public:
    uint8_t *inputJSON = nullptr;
    std::vector<int> resultSizes;
    std::vector<int> resultSizesPrefix;
    int32_t *structural = nullptr;
    int32_t *pair_pos = nullptr;
    int jsonDepth = 0;
    int bufferSize = 0;
    int totalResultSize = 0;
    int chunkCount = 0;
    int currentChunkIndex = 0;
    int fileSize = 0;
    token_type nodeType = OBJECT;
    int node = 0;
    size_t len = 0;

    // As the reference (:80-100): the JSON text is read again from the file.
    cuJSONIterator(cuJSONResult *parsedTree, const char *filePath) : result_(parsedTree) {
        if (!readFile(filePath, inputJSON, len)) std::cout << "Failed to Open File for Query!\n";
        owns_input_ = inputJSON != nullptr;
        attach(parsedTree);
    }
    // The JSON text given directly (not copied; it must outlive the iterator).
    cuJSONIterator(cuJSONResult *parsedTree, const uint8_t *data, size_t size)
        : result_(parsedTree) {
        inputJSON = const_cast<uint8_t *>(data);
        len = size;
        attach(parsedTree);
    }

    // ---- navigation -------------------------------------------------------
    void reset() {
        node = 0;
        nodeType = OBJECT;
    }

    // Moves by `index` nodes and classifies the node reached; 0 when its
    // character is not structural.
    int increamentIndex(int index) {
        node += index;
        switch (getChar(node)) {
            case '{': nodeType = OBJECT; break;
            case '[': nodeType = ARRAY; break;
            case ',':
            case '\n': nodeType = VALUE; break;
            case ':': nodeType = KEYVALUE; break;
            case ']':
            case '}': nodeType = CLOSING; break;
            default: return 0;
        }
        return node;
    }

    // Offset from the current node to the ':' of `key` in the object at (or
    // right after) the current node; 0 when absent.
    int findKey(std::string key) {
        const char here = getChar(node);
        int obj = node;
        if (here == ':' || here == ',' || here == '\n') obj = node + 1;
        if (here == '[' && getChar(node + 1) == '{') obj = node + 1;
        if (getChar(obj) != '{') {
            std::cout << "ERROR: Node is not an object to find key!" << std::endl;
            return 0;
        }
        const int end = jumpOpeningForward(obj);
        for (int i = obj + 1; i < end; i++) {
            const char c = getChar(i);
            if (c == '}') break;
            if (c == '[' || c == '{') {
                i = jumpOpeningForward(i);  // nested values are skipped whole
                continue;
            }
            if (c == ':' && keyBefore(i) == key) return i - node;
        }
        return 0;
    }

    int gotoKey(std::string key) { return increamentIndex(findKey(key)); }

    // Moves to element `index` (from 0) of the array at (or right after) the
    // current node; 0 when the array is shorter.
    int gotoArrayIndex(int index) {
        int remaining = index + 1;
        const char here = getChar(node);
        if (here == ',' || here == '\n' || here == ':') increamentIndex(1);
        int next = node + 1;
        char c = getChar(next);
        while (remaining != 1 && c != ']' && next != totalResultSize - 1) {
            if (c == '[' || c == '{') next = jumpOpeningForward(next);
            if (c == ',' || c == '\n') remaining--;
            c = getChar(++next);
        }
        if (remaining != 1) return 0;
        node = next - 1;
        if (c == '{') {
            nodeType = OBJECT;
        } else if (c == '[') {
            nodeType = ARRAY;
            node = next;
        } else if (c == ',' || here == '\n' || c == ']') {
            nodeType = VALUE;
        }
        return node;
    }

    // ---- access -------------------------------------------------------------
    // Key of the ':' node (text between its quotes).
    std::string getKey() {
        if (getChar(node) != ':') {
            std::cout << "ERROR! The iterator must point to a colon.";
#ifndef CUJSON_AMD_NO_EXIT
            std::exit(0);
#endif
            return std::string();
        }
        return keyBefore(node);
    }

    // Value after the current ':' / ',' / '\n' node, or the first element
    // after a '[' node: an object or array whole, a string without its
    // quotes, other scalars as written.
    std::string getValue() {
        const char here = getChar(node);
        if (here != ',' && here != '\n' && here != ':' && here != '[') {
            std::cout << "ERROR: iterator is in wrong place." << std::endl;
            return std::string();
        }
        const int first = node + 1;
        const char c = getChar(first);
        if (c == '[' || c == '{') {
            const int a = structural[first] - 1, b = structural[jumpOpeningForward(first)] - 1;
            return std::string(reinterpret_cast<const char *>(inputJSON) + a, b - a + 1);
        }
        const int a = skipValueStart(structural[node] - 1);
        const int b = skipValueEnd(structural[first] - 1);
        return std::string(reinterpret_cast<const char *>(inputJSON) + a, b - a + 1);
    }

    // Character of structural index idx ('\n' reads as ','; the two virtual
    // ends as '[' and ']').
    char getChar(int idx) {
        if (idx == 0) return '[';
        if (idx == totalResultSize - 1) return ']';
        const char c = static_cast<char>(inputJSON[structural[idx] - 1]);
        return c == '\n' ? ',' : c;
    }

    int jumpOpeningForward(int idx) { return pair_pos[idx]; }

    // Releases the text read by the file constructor and the parse result.
    void freeJson() {
        if (owns_input_) delete[] inputJSON;
        inputJSON = nullptr;
        owns_input_ = false;
        if (result_) cujson_free_result(*result_);
        structural = pair_pos = nullptr;
    }

    bool readFile(const char *file, uint8_t *&buffer, size_t &size) {
        FILE *f = std::fopen(file, "rb");
        if (!f) {
            std::cerr << "File not found!" << std::endl;
            return false;
        }
        std::fseek(f, 0, SEEK_END);
        size = (size_t)std::ftell(f);
        std::fseek(f, 0, SEEK_SET);
        buffer = new uint8_t[size ? size : 1];
        const size_t got = std::fread(buffer, 1, size, f);
        std::fclose(f);
        if (got != size) {
            std::cerr << "Reading error!" << std::endl;
            delete[] buffer;
            buffer = nullptr;
            return false;
        }
        return true;
    }

  private:
    cuJSONResult *result_ = nullptr;
    bool owns_input_ = false;

    void attach(cuJSONResult *r) {
        resultSizes = r->resultSizes;
        resultSizesPrefix = r->resultSizesPrefix;
        totalResultSize = r->totalResultSize;
        structural = r->structural;
        pair_pos = r->pair_pos;
        fileSize = r->fileSize;
        jsonDepth = r->depth;
        bufferSize = r->bufferSize;
        chunkCount = r->chunkCount;
        currentChunkIndex = 0;
        if (structural && totalResultSize > 1) {  // the virtual ends, as the reference sets them
            structural[0] = 0;
            structural[totalResultSize - 1] = fileSize - 1;
            pair_pos[0] = totalResultSize - 1;
        }
    }

    // Text between the quotes of the key that ends at the ':' of index i.
    std::string keyBefore(int i) {
        int a = structural[i - 1];  // byte after the previous structural
        int b = structural[i] - 2;  // byte before the ':'
        while (inputJSON[a] == ' ') a++;
        while (inputJSON[b] == ' ') b--;
        return std::string(reinterpret_cast<const char *>(inputJSON) + a + 1, std::abs(b - a - 1));
    }
    // First byte of a scalar after structural byte p (an opening quote skipped).
    int skipValueStart(int p) {
        for (int q = p + 1;; q++) {
            if (inputJSON[q] == '"') return q + 1;
            if (inputJSON[q] != ' ') return q;
        }
    }
    // Last byte of a scalar before structural byte p (a closing quote skipped).
    int skipValueEnd(int p) {
        for (int q = p - 1;; q--) {
            if (inputJSON[q] == '"') return q - 1;
            if (inputJSON[q] != ' ') return q;
        }
    }
};

#endif  // CUJSON_AMD_QUERY_ITERATOR_H
