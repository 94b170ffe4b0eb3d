// cujson/query_iterator.h - navigation over a parse result, with the public
// interfaces of the reference's iterators:
//   cuJSONIterator       standard JSON (cujson/query/query_iterator_standard_json.cpp:22-133)
//   cuJSONLinesIterator  JSON Lines    (cujson/query/query_iterator.cpp:22-120)
// The same member names, argument meanings and return values, over this
// build's cuJSONResult (the reference's layout).  Written for this build;
// freeJson() releases the result through cujson_free_result instead of
// cudaFreeHost.
//
// Index space: structural indices 0 .. totalResultSize-1, where 0 is a virtual
// '[' paired with the virtual ']' at totalResultSize-1 and 1..N are the
// structural characters (a '\n' of JSON Lines reads as ',', so the records of
// a JSON Lines file are the elements of one virtual top-level array).  "node"
// is the current index; key and value strings are cut from the input bytes
// between structural positions.
#ifndef CUJSON_AMD_QUERY_ITERATOR_H
#define CUJSON_AMD_QUERY_ITERATOR_H

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "cjx_bridge.h"

namespace cujson_detail {

// What both iterators share: the result's arrays, the text, character and
// value access, key search and single steps.
class IteratorCore {
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
        const int a = jumpValueForward(structural[node] - 1);
        const int b = jumpValueBackward(structural[first] - 1);
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

    // First / last byte at or from byte pos that is not a space.
    int jumpSpacesForward(int pos) {
        while (inputJSON[pos] == ' ') pos++;
        return pos;
    }
    int jumpSpacesBackward(int pos) {
        while (inputJSON[pos] == ' ') pos--;
        return pos;
    }
    // First byte of a scalar after the structural byte pos (an opening quote
    // skipped) / last byte of one before it (a closing quote skipped).
    int jumpValueForward(int pos) {
        for (int q = pos + 1;; q++) {
            if (inputJSON[q] == '"') return q + 1;
            if (inputJSON[q] != ' ') return q;
        }
    }
    int jumpValueBackward(int pos) {
        for (int q = pos - 1;; q--) {
            if (inputJSON[q] == '"') return q - 1;
            if (inputJSON[q] != ' ') return q;
        }
    }

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

  protected:
    // As the reference's constructors: the JSON text is read again from the
    // file (`failed` is printed when that fails) ...
    IteratorCore(cuJSONResult *r, const char *filePath, const char *failed) : result_(r) {
        if (!readFile(filePath, inputJSON, len)) std::cout << failed;
        owns_input_ = inputJSON != nullptr;
        attach(r);
    }
    // ... or given directly (not copied; it must outlive the iterator).
    IteratorCore(cuJSONResult *r, const uint8_t *data, size_t size) : result_(r) {
        inputJSON = const_cast<uint8_t *>(data);
        len = size;
        attach(r);
    }

    // Offset from the current node to the ':' of `key` in the object at (or
    // right after) the current node; 0 when absent or when there is no
    // object there (then `not_object` is printed, if given).
    int findKeyIn(const std::string &key, const char *not_object) {
        const char here = getChar(node);
        int obj = node;
        if (here == ':' || here == ',' || here == '\n') obj = node + 1;
        if (here == '[' && getChar(node + 1) == '{') obj = node + 1;
        if (getChar(obj) != '{') {
            if (not_object) std::cout << not_object << std::endl;
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

    // Text between the quotes of the key that ends at the ':' of index i.
    std::string keyBefore(int i) {
        const int a = jumpSpacesForward(structural[i - 1]);  // byte after the previous structural
        const int b = jumpSpacesBackward(structural[i] - 2);  // byte before the ':'
        return std::string(reinterpret_cast<const char *>(inputJSON) + a + 1, std::abs(b - a - 1));
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
};

}  // namespace cujson_detail

// Standard JSON (query_iterator_standard_json.cpp).
class cuJSONIterator : public cujson_detail::IteratorCore {
  public:
    cuJSONIterator(cuJSONResult *parsedTree, const char *filePath)
        : IteratorCore(parsedTree, filePath, "Failed to Open File for Query!\n") {}
    cuJSONIterator(cuJSONResult *parsedTree, const uint8_t *data, size_t size)
        : IteratorCore(parsedTree, data, size) {}

    void reset() {
        node = 0;
        nodeType = OBJECT;
    }

    // Offset from the current node to the ':' of `key` in the object at (or
    // right after) the current node; 0 when absent.
    int findKey(std::string key) {
        return findKeyIn(key, "ERROR: Node is not an object to find key!");
    }
    int gotoKey(std::string key) { return increamentIndex(findKey(key)); }

    // Moves to element `index` (from 0) of the array at (or right after) the
    // current node; 0 when the array is shorter.  An element that is itself
    // an array leaves the iterator on its '['.
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
};

// JSON Lines (query_iterator.cpp): the records are the elements of the
// virtual top-level array, so gotoArrayIndex(i) from the root reaches record
// i, and gotoNextSibling steps from record to record.
class cuJSONLinesIterator : public cujson_detail::IteratorCore {
  public:
    int currArrayNode = 0;  // separator before the element last reached in an array walk

    cuJSONLinesIterator(cuJSONResult *parsedTree, const char *filePath)
        : IteratorCore(parsedTree, filePath,
                       "[ERR] File loading for the query iterator failed. Please check the "
                       "file path.\n") {}
    cuJSONLinesIterator(cuJSONResult *parsedTree, const uint8_t *data, size_t size)
        : IteratorCore(parsedTree, data, size) {}

    void reset() {
        node = 0;
        currArrayNode = 0;
        nodeType = OBJECT;
    }

    // As the standard iterator's, without the message (query_iterator.cpp:448-488).
    int findKey(std::string key) { return findKeyIn(key, nullptr); }
    int gotoKey(std::string key) { return increamentIndex(findKey(key)); }

    // Moves to element `index` (from 0) of the array at (or right after) the
    // current node; 0 when the array is shorter.  The iterator stops on the
    // separator (or '[') before the element, also when the element is an
    // array (query_iterator.cpp:233-299).  Every separator passed is kept in
    // currArrayNode.
    int gotoArrayIndex(int index) {
        const char here = getChar(node);
        if (here == ',' || here == ':') increamentIndex(1);
        return walkElements(index, false);
    }

    // From the separator kept in currArrayNode, moves `index` elements on in
    // the same array (query_iterator.cpp:358-372); 0 past the array's end.
    int gotoNextSibling(int index) {
        node = currArrayNode;
        return gotoArrayIndexSiblingHelper(index);
    }

    // Moves `index` elements on from the current node, which must be the
    // separator (or '[') before an element, and keeps the node reached in
    // currArrayNode (query_iterator.cpp:301-356).
    int gotoArrayIndexSiblingHelper(int index) { return walkElements(index, true); }

    // Whether the object at (or right after) the current node holds `key`
    // with the value text `value` (as getValue() gives it).  The node is
    // left where it was; nodeType is that of the key's ':'
    // (query_iterator.cpp:601-629).
    bool checkKeyValue(std::string key, std::string value) {
        const int d = findKey(key);
        if (d == 0) return false;
        const int here = node;
        increamentIndex(d);
        const bool same = getValue() == value;
        node = here;
        return same;
    }

    // ']' for the virtual end, '\0' for every other index.
    char getArtificalChar(int idx) { return idx == totalResultSize - 1 ? ']' : '\0'; }

    // First index after idx whose character is not a closing bracket (the
    // virtual end when only closers follow).
    int jumpForwardStructural(int idx) {
        int i = idx + 1;
        while (i < totalResultSize - 1) {
            const char c = getChar(i);
            if (c != '}' && c != ']') break;
            i++;
        }
        return i;
    }

  private:
    // Passes `index` separators of the array level that starts after the
    // current node, skipping nested values whole; stops on the last one.
    int walkElements(int index, bool sibling) {
        int remaining = index;
        int next = node + 1;
        char c = getChar(next);
        while (remaining != 0 && c != ']' && next != totalResultSize - 1) {
            if (c == '[' || c == '{') {
                next = jumpOpeningForward(next);
            } else if (c == ',') {
                if (!sibling) currArrayNode = next;
                remaining--;
            }
            c = getChar(++next);
        }
        if (remaining != 0) return 0;
        node = next - 1;
        if (sibling) currArrayNode = node;
        if (c == '{') {
            nodeType = OBJECT;
        } else if (c == '[') {
            nodeType = ARRAY;
        } else if (c == ',' || c == ']') {
            nodeType = VALUE;
        }
        return node;
    }
};

#endif  // CUJSON_AMD_QUERY_ITERATOR_H
