// pison_shim.cpp - C entry point over Pison, the CPU parser the reference
// compares against (paper_reproduced/related_works/pison/standard_JSON/src),
// compiled from the reference's vendored sources by oracle/Makefile.  It is
// a CPU baseline only: it times the leveled-bitmap construction that the
// reference's harness times (example/twitter.cpp:104-111: construct plus
// getIterator; record loading excluded).
#include <chrono>
#include <cstdlib>
#include <cstring>

#include "BitmapConstructor.h"
#include "Records.h"

extern "C" double pison_construct_ms(const char *data, size_t n, int threads, int levels,
                                     int reps) {
    // The loader's layout (RecordLoader.cpp:7-41): 64-byte aligned, padded
    // with 'd' to a multiple of 64, then a NUL.
    const size_t pad = 64 - (n % 64);
    void *p = nullptr;
    if (posix_memalign(&p, 64, n + pad + 64) != 0) return -1.0;
    char *text = static_cast<char *>(p);
    std::memcpy(text, data, n);
    std::memset(text + n, 'd', pad);
    text[n + pad] = '\0';
    Record rec;
    rec.text = text;
    rec.rec_start_pos = 0;
    rec.rec_length = (long)(n + pad);
    rec.can_delete_text = false;
    double best = -1.0;
    for (int r = 0; r < reps; r++) {
        const auto t0 = std::chrono::steady_clock::now();
        Bitmap *bm = BitmapConstructor::construct(&rec, threads, levels);
        BitmapIterator *it = BitmapConstructor::getIterator(bm);
        const auto t1 = std::chrono::steady_clock::now();
        delete it;
        delete bm;
        const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        if (best < 0 || ms < best) best = ms;
    }
    std::free(text);
    return best;
}
