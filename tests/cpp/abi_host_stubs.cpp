// Host-only stand-ins for the C ABI entry points that need a HIP device
// (include/cujson_amd.h), for the sanitizer build of the reference-shaped
// headers (include/cujson/) and their test drivers, which links no HIP
// runtime.  Results and inputs live in malloc'd memory here instead of
// pinned memory; the parse entry points report that no device is present.
// The chunker is the library's own (cjx_chunk.cpp is linked in).
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "cujson_amd.h"

extern "C" {

const char *cjx_strerror(int code) {
    switch (code) {
    case CJX_OK: return "ok";
    case CJX_ERR_INPUT: return "invalid input";
    case CJX_ERR_UTF8: return "invalid UTF-8";
    case CJX_ERR_STRUCTURE: return "invalid structure";
    case CJX_ERR_ALLOC: return "allocation failed";
    case CJX_ERR_DEVICE: return "no HIP device (host-only build)";
    default: return "error";
    }
}

int cjx_context_create(int, cjx_context **out) {
    if (out) *out = nullptr;
    return CJX_ERR_DEVICE;
}

void cjx_context_destroy(cjx_context *) {}

int cjx_load(const char *path, cjx_input *out) {
    if (!path || !out) return CJX_ERR_INPUT;
    out->data = nullptr;
    out->size = 0;
    FILE *f = std::fopen(path, "rb");
    if (!f) return CJX_ERR_INPUT;
    std::fseek(f, 0, SEEK_END);
    const long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    uint8_t *buf = static_cast<uint8_t *>(std::malloc(sz > 0 ? (size_t)sz : 1));
    if (!buf) {
        std::fclose(f);
        return CJX_ERR_ALLOC;
    }
    const size_t got = std::fread(buf, 1, (size_t)sz, f);
    std::fclose(f);
    if (got != (size_t)sz) {
        std::free(buf);
        return CJX_ERR_INPUT;
    }
    out->data = buf;
    out->size = (size_t)sz;
    return CJX_OK;
}

void cjx_free_input(cjx_input *in) {
    if (!in) return;
    std::free(in->data);
    in->data = nullptr;
    in->size = 0;
}

void cjx_free_result(cjx_result *r) {
    if (!r) return;
    std::free(r->buf);
    std::free(r->records);
    std::free(r->result_sizes);
    std::free(r->result_sizes_prefix);
    std::memset(r, 0, sizeof(*r));
    r->error_chunk = -1;
}

int cjx_parse_standard(cjx_context *, const uint8_t *, size_t, unsigned, cjx_result *out) {
    if (out) {
        std::memset(out, 0, sizeof(*out));
        out->error = CJX_ERR_DEVICE;
        out->error_chunk = -1;
    }
    return CJX_ERR_DEVICE;
}

int cjx_parse_lines(cjx_context *, const uint8_t *, size_t, const size_t *, const size_t *,
                    size_t, unsigned, cjx_result *out) {
    if (out) {
        std::memset(out, 0, sizeof(*out));
        out->error = CJX_ERR_DEVICE;
        out->error_chunk = -1;
    }
    return CJX_ERR_DEVICE;
}

}  // extern "C"
