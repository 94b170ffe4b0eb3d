"""C-ABI library checks that need no GPU: it loads, exports every symbol the
public header declares, and its host-only entry points behave."""
import ctypes
import os
import re

import numpy as np
import pytest

import cujson_amd as cj
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cujson_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cjx_[a-z_0-9]+)\s*\(", text)))


def test_library_loads_and_exports_header_symbols():
    L = cj.lib()
    syms = declared_symbols()
    assert len(syms) >= 12
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(cj.ABI_SYMBOLS)


def test_strerror_and_version():
    assert cj.strerror(cj.CJX_OK) == "ok"
    assert "UTF-8" in cj.strerror(cj.CJX_ERR_UTF8)
    assert cj.lib().cjx_version().startswith(b"cujson_amd")


@pytest.mark.parametrize("cc", [1, 2, 4, 7, 51, 60])
def test_chunk_lines_count_mode_matches_loader_rule(twitter_small, cc):
    offs, sizes = cj.chunk_lines(twitter_small, 0, cc)
    assert list(zip(offs, sizes)) == oracle.chunks_by_count(twitter_small, cc)


@pytest.mark.parametrize("nb", [1, 100, 4096, 65536, 10 ** 7])
def test_chunk_lines_bytes_mode_matches_loader_rule(twitter_small, nb):
    offs, sizes = cj.chunk_lines(twitter_small, 1, nb)
    assert list(zip(offs, sizes)) == oracle.chunks_by_bytes(twitter_small, nb)


def test_chunk_lines_threaded_scan_matches_loader_rule():
    """From 64 MiB the host chunker scans the text on several threads: an
    80 MB text of random line lengths (blank lines, an unterminated last
    line) gives the loader rule's chunks in both modes."""
    rng = np.random.default_rng(5)
    lens = rng.integers(0, 2000, 80_000)
    lens[rng.integers(0, lens.size, 2000)] = 0  # blank lines
    a = np.full(int(lens.sum()) + lens.size + 17, ord("x"), np.uint8)
    a[np.cumsum(lens + 1)[:-1] - 1] = 10
    a[-17:] = ord("y")  # unterminated last line
    assert a.size >= 64 << 20
    for mode, param, want in ((0, 997, oracle.chunks_by_count(a, 997)),
                              (1, 1 << 20, oracle.chunks_by_bytes(a, 1 << 20))):
        offs, sizes = cj.chunk_lines(a, mode, param)
        assert list(zip(offs, sizes)) == want, mode


def test_chunk_lines_unterminated_last_line():
    data = b'{"a":1}\n{"b":2}'
    offs, sizes = cj.chunk_lines(data, 0, 2)
    assert list(zip(offs, sizes)) == [(0, 8), (8, 7)]


def test_entry_points_reject_bad_arguments_without_device():
    r = cj._Result()
    assert cj.lib().cjx_parse_standard(None, None, 0, 0, ctypes.byref(r)) == cj.CJX_ERR_INPUT
    assert r.error == cj.CJX_ERR_INPUT and not r.buf
    assert cj.lib().cjx_parse_lines(None, None, 0, None, None, 0, 0,
                                    ctypes.byref(r)) == cj.CJX_ERR_INPUT


def test_tokenizer_shape_entries_reject_bad_arguments_without_device():
    L = cj.lib()
    assert L.cjx_context_set_tokenizer_shape(None, 4, 4) == cj.CJX_ERR_INPUT
    t, s = ctypes.c_int(), ctypes.c_int()
    assert L.cjx_context_tokenizer_shape(None, ctypes.byref(t), ctypes.byref(s)) == cj.CJX_ERR_INPUT


def test_loader_entries_reject_bad_arguments_without_device():
    L = cj.lib()
    n = ctypes.c_size_t(7)
    assert L.cjx_scan_lines_device(None, None, 0, None, 0, ctypes.byref(n)) == cj.CJX_ERR_INPUT
    assert n.value == 0
    n = ctypes.c_size_t(7)
    assert L.cjx_chunk_lines_device(None, None, 0, 0, 1, None, None, 0,
                                    ctypes.byref(n)) == cj.CJX_ERR_INPUT
    assert n.value == 0
    assert L.cjx_load_device(None, b"/nonexistent", None, 0, ctypes.byref(n)) == cj.CJX_ERR_INPUT
    # chunk rules: mode 0 needs 1 .. 2^32 chunks, modes other than 0 and 1 are rejected
    text = (ctypes.c_ubyte * 2)(ord("a"), 10)
    for mode, p in ((0, 0), (0, (1 << 32) + 1), (2, 5), (-1, 5)):
        assert L.cjx_chunk_lines(text, 2, mode, p, None, None, 0,
                                 ctypes.byref(n)) == cj.CJX_ERR_INPUT, (mode, p)


def test_chunk_rules_against_oracle_edge_cases():
    """The host chunker (whose rule functions the device chunker shares)
    against the oracle's restatement of load_file.cu:83-103 / :153-176: lines
    longer than the size limit (empty chunks), blank lines, unterminated last
    lines, more chunks than lines."""
    rng = np.random.default_rng(3)
    cases = [b"", b"\n", b"\n\n\n", b"abc", b"abc\n", b"a\nbb\nccc\ndddd\n",
             b"x" * 50 + b"\n" + b"y" * 3 + b"\n" + b"z" * 70]
    for _ in range(40):
        lines = [b"q" * int(k) for k in rng.integers(0, 40, rng.integers(1, 30))]
        d = b"\n".join(lines) + (b"\n" if rng.random() < 0.5 else b"")
        cases.append(d)
    for d in cases:
        for cc in (1, 2, 3, 5, 40):
            offs, sizes = cj.chunk_lines(d, 0, cc)
            assert list(zip(offs, sizes)) == oracle.chunks_by_count(d, cc), (d, cc)
        for b in (1, 2, 5, 10, 33, 64, 1000):
            offs, sizes = cj.chunk_lines(d, 1, b)
            assert list(zip(offs, sizes)) == oracle.chunks_by_bytes(d, b), (d, b)


def test_server_rejects_bad_arguments_without_device():
    h = ctypes.c_void_p()
    L = cj.lib()
    assert L.cjx_server_create(0, 1 << 30, 0, 64, ctypes.byref(h)) == cj.CJX_ERR_INPUT
    assert L.cjx_server_create(0, 0, 2, 64, ctypes.byref(h)) == cj.CJX_ERR_INPUT
    assert L.cjx_server_create(0, 1 << 30, 2, 0, ctypes.byref(h)) == cj.CJX_ERR_INPUT
    assert not h.value
    t = ctypes.c_uint64()
    assert L.cjx_server_submit(None, None, 0, 0, ctypes.byref(t)) == cj.CJX_ERR_INPUT
    n = ctypes.c_size_t()
    assert L.cjx_server_wait(None, None, 0, ctypes.byref(n), 0) == cj.CJX_ERR_INPUT
    L.cjx_server_destroy(None)


def test_context_without_gpu_fails_loudly():
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    assert cj.lib().cjx_context_create(0, ctypes.byref(h)) == cj.CJX_ERR_DEVICE
    with pytest.raises(RuntimeError):
        cj.Context(0)
    with pytest.raises(RuntimeError):
        cj.Server(0, 1 << 30, 2, 64)


def test_mask_transpose():
    """The tokenizer's 8x8 bit transpose (host run of the same function):
    word K, byte b, bit d must be bit K of byte b of input word d."""
    f = cj.lib().cjx_debug_transpose8x8
    f.argtypes = [ctypes.POINTER(ctypes.c_uint32)]
    rng = np.random.default_rng(5)
    cases = [rng.integers(0, 2 ** 32, 8, dtype=np.uint64) for _ in range(200)]
    cases += [np.array([1 << (8 * (d % 4) + d) for d in range(8)], np.uint64),
              np.full(8, 0xFFFFFFFF, np.uint64), np.zeros(8, np.uint64)]
    for c in cases:
        t = (ctypes.c_uint32 * 8)(*[int(x) for x in c])
        f(t)
        for K in range(8):
            want = 0
            for b in range(4):
                for d in range(8):
                    want |= ((int(c[d]) >> (8 * b + K)) & 1) << (8 * b + d)
            assert t[K] == want, (K, hex(t[K]), hex(want))


def test_byte_class_tables():
    """The tokenizer's three-lookup byte classifier (host model of the same
    tables and of v_perm) against the class definitions, for all 256 bytes."""
    f = cj.lib().cjx_debug_byte_class
    f.argtypes = [ctypes.c_uint32, ctypes.c_int]
    f.restype = ctypes.c_uint32
    for jsonl in (0, 1):
        for b in range(256):
            want = 0
            if b == 0x3A or (jsonl and b == 0x0A):
                want |= 0x01
            if 0x80 <= b <= 0xBF:
                want |= 0x02
            if b in (0x5B, 0x5D, 0x7B, 0x7D):
                want |= 0x04
            if b == 0x2C:
                want |= 0x08
            if b == 0x5C:
                want |= 0x10
            if b >= 0xE0:
                want |= 0x20
            if b >= 0xF0:
                want |= 0x40
            if b == 0x22:
                want |= 0x80
            assert f(b, jsonl) == want, (jsonl, hex(b), hex(f(b, jsonl)), hex(want))


def test_noalloc_entry_rejects_before_launching():
    """cjx_parse_device_noalloc checks its arguments on the host before any
    HIP call: missing pointers or an unaligned input are CJX_ERR_INPUT, flags
    other than the UTF-8 and depth flags are CJX_ERR_INPUT, and an input
    above the stated maximum is CJX_ERR_TOO_LARGE."""
    L = cj.lib()
    mx = cj.noalloc_max_bytes()
    assert mx == 1 << 20
    fake = 1 << 40  # never dereferenced: every case below returns first
    f = cj.parse_device_noalloc
    assert f(0, 100, fake, fake, 100, fake) == cj.CJX_ERR_INPUT
    assert f(fake, 0, fake, fake, 100, fake) == cj.CJX_ERR_INPUT
    assert f(fake, 100, fake, fake, 100, 0) == cj.CJX_ERR_INPUT
    assert f(fake + 4, 100, fake, fake, 100, fake) == cj.CJX_ERR_INPUT
    for flags in (cj.CJX_FLAG_GRAPH, cj.CJX_FLAG_JSONL, cj.CJX_FLAG_RECORDS, 0x100):
        assert f(fake, 100, fake, fake, 100, fake, flags) == cj.CJX_ERR_INPUT, flags
    assert f(fake, mx + 1, fake, fake, mx + 1, fake) == cj.CJX_ERR_TOO_LARGE
    assert "maximum" in cj.strerror(cj.CJX_ERR_TOO_LARGE)
    assert L.cjx_noalloc_max_bytes() == mx


def test_chunk_lines_thread_fallback():
    """The threaded line scan of cjx_chunk_lines, forced to find that only
    one or three of its threads can be created: it joins the threads it
    started, scans on the calling thread instead, and gives the loader rule's
    chunks; nothing escapes into the C ABI."""
    f = cj.lib().cjx_debug_fail_threads
    f.argtypes = [ctypes.c_int]
    f.restype = None
    rng = np.random.default_rng(9)
    lens = rng.integers(0, 3000, 50_000)
    a = np.full(int(lens.sum()) + lens.size, ord("q"), np.uint8)
    a[np.cumsum(lens + 1) - 1] = 10
    assert a.size >= 64 << 20
    want = oracle.chunks_by_bytes(a, 1 << 20)
    try:
        for k in (1, 3):
            f(k)
            offs, sizes = cj.chunk_lines(a, 1, 1 << 20)
            assert list(zip(offs, sizes)) == want, k
    finally:
        f(0)
    offs, sizes = cj.chunk_lines(a, 0, 333)
    assert list(zip(offs, sizes)) == oracle.chunks_by_count(a, 333)


def test_strict_utf8_lead_table():
    """The strict UTF-8 check's lead table (host model of the same v_perm
    lookups) against its class definitions, for all 256 bytes."""
    f = cj.lib().cjx_debug_lead_class
    f.argtypes = [ctypes.c_uint32]
    f.restype = ctypes.c_uint32
    for b in range(256):
        want = ((b in (0xC0, 0xC1)) | ((0xF5 <= b <= 0xF7) << 1) | ((b >= 0xF8) << 2) |
                ((b == 0xF4) << 3) | ((b == 0xF0) << 4) | ((b == 0xE0) << 5) |
                ((b == 0xED) << 6) | ((b >= 0xC0) << 7))
        assert f(b) == want, (hex(b), hex(f(b)), hex(want))


def test_strict_utf8_word_steps_against_oracle():
    """Host run of the strict check's word-level steps (transposed words,
    both classifications, the carry between 16-byte lanes) against the
    oracle's RFC 3629 validation: every short sequence at every offset of a
    48-byte window, and random byte strings."""
    f = cj.lib().cjx_debug_strict_utf8
    f.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    f.restype = ctypes.c_int
    seqs = [b"\xc3\xa9", b"\xe2\x82\xac", b"\xf0\x9f\x98\x80", b"\xf4\x8f\xbf\xbf", b"\xc3",
            b"\xe2\x82", b"\xf0\x9f\x98", b"\xc0\xaf", b"\xc1\xbf", b"\xe0\x80\xaf",
            b"\xe0\xa0\x80", b"\xed\xa0\x80", b"\xed\x9f\xbf", b"\xf0\x80\x80\x80",
            b"\xf0\x90\x80\x80", b"\xf4\x90\x80\x80", b"\xf5\x80", b"\xff", b"\x80", b"\xbf\xbf",
            b"\xc3\xa9\xa9", b"\xe2\x82\xac\x80", b"\xdf\xbf", b"\xef\xbf\xbf", b"\xf8\x88\x80\x80"]
    for sq in seqs:
        for off in range(0, 48 - len(sq)):
            for tail in (0, 1, 5):
                d = b"a" * off + sq + b"b" * tail
                assert f(d, len(d)) == oracle.utf8_check(d, oracle.FLAG_UTF8_STRICT), (sq, off, tail)
    rng = np.random.default_rng(17)
    pool = np.array([0x41, 0x80, 0x8F, 0x90, 0x9F, 0xA0, 0xBF, 0xC0, 0xC2, 0xDF, 0xE0, 0xE1,
                     0xED, 0xEF, 0xF0, 0xF1, 0xF4, 0xF5, 0xFF], np.uint8)
    for _ in range(20000):
        d = rng.choice(pool, rng.integers(1, 70)).tobytes()
        assert f(d, len(d)) == oracle.utf8_check(d, oracle.FLAG_UTF8_STRICT), d
    for _ in range(300):  # mostly valid text with some invalid bytes
        s = "".join(chr(int(c)) for c in rng.choice([0x41, 0xE9, 0x20AC, 0x1F600, 0x7F], 300))
        d = bytearray(s.encode())
        if rng.random() < 0.5:
            d[int(rng.integers(0, len(d)))] = int(rng.choice(pool))
        d = bytes(d)
        assert f(d, len(d)) == oracle.utf8_check(d, oracle.FLAG_UTF8_STRICT), d


def _lane_class_want(b, jsonl):
    """Class bits of one byte: bs, q, br, op, b1, b5 (as the tokenizer's masks)."""
    br = b in (0x5B, 0x5D, 0x7B, 0x7D)
    nl = jsonl and b == 0x0A
    return (b == 0x5C, b == 0x22, br, br or b in (0x3A, 0x2C) or nl,
            br and b in (0x5B, 0x7B), (br and b in (0x7B, 0x7D)) or nl)


def test_lane_classes_from_bit_planes():
    """The tokenizer's character classes from bit planes (host run of the
    same byte transposes, bit transpose and plane products): every byte value
    at every one of a lane's 32 positions, both modes."""
    f = cj.lib().cjx_debug_lane_classes
    f.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_uint32)]
    f.restype = None
    rng = np.random.default_rng(21)
    lanes = [bytes((v + 37 * i) & 0xFF for i in range(32)) for v in range(256)]
    lanes += [rng.integers(0, 256, 32, dtype=np.uint8).tobytes() for _ in range(300)]
    lanes += [b'{"a":[1,2],"b\\\\":"x"}\n[]{}::,,\\"\n\n\x0a\x22\x5c'[:32].ljust(32, b" ")]
    out = (ctypes.c_uint32 * 6)()
    for jsonl in (0, 1):
        for lane in lanes:
            assert len(lane) == 32
            f(lane, jsonl, out)
            for i, b in enumerate(lane):
                want = _lane_class_want(b, jsonl)
                got = tuple(bool((out[k] >> i) & 1) for k in range(6))
                assert got == want, (jsonl, hex(b), i, got, want)


def test_weak_utf8_bit_planes_against_oracle():
    """Host run of the tokenizer's reference-rule UTF-8 check on bit planes
    (lane chain, wave starts carried from the four bytes before them, word
    validity at the end) against the oracle's per-word restatement of the
    reference's checkUTF8: every short sequence at every offset of a 96-byte
    window for three wave sizes, and random byte strings."""
    f = cj.lib().cjx_debug_weak_utf8
    f.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_size_t]
    f.restype = ctypes.c_int
    seqs = [b"\xc3\xa9", b"\xe2\x82\xac", b"\xf0\x9f\x98\x80", b"\xc3", b"\xe2\x82",
            b"\xf0\x9f\x98", b"\xc0\xaf", b"\xe0\x80\xaf", b"\xed\xa0\x80", b"\xf5\x80", b"\xff",
            b"\x80", b"\xbf\xbf", b"\xc3\xa9\xa9", b"\xe2\x82\xac\x80", b"\xf8\x88\x80\x80",
            b"\xe2", b"\xf0", b"\xf0\x9f"]
    n_err = 0
    for sq in seqs:
        for off in range(0, 96 - len(sq)):
            for tail in (0, 1, 3, 6):
                d = b"a" * off + sq + b"b" * tail
                want = oracle.utf8_check(d, 0)
                n_err += want
                for wb in (32, 64, 2048):
                    assert f(d, len(d), wb) == want, (sq, off, tail, wb)
    assert n_err > 0
    rng = np.random.default_rng(23)
    pool = np.array([0x41, 0x22, 0x80, 0x8F, 0xA0, 0xBF, 0xC0, 0xC2, 0xDF, 0xE0, 0xE1, 0xED,
                     0xEF, 0xF0, 0xF4, 0xF5, 0xFF], np.uint8)
    for _ in range(6000):
        d = rng.choice(pool, rng.integers(1, 200)).tobytes()
        want = oracle.utf8_check(d, 0)
        for wb in (32, 128):
            assert f(d, len(d), wb) == want, (d, wb)
    for _ in range(200):  # valid text with an occasional stray byte
        s = "".join(chr(int(c)) for c in rng.choice([0x41, 0xE9, 0x20AC, 0x1F600, 0x7F], 400))
        d = bytearray(s.encode())
        if rng.random() < 0.5:
            d[int(rng.integers(0, len(d)))] = int(rng.choice(pool))
        d = bytes(d)
        want = oracle.utf8_check(d, 0)
        for wb in (32, 256, 2048):
            assert f(d, len(d), wb) == want, (d, wb)
