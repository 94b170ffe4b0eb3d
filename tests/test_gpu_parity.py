"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Integer/byte/index work, so the bar is bit-exact: same error code, same N,
and the same 2(N+2)-cell host buffer (structural row, sentinels, pair row
with -1 in the cells the reference never writes)."""
import random

import numpy as np
import pytest

import cujson_amd as cj
import oracle
from cujson_amd import gen

pytestmark = pytest.mark.gpu

ALPHABET = [ord(c) for c in '{}[]:,"\\a \n'] + [0xC3, 0xA9, 0xE2, 0x82, 0xAC, 0xF0]


@pytest.fixture(scope="module")
def ctx():
    c = cj.Context(0)
    yield c
    c.close()


def check_standard(ctx, data, flags=0, label=""):
    data = bytes(data) if not isinstance(data, np.ndarray) else data
    g = cj.parse_standard_json(data, flags, ctx)
    o = oracle.parse_standard(data, flags)
    assert g.error == o.code, (label, g.error, o.code)
    if o.code == oracle.ERR_INPUT:
        return g
    assert g.N == o.N, (label, g.N, o.N)
    if not np.array_equal(g.buf, o.buf):
        bad = np.flatnonzero(g.buf != o.buf)
        raise AssertionError(f"{label}: {bad.size} cells differ, first {bad[:8]} "
                             f"gpu={g.buf[bad[:8]]} oracle={o.buf[bad[:8]]} N={o.N}")
    return g


def check_lines(ctx, data, chunks, flags=0, label=""):
    offs = [o for o, _ in chunks]
    sizes = [s for _, s in chunks]
    inp = cj.cuJSONLinesInput(np.frombuffer(bytes(data), np.uint8), len(chunks), len(data), offs,
                              sizes)
    g = cj.parse_json_lines(inp, flags, ctx)
    o = oracle.parse_lines(data, chunks, flags)
    assert g.error == o.code, (label, g.error, o.code)
    if o.code != oracle.OK:
        return g
    assert g.N == o.N
    assert np.array_equal(g.buf, o.buf), label
    assert g.resultSizes == o.result_sizes
    assert g.resultSizesPrefix == o.result_sizes_prefix
    return g


def test_twitter_large(ctx, twitter_large):
    g = check_standard(ctx, twitter_large, label="twitter_large")
    assert g.N == 22388
    assert g.pair_pos[0] == g.N + 1


def test_config1(ctx):
    check_standard(ctx, gen.config1(), label="config1")


@pytest.mark.parametrize("cc", [1, 2, 4, 8, 51])
def test_twitter_small_records_chunked(ctx, twitter_small, cc):
    check_lines(ctx, twitter_small, oracle.chunks_by_count(twitter_small, cc), label=f"cc={cc}")


def test_small_sizes_and_tile_edges(ctx):
    # lane (32 B), wave round (2 KiB), wave (16 KiB) and tile (64 KiB) edges
    base = bytes(gen.config1()[:270000])
    for n in [1, 2, 3, 4, 5, 15, 16, 17, 31, 32, 33, 63, 64, 65, 1023, 1024, 1025, 2047, 2048,
              2049, 4095, 4096, 4097, 16383, 16384, 16385, 32767, 32768, 32769, 65535, 65536,
              65537, 131071, 131072, 131073, 131072 + 7, 262143, 262144, 262145]:
        check_standard(ctx, base[:n], label=f"prefix n={n}")


def test_empty_and_null(ctx):
    g = cj.parse_standard_json(b"", 0, ctx)
    assert g.error == cj.CJX_ERR_INPUT and g.structural is None


@pytest.mark.parametrize("boundary", [16, 1024, 2048, 4096, 16384, 32768, 65536, 131072])
def test_backslash_runs_across_boundaries(ctx, boundary):
    for k in range(1, 40):
        for tail in (b'"]', b'a"]', b'\\"]'):
            for shift in (-2, -1, 0, 1):
                pre = b'["' + b"x" * (boundary - 2 - k + shift)
                check_standard(ctx, pre + b"\\" * k + tail + b',{"a":[1]}]',
                               label=f"b={boundary} k={k} s={shift} tail={tail}")


def test_all_backslash_tiles(ctx):
    for k in (16384 - 1, 16384, 16384 + 1, 2 * 16384, 2 * 16384 + 1, 3 * 16384 + 5, 65535,
              65536, 65537, 2 * 65536 + 1, 3 * 65536 + 5):
        for head in (b'["', b'[', b'"'):
            check_standard(ctx, head + b"\\" * k + b'"]",[1]]', label=f"allbs k={k} head={head}")


def test_strings_across_tiles(ctx):
    for L in (16383, 16384, 16385, 50000, 65535, 65536, 65537, 200000):
        doc = b'[{"a":"' + (b"{[:,]}" * (L // 6 + 1))[:L] + b'"},[1,2,{"b":[]}]]'
        check_standard(ctx, doc, label=f"long string L={L}")


def test_random_byte_soup(ctx):
    rng = random.Random(2024)
    for it in range(40):
        n = rng.choice([100, 5000, 20000, 70000, 200000])
        data = bytes(rng.choice(ALPHABET) for _ in range(n))
        check_standard(ctx, data, label=f"soup it={it} n={n}")
        check_standard(ctx, data, oracle.FLAG_DEPTH_U16, label=f"soup u16 it={it}")


def test_deep_nesting(ctx):
    for d in (255, 256, 257, 300, 512, 1000):
        doc = b"[" * d + b"1" + b"]" * d
        check_standard(ctx, doc, label=f"depth {d}")
        check_standard(ctx, doc, oracle.FLAG_DEPTH_U16, label=f"depth {d} u16")
    doc = b"".join((b"[" if i % 2 == 0 else b'{"d":') for i in range(600)) + b"0" + b"".join(
        (b"]" if i % 2 == 0 else b"}") for i in reversed(range(600)))
    check_standard(ctx, doc, label="alternating 600")
    check_standard(ctx, doc, oracle.FLAG_DEPTH_U16, label="alternating 600 u16")


def test_faults(ctx):
    base = gen.config1()
    for kind in range(5):
        for seed in range(3):
            d = base.copy()
            assert gen.plant_fault(d, kind, 100 + seed)
            check_standard(ctx, d, label=f"fault {kind} seed {seed}")


def test_utf8_at_word_and_tile_edges(ctx):
    seqs = [b"\xc3\xa9", b"\xe2\x82\xac", b"\xf0\x9f\x98\x80", b"\xc3", b"\xe2\x82", b"\xf0\x9f",
            b"\xc0\xaf", b"\xed\xa0\x80"]
    for edge in (4, 16, 1024, 2048, 4096, 16384, 65536):
        for s in seqs:
            for shift in range(-4, 2):
                pre = b'["' + b"a" * (edge - 2 + shift)
                check_standard(ctx, pre + s + b'bcdefgh"]', label=f"utf8 edge={edge} {s} {shift}")


STRICT_SEQS = [b"\xc3\xa9", b"\xe2\x82\xac", b"\xf0\x9f\x98\x80", b"\xf4\x8f\xbf\xbf",
               b"\xc3", b"\xe2\x82", b"\xf0\x9f\x98", b"\xc0\xaf", b"\xc1\xbf",
               b"\xe0\x80\xaf", b"\xe0\xa0\x80", b"\xed\xa0\x80", b"\xed\x9f\xbf",
               b"\xf0\x80\x80\x80", b"\xf0\x90\x80\x80", b"\xf4\x90\x80\x80", b"\xf5\x80",
               b"\xff", b"\x80", b"\xbf\xbf", b"\xc3\xa9\xa9", b"\xe2\x82\xac\x80"]


def test_strict_utf8_at_edges(ctx):
    for edge in (4, 64, 2048, 16384, 65536):
        for sq in STRICT_SEQS:
            for shift in range(-5, 2):
                pre = b'["' + b"a" * (edge - 2 + shift)
                check_standard(ctx, pre + sq + b'bc"]', oracle.FLAG_UTF8_STRICT,
                               label=f"strict edge={edge} {sq} {shift}")


def test_strict_utf8_at_end_and_start(ctx):
    for sq in STRICT_SEQS:
        check_standard(ctx, sq, oracle.FLAG_UTF8_STRICT, label=f"strict alone {sq}")
        check_standard(ctx, b'["x",' + sq, oracle.FLAG_UTF8_STRICT, label=f"strict tail {sq}")
        check_standard(ctx, sq + b'["x"]', oracle.FLAG_UTF8_STRICT, label=f"strict head {sq}")


def test_strict_utf8_random(ctx):
    rng = random.Random(99)
    pool = [bytes([b]) for b in range(0x20, 0x7F, 7)] + STRICT_SEQS
    for it in range(60):
        n = rng.choice([10, 300, 5000, 70000])
        data = b"".join(rng.choice(pool) for _ in range(n))
        check_standard(ctx, data, oracle.FLAG_UTF8_STRICT, label=f"strict soup {it}")
    d = gen.config1()
    check_standard(ctx, d, oracle.FLAG_UTF8_STRICT, label="strict config1")
    for kind in range(5):
        f = d.copy()
        assert gen.plant_fault(f, kind, 7)
        check_standard(ctx, f, oracle.FLAG_UTF8_STRICT, label=f"strict fault {kind}")


def test_config2_shape_64mb(ctx):
    d = gen.config2(target_bytes=64 << 20)
    check_standard(ctx, d, label="config2 64MB")


@pytest.fixture(scope="module")
def doc96():
    d = gen.config2_parallel(total_bytes=96 << 20)
    assert d.size >= 64 << 20  # large enough for the chunked host copy
    return d


@pytest.mark.parametrize("flags", [0, oracle.FLAG_UTF8_STRICT, oracle.FLAG_DEPTH_U16])
def test_chunked_host_input(ctx, doc96, tmp_path, flags):
    """Host input of at least 64 MiB is copied in parts while the tokenizer
    runs on the parts already in (one tokenizer launch per part): pageable
    input and pinned input (loadJSON) both equal the oracle."""
    o = oracle.parse_standard(doc96, flags)
    for label, inp in (("pageable", doc96), ("pinned", None)):
        if inp is None:
            p = tmp_path / "doc.json"
            p.write_bytes(doc96.tobytes())
            inp = cj.loadJSON(str(p))
        g = cj.parse_standard_json(inp, flags, ctx)
        assert g.error == o.code, (label, g.error, o.code)
        assert g.N == o.N, label
        assert np.array_equal(g.buf, o.buf), label


def test_chunked_host_input_fault(ctx, doc96):
    d = doc96.copy()
    assert gen.plant_fault(d, 3, 99)
    check_standard(ctx, d, label="chunked with a fault")


@pytest.mark.parametrize("shift", [0, 3, 9, 15])
def test_lines_host_span_placement(ctx, shift):
    """More than 16 chunks from host memory: the source span goes in as one
    copy and the device places every chunk at its aligned offset (source at
    an aligned and at an odd address); chunks far apart in their buffer (a
    sparse span) take the staging copy instead.  All equal the oracle."""
    d = bytes(gen.lines(3000))
    chunks = oracle.chunks_by_count(d, 40)
    o = oracle.parse_lines(d, chunks)
    assert len(chunks) > 16 and o.code == oracle.OK
    buf = np.zeros(len(d) + 16, np.uint8)
    buf[shift:shift + len(d)] = np.frombuffer(d, np.uint8)
    view = buf[shift:shift + len(d)]
    inp = cj.cuJSONLinesInput(view, len(chunks), len(d), [c[0] for c in chunks],
                              [c[1] for c in chunks])
    g = cj.parse_json_lines(inp, 0, ctx)
    assert g.error == 0 and g.N == o.N and np.array_equal(g.buf, o.buf), shift
    # sparse: every chunk followed by a 64 KiB gap in the caller's buffer
    gap = 64 << 10
    sparse = np.zeros(len(d) + gap * len(chunks), np.uint8)
    offs = []
    for i, (a, z) in enumerate(chunks):
        offs.append(a + gap * i)
        sparse[offs[-1]:offs[-1] + z] = np.frombuffer(d[a:a + z], np.uint8)
    r = cj.parse_json_lines(cj.cuJSONLinesInput(sparse, len(chunks), sparse.size, offs,
                                                [c[1] for c in chunks]), 0, ctx)
    assert r.error == 0 and r.N == o.N
    # positions count bytes of the chunks only, so the layout is the same
    assert np.array_equal(r.buf, o.buf)


def test_chunked_host_sparse_head(ctx):
    """The structural rows of a chunked host parse are copied out part by
    part into a buffer sized from the first part's density.  Here the first
    part lies inside one 20 MiB string, so the rest outgrows that buffer and
    the result is copied after the parse instead: same output."""
    body = b'{"key":12345},' * ((52 << 20) // 14)
    d = np.frombuffer(b'["' + b"a" * (20 << 20) + b'",' + body + b"0]", np.uint8)
    assert d.size >= 64 << 20
    check_standard(ctx, d, label="sparse head")


def test_chunked_host_dense_rerun(ctx):
    """More than one structural per three bytes overflows the first
    attempt's row region: the parse re-runs with full regions and the rows
    streamed during the first attempt are not used."""
    d = np.frombuffer(b"[" + b"[]," * ((66 << 20) // 3) + b"0]", np.uint8)
    check_standard(ctx, d, label="dense re-run")


def test_jsonl_generated_records(ctx):
    d = gen.lines(3000)
    data = bytes(d)
    for cc in (1, 3, 16):
        check_lines(ctx, data, oracle.chunks_by_count(data, cc), label=f"lines cc={cc}")
    check_lines(ctx, data, oracle.chunks_by_bytes(data, 1 << 16), label="lines 64KiB")


def nested_doc(rng, n_brackets, max_depth, mismatches=0):
    """Random well-nested document of about n_brackets brackets whose depth
    stays within max_depth, with `mismatches` closers of the wrong type."""
    out, stack = [b"["], [b"]"]
    made = 1
    while made < n_brackets or len(stack) > 0:
        if made >= n_brackets:
            out.append(stack.pop())
            continue
        up = len(stack) < max_depth and (not stack or rng.random() < 0.5)
        if up:
            kind = rng.random() < 0.5
            out.append(b'{"k":' if kind else b"[1,")
            stack.append(b"}" if kind else b"]")
        else:
            out.append(stack.pop() + (b"," if stack else b""))
            if stack and out[-1].endswith(b","):
                out.append(b"0")
        made += 1
    doc = bytearray(b"".join(out))
    closers = [i for i, c in enumerate(doc) if c in (0x5D, 0x7D)]
    for i in rng.sample(closers, mismatches):
        doc[i] = 0x7D if doc[i] == 0x5D else 0x5D
    return bytes(doc)


@pytest.mark.parametrize("max_depth", [3, 64, 255, 256, 257])
def test_nested_depth_limits(ctx, max_depth):
    rng = random.Random(max_depth)
    for n, bad in ((3000, 0), (200000, 0), (200000, 1), (50000, 3)):
        doc = nested_doc(rng, n, max_depth, bad)
        check_standard(ctx, doc, label=f"nested d<={max_depth} n={n} bad={bad}")
        if max_depth >= 255:
            check_standard(ctx, doc, oracle.FLAG_DEPTH_U16, label=f"u16 d<={max_depth} n={n}")


@pytest.mark.parametrize("max_depth", [2, 40, 255, 256, 257])
def test_nested_wave_units(ctx, max_depth):
    """Units of 1 to 1100 brackets (one wave matches those of at most 1024,
    in groups of 64; deeper or larger ones go to the sort or chunk path):
    random well-nested documents at the depth limits, with and without
    planted type mismatches, alone and in a batch."""
    rng = random.Random(100 + max_depth)
    docs = []
    for n in (1, 2, 63, 64, 65, 100, 127, 128, 129, 500, 1000, 1023, 1024, 1025, 1100):
        for bad in (0, 1):
            d = nested_doc(rng, n, max_depth, bad if n > 2 else 0)
            docs.append(d)
            check_standard(ctx, d, label=f"wave unit n={n} d<={max_depth} bad={bad}")
            if max_depth >= 255:
                check_standard(ctx, d, oracle.FLAG_DEPTH_U16, label=f"wave u16 n={n}")
    offs, pos = [], 0
    for d in docs:
        offs.append(pos)
        pos += len(d)
    data = np.frombuffer(b"".join(docs), np.uint8)
    r = cj.parse_batch(data, offs, [len(d) for d in docs], 0, ctx)
    assert r.rc == 0
    for i, d in enumerate(docs):
        want = oracle.parse_standard(d)
        err, buf = r.doc(i)
        assert err == want.code and np.array_equal(buf, want.buf), (max_depth, i)


def test_many_match_chunks(ctx):
    """More than 4096 chunks of 4096 brackets under one top-level array."""
    rng = random.Random(7)
    parts = [b"[[[]]]", b'{"a":[{}]}', b"[{},[1]]", b'[[{"b":[[]]}]]']
    body = b",".join(rng.choice(parts) for _ in range(3_000_000))
    doc = b'{"top":[' + body + b"]}"
    check_standard(ctx, doc, label="many chunks")


def test_jsonl_bad_chunks(ctx, twitter_small):
    ch = [(0, 0)] + oracle.chunks_by_count(twitter_small, 2)
    g = cj.parse_json_lines(cj.cuJSONLinesInput(np.frombuffer(twitter_small, np.uint8), 3,
                                                len(twitter_small), [c[0] for c in ch],
                                                [c[1] for c in ch]), 0, ctx)
    assert g.error == cj.CJX_ERR_INPUT and g.structural is None


def test_device_resident_entry(ctx):
    from cujson_amd.hip import DeviceBuffer
    d = gen.config1()
    n = d.size
    dev = DeviceBuffer(((n + 15) // 16) * 16)
    dev.upload(d)
    cap = n
    st = DeviceBuffer(cap * 4)
    pr = DeviceBuffer(cap * 4)
    assert ctx.parse_device(dev.ptr, n, st.ptr, pr.ptr, cap) == 0
    rc, N, M, over = ctx.wait()
    o = oracle.parse_standard(d)
    assert rc == o.code and N == o.N and not over
    assert np.array_equal(st.download(np.int32, N), o.structural[1:N + 1])
    assert np.array_equal(pr.download(np.int32, N), o.pair_pos[1:N + 1])
    # capacity overflow is reported, never written past
    small = 1000
    st2 = DeviceBuffer((small + 64) * 4)
    pr2 = DeviceBuffer((small + 64) * 4)
    st2.fill(7)
    pr2.fill(7)
    assert ctx.parse_device(dev.ptr, n, st2.ptr, pr2.ptr, small) == 0
    rc, N, M, over = ctx.wait()
    assert over and rc == cj.CJX_ERR_CAPACITY
    assert np.all(st2.download(np.uint8, 64 * 4, small * 4) == 7)
    assert np.all(pr2.download(np.uint8, 64 * 4, small * 4) == 7)


def test_timings(ctx):
    ctx.set_timing(True)
    try:
        g = cj.parse_standard_json(gen.config1(), 0, ctx)
        assert g.error == 0
        tok, p3, tot = ctx.timings()
        assert tok > 0 and p3 > 0 and tot >= tok
    finally:
        ctx.set_timing(False)
