"""Typed value tables (cjx_values, cjx_values_device) against a gap-by-gap
restatement in Python: the structural positions come from the oracle's
parse, each gap's trimmed text is classified and decoded with json.loads.
Every entry must match: type, text span, flags, the int64 / binary64 value
bit for bit, the unescaped length, and the unescaped bytes."""
import json
import math
import random
import struct
from decimal import Decimal, getcontext

import numpy as np
import pytest

import cujson_amd as cj
import oracle
from cujson_amd import gen
from cujson_amd.hip import DeviceBuffer

pytestmark = pytest.mark.gpu

WS = b" \t\n\r"


@pytest.fixture(scope="module")
def ctx():
    c = cj.Context(0)
    yield c
    c.close()


def bits(x: float) -> int:
    return struct.unpack("<q", struct.pack("<d", x))[0]


def expected_entry(data: bytes, a: int, b: int, nxt):
    """(type, text_off, text_len, flags, v, unescaped bytes or None)."""
    while a < b and data[a] in WS:
        a += 1
    while b > a and data[b - 1] in WS:
        b -= 1
    text = data[a:b]
    if not text:
        return cj.CJX_VALUE_NONE, a, 0, 0, 0, None
    if text[:1] == b'"':
        off, ln = a + 1, max(len(text) - 2, 0)
        try:
            s = json.loads(text) if len(text) >= 2 and text[-1:] == b'"' else None
        except ValueError:
            s = None
        if not isinstance(s, str):
            return cj.CJX_VALUE_INVALID, off, ln, 0, 0, None
        u = s.encode("utf-8", "surrogatepass")
        fl = ((cj.CJX_VALUE_ESCAPED if b"\\" in text[1:-1] else 0) |
              (cj.CJX_VALUE_UNPAIRED if any(0xD800 <= ord(c) <= 0xDFFF for c in s) else 0))
        t = cj.CJX_VALUE_KEY if nxt == ord(":") else cj.CJX_VALUE_STRING
        return t, off, ln, fl, len(u), u
    lit = {b"true": cj.CJX_VALUE_TRUE, b"false": cj.CJX_VALUE_FALSE, b"null": cj.CJX_VALUE_NULL}
    if text in lit:
        return lit[text], a, len(text), 0, 0, None
    s = text.decode("latin-1")
    try:
        o = None if s in ("NaN", "Infinity", "-Infinity") else json.loads(s)
    except ValueError:
        o = None
    if o is None or isinstance(o, (bool, str, list, dict)):
        return cj.CJX_VALUE_INVALID, a, len(text), 0, 0, None
    if isinstance(o, int):
        if -2 ** 63 <= o < 2 ** 63:
            return cj.CJX_VALUE_INT, a, len(text), 0, o, None
        return cj.CJX_VALUE_FLOAT, a, len(text), cj.CJX_VALUE_BIG_INT, bits(float(o)), None
    return cj.CJX_VALUE_FLOAT, a, len(text), 0, bits(o), None


def check_table(data: bytes, st, table, strings, label):
    """st: 1-based positions of structurals 1..N."""
    N, n = len(st), len(data)
    assert table.size == N + 1, label
    bad = []
    for h in range(N + 1):
        a = 0 if h == 0 else int(st[h - 1])
        b = n if h == N else int(st[h]) - 1
        nxt = data[b] if h < N else None
        t, off, ln, fl, v, u = expected_entry(data, a, b, nxt)
        e = table[h]
        got = (int(e["type"]), int(e["text_off"]), int(e["text_len"]))
        ok = got == (t, off, ln) and int(e["row"]) == h
        if ok and t not in (cj.CJX_VALUE_INVALID, cj.CJX_VALUE_NONE, cj.CJX_VALUE_TRUE,
                            cj.CJX_VALUE_FALSE, cj.CJX_VALUE_NULL):
            ok = int(e["flags"]) == fl and int(e["v"]) == v
            if ok and u is not None and strings is not None:
                ok = strings[off:off + len(u)] == u
        if not ok:
            bad.append((h, data[a:b][:80], tuple(e.tolist()), (t, off, ln, fl, v)))
            if len(bad) >= 5:
                break
    assert not bad, (label, bad)


def edge_document() -> bytes:
    getcontext().prec = 100
    mid = (Decimal(1.0) + Decimal(math.nextafter(1.0, 2.0))) / 2
    return (b'{"k1": 0, "k2": -0.0, "k3": 1e400, "k4": 9007199254740993, '
            b'"k5": "esc\\"aped\\n\\u00e9\\ud83d\\ude00", "k6": "lone\\ud800", "k7": tru, '
            b'"k8": 01, "k9": NaN, "k10": "bad\\x", "k11": "ctl\x01", '
            b'"k12": 18446744073709551616, "k13": [true, false, null, 1.5e-320, '
            b'2.2250738585072011e-308, ' + format(mid, "e").encode() + b'], "k14": "", '
            b'"k15": "a" "b", "k16": 4503599627370496.5, "k17":   \t 42 \r\n , '
            b'"k18": {"nested": "v", "e": {}}, "k19": [ ], "\\u006b20": -1.5E-7, '
            b'"k21": 123456789012345678901234567890.5e-10, "k22": "caf\xc3\xa9"}')


def numbers_document(k: int = 30000) -> bytes:
    rng = random.Random(21)
    getcontext().prec = 1200
    out = []
    while len(out) < k:
        x = struct.unpack("<d", struct.pack("<Q", rng.getrandbits(64)))[0]
        if math.isnan(x) or math.isinf(x):
            continue
        r = rng.random()
        if r < 0.5:
            out.append(repr(x))
        elif r < 0.8:
            out.append("%.17g" % x)
        else:  # a halfway point (exact step) or a near miss
            y = math.nextafter(x, math.inf)
            if math.isinf(y):
                continue
            mid = (Decimal(x) + Decimal(y)) / 2
            out.append(format(mid, "e"))
        if rng.random() < 0.2:
            out.append(str(rng.randint(-2 ** 70, 2 ** 70)))
    return ("[" + ", ".join(out) + "]").encode()


def run_standard(ctx, data: bytes, label: str):
    r = cj.parse_standard_json(cj.cuJSONInput(np.frombuffer(data, np.uint8), len(data)), 0, ctx)
    o = oracle.parse_standard(data)
    assert r.error == 0 and o.code == 0, label
    table, strings = ctx.values(data, r)
    check_table(data, o.structural[1:o.N + 1], table, strings, label)
    return table


def test_twitter_large(ctx, twitter_large):
    run_standard(ctx, bytes(twitter_large), "twitter large")


def test_generated_document(ctx):
    run_standard(ctx, bytes(gen.config1()), "config1")


def test_edge_document(ctx):
    table = run_standard(ctx, edge_document(), "edge")
    types = set(int(t) for t in table["type"])
    assert {cj.CJX_VALUE_INVALID, cj.CJX_VALUE_KEY, cj.CJX_VALUE_STRING, cj.CJX_VALUE_INT,
            cj.CJX_VALUE_FLOAT, cj.CJX_VALUE_TRUE, cj.CJX_VALUE_FALSE,
            cj.CJX_VALUE_NULL} <= types


def test_json_lines(ctx, twitter_small):
    data = bytes(twitter_small)
    chunks = oracle.chunks_by_count(data, 3)
    inp = cj.cuJSONLinesInput(np.frombuffer(data, np.uint8), len(chunks), len(data),
                              [c[0] for c in chunks], [c[1] for c in chunks])
    r = cj.parse_json_lines(inp, 0, ctx)
    o = oracle.parse_lines(data, chunks)
    assert r.error == 0 and o.code == 0
    table, strings = ctx.values(data, r)
    check_table(data, o.structural[1:o.N + 1], table, strings, "lines")


def test_device_entry_and_exact_step(ctx):
    """Many numbers, a third of them halfway points: the exact step runs on
    the device, and the device entry equals the host one."""
    data = numbers_document()
    table = run_standard(ctx, data, "numbers")
    n = len(data)
    dev = DeviceBuffer(n + 64)
    dev.upload(np.frombuffer(data, np.uint8))
    cap = n
    st, pr = DeviceBuffer(cap * 4), DeviceBuffer(cap * 4)
    assert ctx.parse_device(dev.ptr, n, st.ptr, pr.ptr, cap) == 0
    rc, N, M, over = ctx.wait()
    assert rc == 0 and not over
    vals = DeviceBuffer((N + 1) * cj.VALUE_DTYPE.itemsize)
    strs = DeviceBuffer(n)
    rc, n_exact = ctx.values_device(dev.ptr, n, st.ptr, N, vals.ptr, strs.ptr)
    assert rc == 0 and n_exact > 1000
    got = np.frombuffer(vals.download(np.uint8, (N + 1) * cj.VALUE_DTYPE.itemsize).tobytes(),
                        dtype=cj.VALUE_DTYPE)
    assert np.array_equal(got, table)


def test_long_strings_and_unaligned_text(ctx):
    """Blocks whose text span exceeds the LDS stage (strings of 20-60 KB)
    read the text in place; device text at an odd address is never staged.
    Both must give the same table and strings as the host form."""
    rng = random.Random(22)
    pieces = ["ab", "c ", "xyz", "\\\\", "\\n", "\\u00e9", "\u00e9"]
    parts = []
    for i in range(40):
        body = "".join(rng.choice(pieces) for _ in range(rng.choice((3, 8000, 25000))))
        parts.append(f'"k{i}": "{body}", "n{i}": {rng.random() * 1e6!r}, "b{i}": true')
    data = ("{" + ", ".join(parts) + "}").encode()
    table = run_standard(ctx, data, "long strings")
    n = len(data)
    dev = DeviceBuffer(n + 64)
    dev.upload(np.frombuffer(data, np.uint8))
    st, pr = DeviceBuffer(n * 4), DeviceBuffer(n * 4)
    assert ctx.parse_device(dev.ptr, n, st.ptr, pr.ptr, n) == 0
    rc, N, M, over = ctx.wait()
    assert rc == 0 and not over and N + 1 == table.size
    for shift in (0, 1):
        text = DeviceBuffer(n + 64)
        buf = np.zeros(n + shift, np.uint8)
        buf[shift:] = np.frombuffer(data, np.uint8)
        text.upload(buf)
        vals = DeviceBuffer((N + 1) * cj.VALUE_DTYPE.itemsize)
        strs = DeviceBuffer(n)
        rc, _ = ctx.values_device(text.ptr + shift, n, st.ptr, N, vals.ptr, strs.ptr)
        assert rc == 0
        got = np.frombuffer(vals.download(np.uint8, (N + 1) * cj.VALUE_DTYPE.itemsize).tobytes(),
                            dtype=cj.VALUE_DTYPE)
        assert np.array_equal(got, table), shift
        sb = strs.download(np.uint8, n).tobytes()
        for e in table:
            if int(e["type"]) in (cj.CJX_VALUE_KEY, cj.CJX_VALUE_STRING):
                o, k = int(e["text_off"]), int(e["v"])
                want = json.loads(b'"' + data[o:o + int(e["text_len"])] + b'"').encode()
                assert sb[o:o + k] == want, (shift, o)


def test_bad_arguments(ctx):
    rc, _ = ctx.values_device(0, 10, 0, 0, 0)
    assert rc == cj.CJX_ERR_INPUT


def test_rows_out_of_order(ctx):
    """Caller-built rows that are not ascending (a block of 256 gaps with
    rows that go back, and rows past the end of the text): the kernel reads
    the text in place for such blocks and treats each gap as gap_of defines
    it (a gap whose start lies after its end is empty at its end; rows past
    the text are clamped to it), never the LDS stage of another span."""
    data = bytes(gen.config1()[:200000])
    data = data[:data.rindex(b",")] + b"]"
    o = oracle.parse_standard(data)
    n = len(data)
    st = o.structural[1:o.N + 1].astype(np.int64).copy()
    rng = np.random.default_rng(4)
    for blk in (3, 17, 40):  # reverse runs inside some 256-gap blocks
        a = blk * 256 + 10
        st[a:a + 90] = st[a:a + 90][::-1]
    idx = rng.integers(0, st.size, 60)
    st[idx[:30]] = rng.integers(0, n, 30)         # rows anywhere in the text
    st[idx[30:]] = n + rng.integers(1, 1 << 20, 30)  # rows past the end
    rows = st.astype(np.uint32).view(np.int32)
    N = rows.size
    dev = DeviceBuffer(n + 64)
    dev.upload(np.frombuffer(data, np.uint8))
    d_st = DeviceBuffer(N * 4)
    d_st.upload(rows)
    vals = DeviceBuffer((N + 1) * cj.VALUE_DTYPE.itemsize)
    strs = DeviceBuffer(n + 64)
    rc, _ = ctx.values_device(dev.ptr, n, d_st.ptr, N, vals.ptr, strs.ptr)
    assert rc == 0
    got = np.frombuffer(vals.download(np.uint8, (N + 1) * cj.VALUE_DTYPE.itemsize).tobytes(),
                        dtype=cj.VALUE_DTYPE)
    bad = []
    for h in range(N + 1):
        b = n if h == N else min(int(st[h]) - 1, n)
        a = 0 if h == 0 else min(int(st[h - 1]), b)
        nxt = data[b] if h < N and b < n else None
        t, off, ln, fl, v, u = expected_entry(data, a, b, nxt)
        e = got[h]
        if (int(e["type"]), int(e["text_off"]), int(e["text_len"]), int(e["row"])) != (t, off, ln, h):
            bad.append((h, a, b, tuple(e.tolist()), (t, off, ln)))
        elif t in (cj.CJX_VALUE_INT, cj.CJX_VALUE_FLOAT, cj.CJX_VALUE_KEY, cj.CJX_VALUE_STRING):
            if int(e["v"]) != v or int(e["flags"]) != fl:
                bad.append((h, a, b, tuple(e.tolist()), (t, off, ln, fl, v)))
        if len(bad) >= 5:
            break
    assert not bad, bad
