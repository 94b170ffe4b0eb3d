"""Typed value decoding (cujson_amd/csrc/cjx_values.h) against Python's own
JSON parser: the host runs of the same decoders the kernel uses.  Numbers
must match bit for bit (Python's float() is correctly rounded, int() exact):
fixed edge cases, random doubles in several printed forms, random digit
strings with wide exponents, and exact halfway points between neighbouring
doubles (with neighbours just above and below).  Strings must unescape to
what json.loads gives (encoded with surrogatepass), or be rejected where it
rejects."""
import ctypes
import json
import math
import random
import struct
from decimal import Decimal, getcontext

import cujson_amd as cj


def number_fn():
    f = cj.lib().cjx_debug_parse_number
    f.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int64),
                  ctypes.POINTER(ctypes.c_uint8)]
    f.restype = ctypes.c_int
    return f


def decode(f, s: str):
    b = s.encode()
    v, fl = ctypes.c_int64(), ctypes.c_uint8()
    t = f(b, len(b), ctypes.byref(v), ctypes.byref(fl))
    return t, v.value, fl.value


def bits(x: float) -> int:
    return struct.unpack("<q", struct.pack("<d", x))[0]


def expected_number(s: str):
    """(type, value, flags) as the table defines them, from json.loads."""
    if s in ("NaN", "Infinity", "-Infinity"):
        return cj.CJX_VALUE_INVALID, None, 0
    try:
        o = json.loads(s)
    except ValueError:
        return cj.CJX_VALUE_INVALID, None, 0
    if isinstance(o, int) and not isinstance(o, bool):
        if -2 ** 63 <= o < 2 ** 63:
            return cj.CJX_VALUE_INT, o, 0
        return cj.CJX_VALUE_FLOAT, bits(float(o)), cj.CJX_VALUE_BIG_INT
    assert isinstance(o, float)
    return cj.CJX_VALUE_FLOAT, bits(o), 0


EDGE = ["0", "-0", "1", "-1", "0.0", "-0.0", "1e0", "1E+2", "1e-2", "123456789012345678",
        "9223372036854775807", "9223372036854775808", "-9223372036854775808",
        "-9223372036854775809", "18446744073709551616", "1e308", "1.7976931348623157e308",
        "1.7976931348623158e308", "1.7976931348623159e308", "2e308", "4.9e-324", "5e-324",
        "2.4703282292062327e-324", "2.4703282292062328e-324", "2.47e-324", "2.5e-324",
        "1e-400", "1e400", "2.2250738585072011e-308", "2.2250738585072014e-308",
        "9007199254740993", "9007199254740992", "9007199254740995", "4503599627370496.5",
        "4503599627370497.5", "0.1", "0.2", "0.3", "1e23", "8.589973e9", "1e22", "1e21",
        "1.00000000000000011102230246251565404236316680908203125",
        "1.00000000000000011102230246251565404236316680908203124",
        "1.00000000000000011102230246251565404236316680908203126",
        "123456789012345678901234567890", "0.000001", "1e-7", "7.2057594037927933e16",
        "1e000000000000000000000000000001", "0e999999999", "-0e-5",
        "01", "1.", ".1", "+1", "-", "--1", "1e", "1e+", "0x10", "NaN", "Infinity", "1.0e1.0",
        "1_000", "1e5.5", "00", "-01", "1.e3"]


def check_numbers(f, strings):
    bad = []
    for s in strings:
        t, v, fl = decode(f, s)
        et, ev, efl = expected_number(s)
        if et == cj.CJX_VALUE_INVALID:
            ok = t == et
        else:
            ok = t == et and v == ev and fl == efl
        if not ok:
            bad.append((s, (t, v, fl), (et, ev, efl)))
    assert not bad, bad[:5]


def test_number_edge_cases():
    check_numbers(number_fn(), EDGE)


def test_random_doubles_in_printed_forms():
    rng = random.Random(11)
    out = []
    while len(out) < 60000:
        x = struct.unpack("<d", struct.pack("<Q", rng.getrandbits(64)))[0]
        if math.isnan(x) or math.isinf(x):
            continue
        out += [repr(x), "%.17g" % x, "%.16g" % x, "%.15g" % x, "%.20e" % x, "%.3e" % x]
    check_numbers(number_fn(), out)


def test_random_digit_strings():
    rng = random.Random(12)
    out = []
    for _ in range(60000):
        nd = rng.randint(1, 60)
        digits = "".join(rng.choice("0123456789") for _ in range(nd)).lstrip("0") or "0"
        e = rng.randint(-370, 340)
        s = digits + "e" + str(e)
        if rng.random() < 0.3 and nd > 2:
            s = digits[:1] + "." + digits[1:] + "e" + str(e)
        if rng.random() < 0.3:
            s = "-" + s
        out.append(s)
    check_numbers(number_fn(), out)


def test_halfway_points():
    """Exact midpoints between neighbouring doubles round to even; a 60th
    digit above or below decides the other way round."""
    getcontext().prec = 1200
    rng = random.Random(13)
    out = []
    for _ in range(4000):
        x = struct.unpack("<d", struct.pack("<Q", rng.getrandbits(63)))[0]
        if math.isnan(x) or math.isinf(x) or x == 0:
            continue
        y = math.nextafter(x, math.inf)
        if math.isinf(y):
            continue
        mid = (Decimal(x) + Decimal(y)) / 2
        out.append(format(mid, "e"))
        step = Decimal(10) ** (mid.adjusted() - 60)
        out += [format(mid + step, "e"), format(mid - step, "e")]
    # subnormal and normal boundary midpoints
    for k in range(1, 64):
        a = 5e-324 * k
        out.append(format((Decimal(a) + Decimal(math.nextafter(a, 1.0))) / 2, "e"))
    check_numbers(number_fn(), out)


def unescape_fn():
    f = cj.lib().cjx_debug_unescape
    f.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_char_p,
                  ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_uint8)]
    f.restype = ctypes.c_int
    return f


def test_unescape_against_json():
    f = unescape_fn()
    rng = random.Random(14)
    pieces = ["a", "Z", " ", '\\"', "\\\\", "\\/", "\\b", "\\f", "\\n", "\\r", "\\t", "é", "中",
              "😀", "\\u0041", "\\u00e9", "\\u4e2d", "\\ud83d\\ude00", "\\ud800", "\\udc00",
              "\\uD834\\uDD1E", "\\u0000", "\\u001f", "\\uFFFF"]
    bad_pieces = ["\\x", "\\u12", "\\u12G4", '"', "\x01", "\x1f", "\\"]
    bad = []
    for _ in range(30000):
        parts = [rng.choice(pieces) for _ in range(rng.randint(0, 12))]
        if rng.random() < 0.15:
            parts.insert(rng.randint(0, len(parts)), rng.choice(bad_pieces))
        body = "".join(parts).encode("utf-8")
        try:
            s = json.loads(b'"' + body + b'"')
            want = s.encode("utf-8", "surrogatepass")
            want_flags = ((cj.CJX_VALUE_ESCAPED if b"\\" in body else 0) |
                          (cj.CJX_VALUE_UNPAIRED if any(0xD800 <= ord(c) <= 0xDFFF for c in s)
                           else 0))
        except ValueError:
            want = None
        out = ctypes.create_string_buffer(len(body) + 1)
        n, fl = ctypes.c_size_t(), ctypes.c_uint8()
        ok = f(body, len(body), out, ctypes.byref(n), ctypes.byref(fl))
        if want is None:
            if ok:
                bad.append((body, "accepted"))
        elif not ok or out.raw[:n.value] != want or fl.value != want_flags:
            bad.append((body, ok, out.raw[:n.value], fl.value, want, want_flags))
    assert not bad, bad[:5]


def test_unescape_word_copy_alignments():
    """Bodies and outputs at the same offset mod 4 take the word-at-a-time
    copy up to the first special byte; every offset, with the special byte
    (escape, quote, control byte) at every position of a word."""
    f = unescape_fn()
    f.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                  ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_uint8)]
    rng = random.Random(15)
    plain = "abcdefghij klmnopqrstuvwxyz0123456789é中"
    bad = []
    for body_len in range(0, 40):
        chars = [rng.choice(plain) for _ in range(body_len)]
        cases = ["".join(chars).encode("utf-8")]
        for pos in range(0, body_len + 1):  # specials between characters
            head, tail = "".join(chars[:pos]).encode(), "".join(chars[pos:]).encode()
            for sp in (b"\\n", b"\\u00e9", b'"', b"\x01", b"\\", b"\\q"):
                cases.append(head + sp + tail)
        for body in cases:
            try:
                s = json.loads(b'"' + body + b'"')
                want = s.encode("utf-8", "surrogatepass")
            except ValueError:
                want = None
            for k in range(4):
                src = ctypes.create_string_buffer(len(body) + 16)
                dst = ctypes.create_string_buffer(len(body) + 16)
                sa, da = ctypes.addressof(src), ctypes.addressof(dst)
                so = (k - sa) % 4  # body at address = k mod 4, output likewise
                do = (k - da) % 4
                ctypes.memmove(sa + so, body, len(body))
                n, fl = ctypes.c_size_t(), ctypes.c_uint8()
                ok = f(sa + so, len(body), da + do, ctypes.byref(n), ctypes.byref(fl))
                got = dst.raw[do:do + n.value]
                if want is None:
                    if ok:
                        bad.append((body, k, "accepted"))
                elif not ok or got != want:
                    bad.append((body, k, ok, got, want))
    assert not bad, bad[:5]
