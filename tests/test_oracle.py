"""Pins the CPU oracle (no GPU).

Each check compares the oracle against something formulated independently
of its word-level arithmetic: the scalar state machine in scalar_ref, brute
force over tiny inputs (with padding so words, carries and backslash walks
cross 32-bit boundaries), stack bracket matching, simdjson 3.2.1 stage 1 built
from the reference's vendored sources, RFC 3629 validation, and known
answers implied by the reference's rules.
"""
import itertools
import json
import random

import numpy as np
import pytest

import oracle
import scalar_ref as S


def test_intrinsics_closed_forms():
    L = oracle.lib()
    assert L.orc_vcmpeq4(0x11223344, 0x11003344) == 0xFF00FFFF
    assert L.orc_vcmpltu4(0x01FF8000, 0x02FE8001) == 0xFF0000FF
    assert L.orc_vcmpgeu4(0x01FF8000, 0x02FE8001) == 0x00FFFF00
    assert L.orc_vcmpgtu4(0x01FF8000, 0x02FE8001) == 0x00FF0000
    assert L.orc_vsubus4(0x10FF0502, 0x20010103) == 0x00FE0400
    # signed saturation: 0x80 - 0x01 saturates at 0x80 (-128); 0x7F - 0xFF = 127+1 -> 0x7F
    assert L.orc_vsubss4(0x80007F10, 0x0100FF01) == 0x80007F0F


ALPHABET = [ord(c) for c in '{}[]:,"\\a \n'] + [0xC3, 0xA9, 0xE2, 0x82, 0xAC, 0xF0]


def _check(data, jsonl=False, u16=False):
    flags = (oracle.FLAG_JSONL if jsonl else 0) | (oracle.FLAG_DEPTH_U16 if u16 else 0)
    r = oracle.parse_unit(data, flags)
    code, pos, row = S.scalar_parse(bytes(data), jsonl, u16)
    assert r.structural.tolist() == pos, data
    assert r.pair.tolist() == row, data
    assert r.code == code, data


def test_bruteforce_short_strings():
    # every string of length <= 4 over a 17-letter alphabet is too many with
    # padding variants; enumerate length <= 3 fully and sample length 4..10
    for L in range(1, 4):
        for tup in itertools.product(ALPHABET, repeat=L):
            _check(bytes(tup))
    rng = random.Random(1234)
    for _ in range(20000):
        L = rng.randint(4, 10)
        _check(bytes(rng.choice(ALPHABET) for _ in range(L)))


@pytest.mark.parametrize("pad", [27, 28, 29, 30, 31, 32, 33, 60, 63, 64, 65, 95, 96])
def test_bruteforce_word_boundaries(pad):
    """Backslash runs, quotes and UTF-8 sequences straddling 32-bit words."""
    rng = random.Random(pad)
    fills = [b"a" * pad, b"\\" * pad, b'"' + b"a" * (pad - 1), b"\\" * (pad - 1) + b'"']
    for _ in range(3000):
        head = rng.choice(fills)
        tail = bytes(rng.choice(ALPHABET) for _ in range(rng.randint(1, 8)))
        _check(head + tail)
        _check(head + tail, jsonl=True)


def test_long_backslash_walk():
    # the escape walk crosses several all-backslash words (PSJ:810-816)
    for k in range(60, 140):
        for tail in (b'"]', b'a"[', b'"', b'\\"'):
            _check(b"\\" * k + tail)
            _check(b'["' + b"\\" * k + tail + b'"]')


def test_random_valid_json_matches_stack_matching():
    rng = random.Random(7)

    def gen(d):
        k = rng.random()
        if d > 6 or k < 0.3:
            return rng.choice(["1", '"a\\"]"', '"\\\\"', "null", '"{x:1}"', '"é€"'])
        if k < 0.65:
            return "[" + ",".join(gen(d + 1) for _ in range(rng.randint(0, 4))) + "]"
        return "{" + ",".join('"k%d":%s' % (i, gen(d + 1)) for i in range(rng.randint(0, 4))) + "}"

    for _ in range(3000):
        s = gen(0).encode()
        json.loads(s)
        r = oracle.parse_unit(s, 0)
        assert r.code == 0
        pos = S.lex(s)
        assert r.structural.tolist() == pos
        ok, pair = S.stack_pairs(s, pos)
        assert ok
        for t, t2 in pair.items():
            assert r.pair[t] == t2 + 1  # host index of the closer
        non_openers = [t for t in range(len(pos)) if t not in pair]
        assert all(r.pair[t] == -1 for t in non_openers)


def test_known_answers_of_the_pairing_rule():
    """Cases the reference's group-of-four rule fixes (PSJ:1441-1509)."""
    code = lambda s: oracle.parse_unit(s, 0).code
    assert code(b"[]") == 0
    assert code(b"][") == 0          # last group of 2: xor only, no order check
    assert code(b"[]][") == 0        # last group of 4: xor only
    assert code(b"[][]][") == 0      # "][" sorts into the last group (key 255)
    assert code(b"[]][][") == 3      # order is checked outside the last group
    assert code(b"[}") == 3
    assert code(b"[") == 3 and code(b"[[]") == 3  # M % 4 in {1, 3}
    assert code(b'["]"]') == 0       # bracket inside a string
    assert code(b"1") == 0           # no brackets: M = 0 is fine
    # depth keys are uint8: nesting 256 passes, 257 fails (S8)
    assert code(b"[" * 256 + b"]" * 256) == 0
    assert code(b"[" * 257 + b"]" * 257) == 3
    assert oracle.parse_unit(b"[" * 257 + b"]" * 257, oracle.FLAG_DEPTH_U16).code == 0
    assert code(b"[" * 300 + b"]" * 300) == 3


def test_utf8_reference_rule_known_answers():
    u = lambda s: oracle.utf8_check(s, 0)
    assert u("é€😀".encode()) == 0
    assert u(b"abc\xc0\xafd") == 0            # overlong: not detected by the weak check
    assert u(b'ab"\xed\xa0\x80') == 0         # surrogate: not detected
    assert u(b"a\xc3Abcd") == 0               # 2-byte truncation inside a word
    assert u(b"abc\xc3defg") == 1             # lead byte before an ASCII word
    assert u(b"a\xe2\x82bcdefg") == 1         # 3-byte truncation
    assert u(b"\xe2\x82\xac\xe2\x82") == 1    # truncated at the end (zero padding)
    assert u(b"abcd\xe2\x82\xac") == 0


def test_utf8_rule_bruteforce_and_strict_mode():
    rng = random.Random(99)
    bytes_pool = [0x41, 0x80, 0x9F, 0xBF, 0xC2, 0xC0, 0xDF, 0xE0, 0xED, 0xEF, 0xF0, 0xF4, 0xF5, 0xFF]
    for _ in range(20000):
        s = bytes(rng.choice(bytes_pool) for _ in range(rng.randint(1, 12)))
        assert oracle.utf8_check(s, 0) == int(S.utf8_word_rule(s)), s
        strict = oracle.utf8_check(s, oracle.FLAG_UTF8_STRICT)
        try:
            s.decode("utf-8")
            valid = True
        except UnicodeDecodeError:
            valid = False
        assert strict == (0 if valid else 1), s


def test_valid_utf8_never_flagged_by_weak_rule():
    rng = random.Random(5)
    for _ in range(3000):
        cps = [rng.choice([rng.randint(0x20, 0x7E), rng.randint(0x80, 0x7FF),
                           rng.randint(0x800, 0xD7FF), rng.randint(0x10000, 0x10FFFF)])
               for _ in range(rng.randint(1, 20))]
        s = "".join(chr(c) for c in cps).encode()
        assert oracle.utf8_check(s, 0) == 0


def test_fixture_twitter_large(twitter_large):
    r = oracle.parse_standard(twitter_large)
    assert r.code == 0
    assert r.N == 22388                    # structural count of the sample
    assert int((np.frombuffer(twitter_large, np.uint8) >= 0x80).sum()) == 0
    pos = S.lex(twitter_large)
    assert r.structural[1:r.N + 1].tolist() == pos
    ok, pair = S.stack_pairs(twitter_large, pos)
    assert ok
    for t, t2 in pair.items():
        assert r.pair_pos[t + 1] == t2 + 1


def test_fixture_twitter_small_records(twitter_small):
    # every chunking of valid JSON Lines gives the same layout (S16)
    base = None
    for cc in (1, 2, 3, 4, 8, 51):
        ch = oracle.chunks_by_count(twitter_small, cc)
        r = oracle.parse_lines(twitter_small, ch)
        assert r.code == 0
        assert r.N == 22438
        if base is None:
            base = r.buf.copy()
        assert np.array_equal(r.buf, base)
        assert sum(r.result_sizes) == r.N
    nl = sum(1 for p in S.lex(twitter_small, True) if twitter_small[p - 1] == 0x0A)
    assert nl == twitter_small.count(b"\n")  # no raw newlines inside strings


def test_chunking_rules(twitter_small):
    ch = oracle.chunks_by_count(twitter_small, 4)
    assert sum(s for _, s in ch) == len(twitter_small)
    assert [o for o, _ in ch] == sorted(o for o, _ in ch)
    # byte-limited: a first line longer than the limit yields a zero-length chunk
    data = b"x" * 100 + b"\n" + b"y\n"
    assert oracle.chunks_by_bytes(data, 10)[0] == (0, 0)
    assert oracle.chunks_by_bytes(data, 1000) == [(0, len(data))]
    assert oracle.chunks_by_count(b"a\nb\nc\n", 2) == [(0, 4), (4, 2)]
    assert oracle.chunks_by_count(b"a\nb", 3) == [(0, 2), (2, 1), (3, 0)]


def test_jsonl_offsets_carry():
    data = b'{"a":[1]}\n{"b":{"c":2}}\n'
    ch = [(0, 10), (10, 15)]
    r = oracle.parse_lines(data, ch)
    assert r.code == 0
    pos = S.lex(data, True)
    assert r.structural[1:r.N + 1].tolist() == pos
    ok, pair = S.stack_pairs(data, pos)
    for t, t2 in pair.items():
        assert r.pair_pos[t + 1] == t2 + 1
    assert r.result_sizes == [6, 7]


def test_sentinels_and_layout():
    r = oracle.parse_standard(b'{"a":[1,2]}')
    N = r.N
    assert r.structural[0] == 0 and r.structural[N + 1] == N + 1
    assert r.pair_pos[0] == N + 1
    assert r.buf.size == 2 * (N + 2)
    assert r.total_result_size == N + 2 == r.file_size


def test_empty_input():
    assert oracle.parse_standard(b"").code == oracle.ERR_INPUT
    assert oracle.parse_unit(b"").code == oracle.ERR_INPUT


# ---------------------------------------------------------------------------
# Cross-checks against simdjson 3.2.1 built from the reference's vendored
# sources (oracle/Makefile, into oracle/_ref/).  Without that build these
# tests are skipped with the reason, never passed silently.
# ---------------------------------------------------------------------------
def _simdjson():
    ref = oracle.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref/libsimdjson_ref.so absent: simdjson 3.2.1 is built by "
                    "oracle/Makefile from the reference tree, which this host lacks")
    return ref


def _config_docs():
    from cujson_amd import gen
    return {"config1": gen.config1(),
            "config2 shape 24 MiB": gen.config2(target_bytes=24 << 20),
            "config4 document": gen.standard_doc(target_bytes=3 << 20, seed=gen.SEED_BASE + 4)}


def test_simdjson_stage1_on_fixture_and_configs(twitter_large):
    """simdjson stage-1 structural indices filtered to { } [ ] : , equal the
    oracle's structural positions (valid JSON: S4-S7 as a textbook lexer)."""
    _simdjson()
    docs = {"twitter_large": np.frombuffer(twitter_large, np.uint8)}
    docs.update(_config_docs())
    for label, d in docs.items():
        r = oracle.parse_unit(d, 0)
        assert r.code == 0, label
        e, idx = oracle.simdjson_op_indices(d)
        assert e == 0, (label, e)
        assert idx.size == r.structural.size, (label, idx.size, r.structural.size)
        assert np.array_equal(idx, r.structural), label


def test_simdjson_validate_utf8_matches_strict_mode():
    ref = _simdjson()
    rng = random.Random(99)
    bytes_pool = [0x41, 0x80, 0x9F, 0xBF, 0xC2, 0xC0, 0xDF, 0xE0, 0xED, 0xEF, 0xF0, 0xF4, 0xF5, 0xFF]
    for _ in range(20000):
        s = bytes(rng.choice(bytes_pool) for _ in range(rng.randint(1, 12)))
        a, p = oracle._u8(s)
        strict = oracle.utf8_check(s, oracle.FLAG_UTF8_STRICT)
        assert ref.sjr_validate_utf8(p, a.size) == (1 - strict), s
    for label, d in _config_docs().items():
        a, p = oracle._u8(d)
        assert ref.sjr_validate_utf8(p, a.size) == 1 - oracle.utf8_check(d, oracle.FLAG_UTF8_STRICT)
