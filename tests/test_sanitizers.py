"""AddressSanitizer + UndefinedBehaviorSanitizer runs of the host code (CPU):
the oracle (oracle/oracle.c), the host chunker (cujson_amd/csrc/cjx_chunk.cpp),
the bridge between the reference-shaped C++ API and the C ABI
(include/cujson/cjx_bridge.h) and the query iterators
(include/cujson/query_iterator.h).

Each is compiled here with gcc/g++ -fsanitize=address,undefined (leak
checking on, the first error fatal) into a driver under tests/cpp/ and run
on the fixtures, generated documents and random byte strings. Every result
is compared with the ordinary build (the oracle through ctypes, the
oracle's chunk rules, Python's json for the queries), so the sanitized runs
check behaviour as well as memory. The C ABI's device entry points are host
stand-ins in this build (tests/cpp/abi_host_stubs.cpp): no HIP runtime is
linked. A planted use-after-free and a planted signed overflow first show
that both sanitizers report errors in this environment."""
import os
import random
import shutil
import struct
import subprocess

import numpy as np
import pytest

import oracle
from cujson_amd import gen

import test_query as tq

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
INC = os.path.join(ROOT, "include")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
       "-g", "-O1"]
ENV = dict(os.environ,
           ASAN_OPTIONS="verify_asan_link_order=0:detect_leaks=1:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                                reason="gcc/g++ not installed")


def build(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]


@pytest.fixture(scope="module")
def bins(tmp_path_factory):
    d = tmp_path_factory.mktemp("san")
    out = {k: str(d / k) for k in ("planted", "oracle", "chunk", "query", "bridge")}
    planted = d / "planted.c"
    planted.write_text("#include <limits.h>\n#include <stdlib.h>\n"
                       "int main(int argc, char **argv) {\n"
                       "    if (argv[1][0] == 'a') { char *p = malloc(8); free(p); return p[0]; }\n"
                       "    int x = INT_MAX - 1; x += argc; return x & 1;\n}\n")
    build(["gcc", *SAN, "-o", out["planted"], str(planted)])
    build(["gcc", *SAN, "-std=c11", "-Wall", "-o", out["oracle"],
           os.path.join(CPP, "oracle_san.c"), os.path.join(ROOT, "oracle", "oracle.c")])
    chunker = os.path.join(ROOT, "cujson_amd", "csrc", "cjx_chunk.cpp")
    build(["g++", *SAN, "-std=c++17", "-Wall", "-I", INC, "-o", out["chunk"],
           os.path.join(CPP, "chunk_san.cpp"), chunker, "-lpthread"])
    stubs = os.path.join(CPP, "abi_host_stubs.cpp")
    build(["g++", *SAN, "-std=c++17", "-Wall", "-I", INC, "-o", out["query"],
           os.path.join(CPP, "query_check.cpp"), stubs, chunker, "-lpthread"])
    build(["g++", *SAN, "-std=c++17", "-Wall", "-I", INC, "-o", out["bridge"],
           os.path.join(CPP, "bridge_san.cpp"), stubs, chunker, "-lpthread"])
    return out


def run_clean(args, timeout=600):
    r = subprocess.run(args, capture_output=True, text=True, timeout=timeout, env=ENV)
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-3000:]
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])
    return r.stdout


def test_sanitizers_report_planted_errors(bins):
    r = subprocess.run([bins["planted"], "a"], capture_output=True, text=True, timeout=60, env=ENV)
    assert r.returncode != 0 and "heap-use-after-free" in r.stderr
    r = subprocess.run([bins["planted"], "u"], capture_output=True, text=True, timeout=60, env=ENV)
    assert r.returncode != 0 and "signed integer overflow" in r.stderr


# ---- the oracle ---------------------------------------------------------------


def _hash(v):
    v = np.asarray(v).astype(np.uint32).astype(np.uint64)
    return int((np.arange(1, v.size + 1, dtype=np.uint64) * v).sum(dtype=np.uint64))


def _expected(kind, flags, data):
    a = np.frombuffer(data, np.uint8)
    if kind == 0:
        r = oracle.parse_unit(a, flags)
        return f"parse {r.code} {r.structural.size} {r.M} {r.utf8_err} " \
               f"{_hash(r.structural)} {_hash(r.pair)}"
    if kind == 1:
        return f"utf8 {oracle.utf8_check(a, flags)}"
    if kind == 2:
        arr, ptr = oracle._u8(a)
        return f"strict {oracle.lib().orc_utf8_valid_strict(ptr, arr.size)}"
    if kind == 3:
        m = oracle.bitmaps(a, flags)
        names = ("bs", "quote", "op", "oc", "rq", "instr", "st", "ocm")
        return f"bitmaps {_hash(np.concatenate([m[k] for k in names]))}"
    # the brackets among the structurals: their bytes and rows
    r = oracle.parse_unit(a, flags)
    pos = r.structural.astype(np.int64) - 1
    inside = (pos >= 0) & (pos < a.size)
    rows = np.flatnonzero(inside & np.isin(a[np.clip(pos, 0, max(a.size - 1, 0))],
                                           list(b"[]{}")))
    oc = a[pos[rows]]
    err, pair, keys, perm = oracle.pair_phase(oc, rows.astype(np.uint32), r.structural.size,
                                              flags)
    return f"pair {err} {_hash(pair)} {_hash(keys)} {_hash(perm)}"


def _cases():
    rng = random.Random(77)
    alphabet = list(b'{}[]:,"\\a \n0') + [0xC3, 0xA9, 0xE2, 0x82, 0xAC, 0xF0, 0x9F, 0x80, 0xFF]
    docs = [open(tq.LARGE, "rb").read(), bytes(gen.config1()[:300000]), b"[" * 300 + b"]" * 300,
            b'{"a":[1,2,{"b":"\\"x\\\\"}]}', b"]", b"[", b"x", b'"\\', b"\xf0\x9f\x98"]
    for _ in range(60):
        n = rng.choice([1, 2, 3, 7, 31, 32, 33, 100, 1000, 5000])
        docs.append(bytes(rng.choice(alphabet) for _ in range(n)))
    cases = [(0, 0, b"")]
    for d in docs:
        for flags in (0, oracle.FLAG_DEPTH_U16, oracle.FLAG_UTF8_STRICT, oracle.FLAG_JSONL):
            cases.append((0, flags, d))
            cases.append((4, flags, d))
        cases += [(1, 0, d), (2, 0, d), (3, 0, d), (3, oracle.FLAG_JSONL, d)]
    return cases


def test_oracle_under_sanitizers(bins, tmp_path):
    cases = _cases()
    f = tmp_path / "cases.bin"
    with open(f, "wb") as fh:
        for kind, flags, d in cases:
            fh.write(struct.pack("<IIQ", kind, flags, len(d)) + d)
    got = run_clean([bins["oracle"], str(f)]).splitlines()
    assert len(got) == len(cases)
    for (kind, flags, d), line in zip(cases, got):
        if kind == 0 and not d:
            assert line.startswith(f"parse {oracle.ERR_INPUT} 0 0"), line
            continue
        assert line == _expected(kind, flags, d), (kind, flags, d[:40])


# ---- the host chunker -----------------------------------------------------------


def _chunks(bins, path, mode, param, limit=None):
    args = [bins["chunk"], str(path), str(mode), str(param)] + ([str(limit)] if limit else [])
    lines = run_clean(args).splitlines()
    rc, count = map(int, lines[0].split())
    return rc, [tuple(map(int, ln.split())) for ln in lines[1:1 + count]]


@pytest.mark.parametrize("mode,param", [(0, 1), (0, 4), (0, 51), (0, 5000), (1, 1), (1, 2000),
                                        (1, 100000), (1, 1 << 30)])
def test_chunker_under_sanitizers(bins, mode, param):
    data = open(tq.SMALL, "rb").read()
    rc, got = _chunks(bins, tq.SMALL, mode, param)
    want = (oracle.chunks_by_count(data, param) if mode == 0
            else oracle.chunks_by_bytes(data, param))
    assert rc == 0 and got == [tuple(c) for c in want]


def test_chunker_threads_and_fallback_under_sanitizers(bins, tmp_path):
    # 70 MB: the threaded newline scan, and its single-thread fallback when
    # threads cannot be created; an unterminated last line; no newline at all
    d = gen.lines_parallel(80_000)
    assert d.size >= 64 << 20
    data = d.tobytes() + b'{"tail":1}'
    p = tmp_path / "big.jsonl"
    p.write_bytes(data)
    want = [tuple(c) for c in oracle.chunks_by_bytes(data, 1 << 20)]
    for limit in (None, 2):
        rc, got = _chunks(bins, p, 1, 1 << 20, limit)
        assert rc == 0 and got == want, limit
    rc, got = _chunks(bins, p, 0, 997)
    assert rc == 0 and got == [tuple(c) for c in oracle.chunks_by_count(data, 997)]
    q = tmp_path / "one.json"
    q.write_bytes(b"[1,2,3]")
    assert _chunks(bins, q, 0, 3) == (0, [(0, 7), (7, 0), (7, 0)])
    assert _chunks(bins, q, 2, 3)[0] == 1  # CJX_ERR_INPUT: unknown mode


# ---- bridge and iterators -----------------------------------------------------------


def test_bridge_under_sanitizers(bins, tmp_path):
    o = oracle.parse_standard(open(tq.LARGE, "rb").read())
    rf = tmp_path / "r.i32"
    o.buf.astype(np.int32).tofile(rf)
    out = run_clean([bins["bridge"], str(rf), tq.SMALL, "4"])
    assert "fail" not in out
    data = open(tq.SMALL, "rb").read()
    got = [tuple(map(int, ln.split()[1:])) for ln in out.splitlines() if ln.startswith("chunk ")]
    assert got == [tuple(c) for c in oracle.chunks_by_count(data, 4)]


def test_query_iterators_under_sanitizers(bins, tmp_path, monkeypatch):
    """The iterator query tests over the oracle's results (the driver's
    --result form), run by the sanitized driver: standard documents, JSON
    Lines queries at two chunkings and the record scan."""
    monkeypatch.setattr(tq, "BIN", bins["query"])
    monkeypatch.setenv("ASAN_OPTIONS", ENV["ASAN_OPTIONS"])
    monkeypatch.setenv("UBSAN_OPTIONS", ENV["UBSAN_OPTIONS"])
    tq.check_doc(tmp_path, "cpu", tq.LARGE, 300, 5)
    cfg = tmp_path / "config1.json"
    gen.config1().tofile(str(cfg))
    tq.check_doc(tmp_path, "cpu", str(cfg), 200, 6)
    deep = tmp_path / "deep.json"
    gen.standard_doc(n_objects=200, deep_depth=40, seed=8).tofile(str(deep))
    tq.check_doc(tmp_path, "cpu", str(deep), 200, 7)
    for chunks in (1, 3):
        tq.check_lines_doc(tmp_path, "cpu", tq.SMALL, 200, 30 + chunks, chunks)
    assert tq.check_scan(tmp_path, "cpu", tq.SMALL, 2, "lang", "en", "id") > 0
