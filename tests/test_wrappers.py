"""The reference-shaped C++ API (include/cujson/cujson.h, cujsonlines.h) driven
by tests/cpp/wrapper_check.cpp the way the reference's example mains call it
(main.cu:17-32, main_jsonlines.cu:16-36)."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "cujson_amd", "cjx_wrapper_check")
LARGE = os.path.join(ROOT, "tests", "golden", "twitter_sample_large_record.json")
SMALL = os.path.join(ROOT, "tests", "golden", "twitter_sample_small_records.json")


def run(*args):
    if not os.path.exists(BIN):
        pytest.fail("cjx_wrapper_check is not built (run __graft_entry__.build())")
    return subprocess.run([BIN, *map(str, args)], capture_output=True, text=True, timeout=300)


def read_dump(path):
    raw = np.fromfile(path, dtype=np.uint8)
    N, cc = np.frombuffer(raw[:16].tobytes(), np.int64)
    body = np.frombuffer(raw[16:].tobytes(), np.int32)
    sizes, prefix = body[:cc], body[cc:2 * cc]
    buf = body[2 * cc:]
    assert buf.size == 2 * (N + 2)
    return int(N), sizes.tolist(), prefix.tolist(), buf


def has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def test_headers_compile_and_program_runs():
    r = run("nonsense", "x")
    assert r.returncode == 2


def test_without_device_the_loader_fails_loudly():
    if has_gpu():
        pytest.skip("a GPU is present")
    r = run("load", SMALL, 3)
    assert "pinned" in r.stderr
    assert json.loads(r.stdout)["chunkCount"] == 0
    r = run("std", LARGE, "/dev/null")
    assert r.returncode == 3  # empty cuJSONResult, as the reference returns on bad input


@pytest.mark.gpu
def test_loader_chunks_match_reference_rule(tmp_path):
    data = open(SMALL, "rb").read()
    for cc in (1, 3, 7):
        out = json.loads(run("load", SMALL, cc).stdout)
        want = oracle.chunks_by_count(data, cc)
        assert out["offsets"] == [o for o, _ in want]
        assert out["sizes"] == [s for _, s in want]


@pytest.mark.gpu
def test_parse_standard_json_through_cpp_api(tmp_path):
    out = tmp_path / "std.bin"
    r = run("std", LARGE, out)
    assert r.returncode == 0, r.stdout + r.stderr
    N, sizes, prefix, buf = read_dump(out)
    o = oracle.parse_standard(open(LARGE, "rb").read())
    assert o.code == 0 and N == o.N
    assert sizes == [N] and prefix == [N]
    assert np.array_equal(buf, o.buf)


@pytest.mark.gpu
@pytest.mark.parametrize("cc", [1, 3, 8])
def test_parse_json_lines_through_cpp_api(tmp_path, cc):
    out = tmp_path / "lines.bin"
    r = run("lines", SMALL, cc, out)
    assert r.returncode == 0, r.stdout + r.stderr
    N, sizes, prefix, buf = read_dump(out)
    data = open(SMALL, "rb").read()
    o = oracle.parse_lines(data, oracle.chunks_by_count(data, cc))
    assert o.code == 0 and N == o.N
    assert sizes == list(o.result_sizes) and prefix == list(o.result_sizes_prefix)
    assert np.array_equal(buf, o.buf)
