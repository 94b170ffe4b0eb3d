"""The JSON Lines loader on the device: the newline scan (record offsets), the
chunk rules of loadJSONLines_chunkCount / _chunkSizeBytes
(cujson/load_file.cu:72-103, :153-176) on device text, the file-to-device
read, and the whole load -> chunk -> parse path with chunks at unaligned
offsets.  Newline offsets compare with numpy, chunks with the host chunker
and the oracle's restatement of the rules, parses with the oracle's parse."""
import numpy as np
import pytest

import cujson_amd as cj
import oracle
from cujson_amd import gen
from cujson_amd.hip import DeviceBuffer
from test_gpu_records import expected_records

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = cj.Context(0)
    yield c
    c.close()


def upload(data: bytes) -> DeviceBuffer:
    a = np.frombuffer(data, np.uint8)
    dev = DeviceBuffer(a.size + 64)
    if a.size:
        dev.upload(a)
    return dev


def newline_offsets(data: bytes) -> np.ndarray:
    return np.flatnonzero(np.frombuffer(data, np.uint8) == 0x0A).astype(np.uint64)


def texts(twitter_small):
    rng = np.random.default_rng(7)
    out = [("twitter", bytes(twitter_small)), ("generated", bytes(gen.lines(3000, seed=11))),
           ("empty", b""), ("one newline", b"\n"), ("one byte", b"a"), ("newlines", b"\n\n\n"),
           ("terminated", b"abc\n"), ("unterminated", b"abc"), ("crlf", b'{"a":1}\r\n' * 100),
           ("all newlines", b"\n" * 70000), ("no newline", b"x" * 200000)]
    for k, size in enumerate([1, 15, 16, 17, 31, 200, 4095, 4097]):
        a = rng.integers(0, 256, size, dtype=np.uint8)
        a[rng.random(size) < 0.1] = 0x0A
        out.append((f"random {size}", a.tobytes()))
    # 64 KiB segment edges, newlines every 37 bytes and on the edges themselves
    for size in (65535, 65536, 65537, 131077):
        a = np.full(size, ord("x"), np.uint8)
        a[36::37] = 0x0A
        for e in (65534, 65535, 65536):
            if e < size:
                a[e] = 0x0A
        out.append((f"segment edges {size}", a.tobytes()))
    return out


def test_scan_newlines(ctx, twitter_small):
    for label, d in texts(twitter_small):
        want = newline_offsets(d)
        dev = upload(d)
        cap = want.size + 1
        out = DeviceBuffer(cap * 8 + 64)
        rc, k = ctx.scan_lines_device(dev.ptr, len(d), out.ptr, cap)
        assert rc == cj.CJX_OK and k == want.size, label
        assert np.array_equal(out.download(np.uint64, k), want), label


def test_scan_capacity_reported_not_overrun(ctx):
    d = b'{"a":1}\n' * 1000
    dev = upload(d)
    out = DeviceBuffer(1000 * 8 + 64)
    out.fill(0x5A)
    rc, k = ctx.scan_lines_device(dev.ptr, len(d), out.ptr, 999)
    assert rc == cj.CJX_ERR_CAPACITY and k == 1000
    assert np.all(out.download(np.uint8, 1000 * 8 + 64) == 0x5A)


def test_chunks_match_host_and_oracle(ctx, twitter_small):
    for label, d in texts(twitter_small):
        dev = upload(d)
        for cc in (1, 2, 3, 7, 64, 1000):
            ho, hs = cj.chunk_lines(d, 0, cc)
            do, ds = ctx.chunk_lines_device(dev.ptr, len(d), 0, cc)
            assert list(do) == ho and list(ds) == hs, (label, cc)
            assert list(zip(ho, hs)) == oracle.chunks_by_count(d, cc), (label, cc)
        for b in (1, 16, 100, 4096, 65536, 1 << 40):
            ho, hs = cj.chunk_lines(d, 1, b)
            do, ds = ctx.chunk_lines_device(dev.ptr, len(d), 1, b)
            assert list(do) == ho and list(ds) == hs, (label, b)
            assert list(zip(ho, hs)) == oracle.chunks_by_bytes(d, b), (label, b)


def test_large_text(ctx):
    d = bytes(gen.lines(300000, seed=13))
    dev = upload(d)
    want = newline_offsets(d)
    out = DeviceBuffer(want.size * 8 + 64)
    rc, k = ctx.scan_lines_device(dev.ptr, len(d), out.ptr, want.size)
    assert rc == cj.CJX_OK and np.array_equal(out.download(np.uint64, k), want)
    for mode, p in ((0, 37), (0, 4096), (1, 262144), (1, 1 << 20)):
        ho, hs = cj.chunk_lines(d, mode, p)
        do, ds = ctx.chunk_lines_device(dev.ptr, len(d), mode, p)
        assert list(do) == ho and list(ds) == hs, (mode, p)


def test_load_device(ctx, tmp_path):
    # more parts than staging buffers (32 MiB parts, 4 buffers)
    d = bytes(gen.lines(160000, seed=17))
    assert len(d) > 4 * (32 << 20)
    path = tmp_path / "big.jsonl"
    path.write_bytes(d)
    buf = DeviceBuffer(len(d) + 64)
    rc, n = ctx.load_device(str(path), buf.ptr, len(d) + 64)
    assert rc == cj.CJX_OK and n == len(d)
    assert buf.download(np.uint8, n).tobytes() == d
    rc, n = ctx.load_device(str(path), buf.ptr, len(d) - 1)
    assert rc == cj.CJX_ERR_CAPACITY and n == len(d)
    rc, n = ctx.load_device(str(tmp_path / "missing.jsonl"), buf.ptr, len(d))
    assert rc == cj.CJX_ERR_INPUT
    empty = tmp_path / "empty.jsonl"
    empty.write_bytes(b"")
    assert ctx.load_device(str(empty), buf.ptr, 16) == (cj.CJX_OK, 0)


def parse_device_chunks(ctx, dev, d, offs, sizes, records):
    cap = len(d) + 64
    st, pr = DeviceBuffer(cap * 4), DeviceBuffer(cap * 4)
    max_rec = d.count(b"\n") + len(offs) + 1
    rec = DeviceBuffer(max_rec * cj.RECORD_DTYPE.itemsize) if records else None
    assert ctx.parse_lines_device(dev.ptr, offs, sizes, st.ptr, pr.ptr, cap,
                                  rec.ptr if rec else 0, max_rec if rec else 0) == 0
    rc, docs, N, M, over = ctx.batch_wait()
    srow, prow = st.download(np.int32, cap), pr.download(np.int32, cap)
    packed_s = np.concatenate([srow[b:e] for _, b, e, _ in docs])
    packed_p = np.concatenate([prow[b:e] for _, b, e, _ in docs])
    got_rec = None
    if rec is not None:
        R = ctx.last_records
        got = np.frombuffer(rec.download(np.uint8, R * cj.RECORD_DTYPE.itemsize).tobytes(),
                            dtype=cj.RECORD_DTYPE)
        got_rec = np.stack([got[f].astype(np.int64) for f in
                            ("byte_begin", "byte_end", "struct_first", "struct_last", "chunk",
                             "error")], axis=1)
    return rc, docs, N, packed_s, packed_p, got_rec


@pytest.mark.parametrize("mode,param", [(1, 8192), (0, 13), (1, 1 << 20)])
def test_load_chunk_parse(ctx, tmp_path, mode, param):
    """File -> device -> device chunking -> device parse of the chunks where
    they lie (unaligned offsets, so the units are relocated first)."""
    d = bytes(gen.lines(20000, seed=19))
    path = tmp_path / "records.jsonl"
    path.write_bytes(d)
    buf = DeviceBuffer(len(d) + 64)
    assert ctx.load_device(str(path), buf.ptr, len(d) + 64) == (cj.CJX_OK, len(d))
    offs, sizes = ctx.chunk_lines_device(buf.ptr, len(d), mode, param)
    chunks = [(int(a), int(s)) for a, s in zip(offs, sizes)]
    want_chunks = (oracle.chunks_by_bytes(d, param) if mode else
                   oracle.chunks_by_count(d, param))
    assert chunks == want_chunks
    assert any(a % 16 for a, _ in chunks)
    o = oracle.parse_lines(d, chunks)
    assert o.code == oracle.OK
    rc, docs, N, packed_s, packed_p, got_rec = parse_device_chunks(ctx, buf, d, offs, sizes, True)
    assert rc == 0 and N == o.N
    assert [int(x) for x in docs[:, 0]] == [a for a, _ in chunks]
    assert np.array_equal(packed_s, o.structural[1:N + 1])
    assert np.array_equal(packed_p, o.pair_pos[1:N + 1])
    assert np.array_equal(got_rec, expected_records(d, chunks, o))


def test_unaligned_batch_documents(ctx):
    """Independent documents packed back to back (unaligned offsets) parse as
    each document alone."""
    docs = [bytes(gen.standard_doc(n_objects=k, seed=30 + k)) for k in (1, 3, 17, 40, 2, 90)]
    docs.insert(2, b'{"a":[1,2}')  # a structure fault
    offs, pos = [], 3
    for x in docs:
        offs.append(pos)
        pos += len(x)
    blob = b"#" * 3 + b"".join(docs)
    dev = upload(blob)
    cap = len(blob) + 64
    st, pr = DeviceBuffer(cap * 4), DeviceBuffer(cap * 4)
    assert ctx.parse_batch_device(dev.ptr, offs, [len(x) for x in docs], st.ptr, pr.ptr, cap) == 0
    rc, res, N, M, over = ctx.batch_wait()
    srow, prow = st.download(np.int32, cap), pr.download(np.int32, cap)
    for i, x in enumerate(docs):
        o = oracle.parse_standard(x)
        byte_base, b, e, code = (int(v) for v in res[i])
        assert byte_base == offs[i] and code == o.code, i
        if o.code == oracle.OK:
            assert np.array_equal(srow[b:e], o.structural[1:o.N + 1]), i
            assert np.array_equal(prow[b:e], o.pair_pos[1:o.N + 1]), i
