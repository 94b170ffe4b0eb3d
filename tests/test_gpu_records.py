"""Per-record table of JSON Lines parses (CJX_FLAG_RECORDS) against records
derived from the oracle's output: a record ends at each '\\n' structural
entry of a chunk, plus an unterminated one when entries follow the chunk's
last '\\n'."""
import numpy as np
import pytest

import cujson_amd as cj
import oracle
from cujson_amd import gen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = cj.Context(0)
    yield c
    c.close()


def expected_records(data, chunks, o):
    recs = []
    st = o.structural
    row = 1
    for ci, ((off, size), cnt) in enumerate(zip(chunks, o.result_sizes)):
        code = oracle.parse_unit(bytes(data[off:off + size]), oracle.FLAG_JSONL).code
        first, prev_val = row, None
        for t in range(row, row + cnt):
            val = int(st[t])
            if data[val - 1] == 0x0A:
                recs.append((prev_val if prev_val is not None else off, val, first, t, ci, code))
                first, prev_val = t + 1, val
        if first <= row + cnt - 1:
            recs.append((prev_val if prev_val is not None else off, off + size, first,
                         row + cnt - 1, ci, code))
        row += cnt
    return np.array(recs, dtype=np.int64).reshape(-1, 6)


def check_records(ctx, data, chunks, label):
    data = bytes(data)
    inp = cj.cuJSONLinesInput(np.frombuffer(data, np.uint8), len(chunks), len(data),
                              [c[0] for c in chunks], [c[1] for c in chunks])
    g = cj.parse_json_lines(inp, cj.CJX_FLAG_RECORDS, ctx)
    o = oracle.parse_lines(data, chunks)
    assert g.error == o.code, label
    if o.code != oracle.OK:  # the reference aborts the file: only the codes compare
        codes = [oracle.parse_unit(bytes(data[a:a + n]), oracle.FLAG_JSONL).code
                 for a, n in chunks]
        assert all(codes[r["chunk"]] == r["error"] for r in g.records), label
        return
    assert np.array_equal(g.buf, o.buf), label
    want = expected_records(data, chunks, o)
    got = np.stack([g.records[f].astype(np.int64) for f in
                    ("byte_begin", "byte_end", "struct_first", "struct_last", "chunk", "error")],
                   axis=1) if g.records.size else np.zeros((0, 6), np.int64)
    assert got.shape == want.shape, (label, got.shape, want.shape)
    if not np.array_equal(got, want):
        bad = np.flatnonzero((got != want).any(axis=1))
        raise AssertionError(f"{label}: {bad.size} records differ, first {bad[:4]}: "
                             f"{got[bad[:2]]} vs {want[bad[:2]]}")


@pytest.mark.parametrize("cc", [1, 3, 51])
def test_records_twitter(ctx, twitter_small, cc):
    check_records(ctx, twitter_small, oracle.chunks_by_count(twitter_small, cc), f"cc={cc}")


def test_records_generated_lines(ctx):
    d = bytes(gen.lines(5000))
    for cc in (1, 7, 64):
        check_records(ctx, d, oracle.chunks_by_count(d, cc), f"lines cc={cc}")
    check_records(ctx, d, oracle.chunks_by_bytes(d, 4096), "lines 4 KiB chunks")


def test_records_edge_cases(ctx):
    cases = [b'{"a":1}\n{"b":[2]}', b'{"a":1}\n\n{"b":2}\n', b'[1,2]\n', b'1\n2\n3',
             b'{"x":"a\nb"}\n{"y":1}\n', b'\n\n\n', b'{"a":[}\n{"b":1}\n']
    for d in cases:
        for cc in (1, 2):
            chunks = [c for c in oracle.chunks_by_count(d, cc)]
            if any(s == 0 for _, s in chunks):
                continue
            check_records(ctx, d, chunks, f"{d!r} cc={cc}")


def test_device_lines_match_host_lines(ctx):
    """cjx_parse_lines_device on chunks placed at 16-byte aligned offsets:
    per-chunk rows, pair values (numbered across chunks) and records equal the
    host path's."""
    from cujson_amd.hip import DeviceBuffer
    d = bytes(gen.lines(3000))
    chunks = oracle.chunks_by_bytes(d, 8192)
    o = oracle.parse_lines(d, chunks)
    assert o.code == 0
    offs, pos = [], 0
    for _, size in chunks:
        offs.append(pos)
        pos += -(-size // 16) * 16
    buf = np.zeros(pos + 64, np.uint8)
    for (a, size), off in zip(chunks, offs):
        buf[off:off + size] = np.frombuffer(d[a:a + size], np.uint8)
    dev = DeviceBuffer(buf.size)
    dev.upload(buf)
    cap = len(d) + 64
    st, pr = DeviceBuffer(cap * 4), DeviceBuffer(cap * 4)
    max_rec = 10000
    rec = DeviceBuffer(max_rec * cj.RECORD_DTYPE.itemsize)
    assert ctx.parse_lines_device(dev.ptr, offs, [s for _, s in chunks], st.ptr, pr.ptr, cap,
                                  rec.ptr, max_rec) == 0
    rc, docs, N, M, over = ctx.batch_wait()
    assert rc == 0 and not over and N == o.N
    srow, prow = st.download(np.int32, cap), pr.download(np.int32, cap)
    packed_s = np.concatenate([srow[b:e] for _, b, e, _ in docs])
    packed_p = np.concatenate([prow[b:e] for _, b, e, _ in docs])
    assert np.array_equal(packed_s, o.structural[1:N + 1])
    assert np.array_equal(packed_p, o.pair_pos[1:N + 1])
    R = ctx.last_records
    got = np.frombuffer(rec.download(np.uint8, R * cj.RECORD_DTYPE.itemsize).tobytes(),
                        dtype=cj.RECORD_DTYPE)
    want = expected_records(d, chunks, o)
    got6 = np.stack([got[f].astype(np.int64) for f in
                     ("byte_begin", "byte_end", "struct_first", "struct_last", "chunk", "error")],
                    axis=1)
    assert np.array_equal(got6, want)
    # a table that is too small is reported, never overrun
    small = DeviceBuffer(16 * cj.RECORD_DTYPE.itemsize + 64)
    small.fill(0x5A)
    assert ctx.parse_lines_device(dev.ptr, offs, [s for _, s in chunks], st.ptr, pr.ptr, cap,
                                  small.ptr, 16) == 0
    rc, docs, N, M, over = ctx.batch_wait()
    assert rc == cj.CJX_ERR_CAPACITY and over
    assert np.all(small.download(np.uint8, 64, 16 * cj.RECORD_DTYPE.itemsize) == 0x5A)


def test_records_over_dirty_device_memory(twitter_small):
    """A fresh context whose buffers reuse memory that held other bytes: the
    record table depends only on the rows of the units (rows beside them in
    the flag buffer hold leftovers)."""
    from cujson_amd.hip import DeviceBuffer
    junk = [DeviceBuffer(64 << 20) for _ in range(4)]
    for j in junk:
        j.fill(0xA7)
    for j in junk:
        j.free()
    c = cj.Context(0)
    try:
        for cc in (3, 7, 51):
            check_records(c, twitter_small, oracle.chunks_by_count(twitter_small, cc),
                          f"dirty cc={cc}")
        # rows that outgrow their guessed regions on the device: the launch
        # reports CJX_ERR_CAPACITY and no records (not the previous counts)
        dense = (b"[" + b"1," * 3000 + b"1]\n") * 40
        ch = oracle.chunks_by_count(dense, 40)
        offs, pos = [], 0
        for _, size in ch:
            offs.append(pos)
            pos += -(-size // 16) * 16
        buf = np.zeros(pos + 64, np.uint8)
        for (a, size), off in zip(ch, offs):
            buf[off:off + size] = np.frombuffer(dense[a:a + size], np.uint8)
        dev = DeviceBuffer(buf.size)
        dev.upload(buf)
        cap = len(dense) - 1000  # below the exact bound: regions are the n/3 + 256 guess
        st, pr = DeviceBuffer(cap * 4), DeviceBuffer(cap * 4)
        rec = DeviceBuffer(10000 * cj.RECORD_DTYPE.itemsize)
        assert c.parse_lines_device(dev.ptr, offs, [z for _, z in ch], st.ptr, pr.ptr, cap,
                                    rec.ptr, 10000) == 0
        rc, docs, N, M, over = c.batch_wait()
        assert rc == cj.CJX_ERR_CAPACITY and over
        assert c.last_records == 0
    finally:
        c.close()
