"""Batches of independent documents (one launch sequence for all of them)
against the oracle run on each document alone: per-document error codes and
the full cuJSONResult buffer of every document, bit for bit."""
import numpy as np
import pytest

import cujson_amd as cj
import oracle
from cujson_amd import gen
from cujson_amd.hip import DeviceBuffer

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = cj.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def batch():
    return gen.many_docs(n_docs=600, max_size=300 << 10, fault_frac=0.2, deep_frac=0.1, seed=11)


def check_batch(ctx, data, offs, sizes, flags, label):
    r = cj.parse_batch(data, offs.tolist(), sizes.tolist(), flags, ctx)
    assert r.rc == 0, (label, r.rc)
    assert r.n_docs == len(offs)
    first = 0
    for i, (o, s) in enumerate(zip(offs, sizes)):
        want = oracle.parse_standard(data[o:o + s].tobytes(), flags)
        err, buf = r.doc(i)
        assert err == want.code, (label, i, err, want.code)
        assert r.docs[i, 0] == o
        if not np.array_equal(buf, want.buf):
            bad = np.flatnonzero(buf != want.buf)
            raise AssertionError(f"{label} doc {i}: {bad.size} cells differ, first {bad[:6]}")
        if err and not first:
            first = err
    assert r.error == first
    return r


@pytest.mark.parametrize("flags", [0, oracle.FLAG_DEPTH_U16, oracle.FLAG_UTF8_STRICT])
def test_batch_matches_oracle_per_document(ctx, batch, flags):
    data, offs, sizes, kinds = batch
    check_batch(ctx, data, offs, sizes, flags, f"flags={flags}")


def test_batch_error_codes_by_fault_kind(ctx, batch):
    data, offs, sizes, kinds = batch
    r = cj.parse_batch(data, offs.tolist(), sizes.tolist(), 0, ctx)
    codes = r.docs[:, 3]
    # overlong and surrogate faults pass the reference's weak UTF-8 rule
    assert set(codes[kinds == 0]) <= {0, 3} and set(codes[kinds == 1]) <= {0, 3}
    assert (codes[kinds == 2] == cj.CJX_ERR_UTF8).any()
    assert (codes[kinds == 3] == cj.CJX_ERR_STRUCTURE).all()
    s = cj.parse_batch(data, offs.tolist(), sizes.tolist(), cj.CJX_FLAG_UTF8_STRICT, ctx)
    assert (s.docs[kinds == 0, 3] == cj.CJX_ERR_UTF8).all()
    assert (s.docs[kinds == 1, 3] == cj.CJX_ERR_UTF8).all()


def test_batch_of_tiny_and_mixed_documents(ctx):
    docs = [b"[]", b"1", b"{}", b'"x"', b"[1,[2,[3]]]", b"][", b"[}", b"[" * 300 + b"]" * 300,
            b'{"a":"\\\\"}', b"\xc3\xa9", b"[\"\xe2\x82\"]", b" ", b"[" * 257 + b"]" * 257]
    docs = docs * 7
    offs, pos, parts = [], 0, []
    for d in docs:
        offs.append(pos)
        parts.append(d)
        pos += len(d)
    data = np.frombuffer(b"".join(parts), np.uint8)
    sizes = np.array([len(d) for d in docs])
    for flags in (0, oracle.FLAG_DEPTH_U16):
        check_batch(ctx, data, np.array(offs), sizes, flags, f"tiny flags={flags}")


def test_batch_round_and_wave_edge_documents(ctx):
    """Documents at the one-round (2 KiB) and one-wave (16 KiB) limits of
    the small-unit kernels and just past them (a tile), strings and
    backslash runs across those edges, in one batch, every flag set."""
    base = bytes(gen.config1()[:40000])
    docs = []
    for n in (1, 2, 31, 32, 33, 2046, 2047, 2048, 2049, 2050, 4096, 16383, 16384, 16385, 20000):
        docs.append(base[:n])
        for k in (1, 2, 3):
            pre = b'["' + b"x" * max(0, n - 6 - k)
            docs.append(pre + b"\\" * k + b'"]",1]')
    offs, pos, parts = [], 0, []
    for d in docs:
        offs.append(pos)
        parts.append(d)
        pos += len(d)
    data = np.frombuffer(b"".join(parts), np.uint8)
    sizes = np.array([len(d) for d in docs])
    for flags in (0, oracle.FLAG_DEPTH_U16, oracle.FLAG_UTF8_STRICT):
        check_batch(ctx, data, np.array(offs), sizes, flags, f"edges flags={flags}")


def test_batch_rejects_empty_document(ctx):
    data = np.frombuffer(b"[1][2]", np.uint8)
    r = cj.parse_batch(data, [0, 3, 3], [3, 0, 3], 0, ctx)
    assert r.rc == cj.CJX_ERR_INPUT


def test_device_batch_matches_host_batch(ctx, batch):
    data, offs, sizes, kinds = batch
    host = cj.parse_batch(data, offs.tolist(), sizes.tolist(), 0, ctx)
    dev = DeviceBuffer(data.size)
    dev.upload(data)
    cap = int(sizes.sum()) + 64
    st = DeviceBuffer(cap * 4)
    pr = DeviceBuffer(cap * 4)
    assert ctx.parse_batch_device(dev.ptr, offs.tolist(), sizes.tolist(), st.ptr, pr.ptr, cap) == 0
    rc, docs, N, M, over = ctx.batch_wait()
    assert rc == 0 and not over and N == host.N
    srow = st.download(np.int32, cap)
    prow = pr.download(np.int32, cap)
    assert np.array_equal(docs[:, 0], offs)
    assert np.array_equal(docs[:, 3], host.docs[:, 3])
    for i in range(len(offs)):
        b, e = docs[i, 1], docs[i, 2]
        hb, he = host.docs[i, 1], host.docs[i, 2]
        assert e - b == he - hb
        assert np.array_equal(srow[b:e], host.structural[hb:he])
        assert np.array_equal(prow[b:e], host.pair_pos[hb:he])
    # a capacity below the guessed regions is refused before launching
    assert ctx.parse_batch_device(dev.ptr, offs.tolist(), sizes.tolist(), st.ptr, pr.ptr,
                                  1000) == cj.CJX_ERR_CAPACITY


def test_batch_twice_in_a_row_smaller_second(ctx, batch):
    data, offs, sizes, kinds = batch
    check_batch(ctx, data, offs, sizes, 0, "first")
    check_batch(ctx, data, offs[:50], sizes[:50], 0, "second, smaller")


def _device_rows(ctx, dev, offs, sizes, cap, st, pr, flags):
    assert ctx.parse_batch_device(dev.ptr, offs, sizes, st.ptr, pr.ptr, cap, flags) == 0
    rc, docs, N, M, over = ctx.batch_wait()
    assert rc == 0 and not over
    srow = st.download(np.int32, cap)
    prow = pr.download(np.int32, cap)
    return docs, [(srow[b:e].copy(), prow[b:e].copy()) for _, b, e, _ in docs]


@pytest.mark.parametrize("u16", [False, True])
def test_graph_replays_match_eager(ctx, batch, u16):
    """The batch pipeline captured as a HIP graph, replayed several times with
    eager parses of other inputs between the replays; every replay must equal
    the eager result."""
    data, offs, sizes, kinds = batch
    offs, sizes = offs.tolist(), sizes.tolist()
    dev = DeviceBuffer(data.size)
    dev.upload(data)
    cap = int(sum(sizes)) + 64
    st, pr = DeviceBuffer(cap * 4), DeviceBuffer(cap * 4)
    base = cj.CJX_FLAG_DEPTH_U16 if u16 else 0
    want_docs, want_rows = _device_rows(ctx, dev, offs, sizes, cap, st, pr, base)
    other = gen.config1()
    odev = DeviceBuffer(other.size + 64)
    odev.upload(other)
    ost, opr = DeviceBuffer(other.size * 4), DeviceBuffer(other.size * 4)
    o_want = oracle.parse_standard(other.tobytes())
    for rep in range(4):
        st.fill(0x7F)
        pr.fill(0x7F)
        docs, rows = _device_rows(ctx, dev, offs, sizes, cap, st, pr, base | cj.CJX_FLAG_GRAPH)
        assert np.array_equal(docs, want_docs), rep
        for i, ((s1, p1), (s2, p2)) in enumerate(zip(rows, want_rows)):
            assert np.array_equal(s1, s2) and np.array_equal(p1, p2), (rep, i)
        # eager work between replays: a different document on the same context
        assert ctx.parse_device(odev.ptr, other.size, ost.ptr, opr.ptr, other.size) == 0
        rc, N, M, over = ctx.wait()
        assert rc == o_want.code and N == o_want.N
        assert np.array_equal(ost.download(np.int32, N), o_want.structural[1:N + 1])
