"""Parity at the sizes the benchmark times, through the device-resident C ABI
entries the benchmark calls, against the oracle (bit-exact rows and codes):

* the 2.0e9-byte standard-JSON document (post-like objects, a depth-64
  branch, every 8th 1 MiB boundary inside a string) through
  cjx_parse_device, launched and as a HIP graph replay;
* the 50,000-document batch (Pareto sizes, faults, deep documents) through
  cjx_parse_batch_device, launched and as a graph replay;
* the first launch of the 8M-record JSON Lines file (every 1 MiB chunk
  below 2^31 bytes, about 2.15 GB in 2,050 chunks) through
  cjx_parse_lines_device with the per-record table;
* a 200 MiB request to the request server among small ones.

Each case generates its input on the host, runs the oracle once and
compares every cell.  Host memory peaks at about 20 GB (the JSON Lines case)."""
import numpy as np
import pytest

import cujson_amd as cj
import oracle
from cujson_amd import autotune, gen
from cujson_amd.hip import DeviceBuffer

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = cj.Context(0)
    yield c
    c.close()


def _first_diff(got, want):
    bad = np.flatnonzero(got != want)
    return f"{bad.size} cells differ, first at {bad[:6]}: {got[bad[:6]]} vs {want[bad[:6]]}"


def test_plain_2gb_document(ctx):
    doc = gen.config2_parallel(total_bytes=2_000_000_000)
    n = doc.size
    want = oracle.parse_unit(doc, 0)
    assert want.code == oracle.OK
    N = want.structural.size
    d_in = DeviceBuffer(((n + 63) // 64) * 64)
    d_in.upload(doc)
    del doc
    cap = n // 4 + 4096
    d_st, d_pr = DeviceBuffer(cap * 4), DeviceBuffer(cap * 4)
    for flags, label in ((0, "launched"), (cj.CJX_FLAG_GRAPH, "graph")):
        for rep in range(2 if flags else 1):  # the graph is captured on the first call
            d_st.fill(0x7F)
            d_pr.fill(0x7F)
            assert ctx.parse_device(d_in.ptr, n, d_st.ptr, d_pr.ptr, cap, flags) == 0
            rc, gN, gM, over = ctx.wait()
            assert rc == want.code and not over, (label, rc, over)
            assert gN == N and gM == want.M, (label, gN, N, gM, want.M)
            got = d_st.download(np.uint32, N)
            assert np.array_equal(got, want.structural), f"{label} structural: " + \
                _first_diff(got, want.structural)
            got = d_pr.download(np.int32, N)
            assert np.array_equal(got, want.pair), f"{label} pair: " + _first_diff(got, want.pair)
            del got


@pytest.fixture(scope="module")
def docs50k():
    return gen.many_docs(n_docs=50_000)


def test_batch_50k_documents(ctx, docs50k):
    data, offs, sizes, kinds = docs50k
    total = int(sizes.sum())
    want_rows, want_pairs, want_codes = [], [], []
    for o, s in zip(offs.tolist(), sizes.tolist()):
        r = oracle.parse_unit(data[o:o + s], 0)
        want_rows.append(r.structural)
        want_pairs.append(r.pair)
        want_codes.append(r.code)
    want_codes = np.array(want_codes)
    want_s = np.concatenate(want_rows).view(np.int32)
    want_p = np.concatenate(want_pairs)
    assert set(np.unique(want_codes)) == {0, 2, 3}  # faults of both kinds are in the batch
    d_in = DeviceBuffer(data.size)
    d_in.upload(data)
    cap = total + 64  # as the benchmark sizes it: exact regions
    d_st, d_pr = DeviceBuffer(cap * 4), DeviceBuffer(cap * 4)
    for flags, label in ((0, "launched"), (cj.CJX_FLAG_GRAPH, "graph"),
                         (cj.CJX_FLAG_GRAPH, "graph replay")):
        d_st.fill(0x7F)
        d_pr.fill(0x7F)
        assert ctx.parse_batch_device(d_in.ptr, offs, sizes, d_st.ptr, d_pr.ptr, cap, flags) == 0
        rc, docs, N, M, over = ctx.batch_wait()
        assert not over and N == want_s.size, (label, rc, over, N, want_s.size)
        assert np.array_equal(docs[:, 0], offs), label
        assert np.array_equal(docs[:, 3], want_codes), \
            (label, np.flatnonzero(docs[:, 3] != want_codes)[:8])
        assert np.array_equal(docs[:, 2] - docs[:, 1], [r.size for r in want_rows]), label
        srow = d_st.download(np.int32, cap)
        prow = d_pr.download(np.int32, cap)
        idx = np.concatenate([np.arange(b, e) for b, e in docs[:, 1:3]])
        got = srow[idx]
        assert np.array_equal(got, want_s), f"{label} structural: " + _first_diff(got, want_s)
        got = prow[idx]
        assert np.array_equal(got, want_p), f"{label} pair: " + _first_diff(got, want_p)


def expected_records(text, chunks, o, codes):
    """The per-record table the oracle's output implies (vectorized form of
    the rule: a record ends at each '\\n' structural entry of a chunk; rows
    after a chunk's last '\\n' form one unterminated record)."""
    N = o.N
    st = o.buf[1:N + 1].astype(np.int64)
    sizes = np.asarray(o.result_sizes, np.int64)
    c_off = np.array([c[0] for c in chunks], np.int64)
    c_size = np.array([c[1] for c in chunks], np.int64)
    c_first = 1 + np.concatenate([[0], np.cumsum(sizes)[:-1]])
    row_chunk = np.repeat(np.arange(len(chunks)), sizes)
    t = np.arange(1, N + 1, dtype=np.int64)
    isnl = text[st - 1] == 10
    nt, nc, nv = t[isnl], row_chunk[isnl], st[isnl]
    same = np.concatenate([[False], nc[1:] == nc[:-1]])
    prev_t = np.concatenate([[0], nt[:-1]])
    prev_v = np.concatenate([[0], nv[:-1]])
    first = np.where(same, prev_t + 1, c_first[nc])
    begin = np.where(same, prev_v, c_off[nc])
    term = np.stack([begin, nv, first, nt, nc, codes[nc]], axis=1)
    # chunks whose last row is not a '\n': one unterminated record each
    last_nl = np.searchsorted(nc, np.arange(len(chunks)), side="right") - 1
    has_nl = (last_nl >= 0) & (nc[np.maximum(last_nl, 0)] == np.arange(len(chunks)))
    c_last = c_first + sizes - 1
    tail = (sizes > 0) & ~(has_nl & (nt[np.maximum(last_nl, 0)] == c_last))
    cs = np.flatnonzero(tail)
    lt = np.where(has_nl[cs], nt[np.maximum(last_nl[cs], 0)] + 1, c_first[cs])
    lb = np.where(has_nl[cs], nv[np.maximum(last_nl[cs], 0)], c_off[cs])
    unterm = np.stack([lb, c_off[cs] + c_size[cs], lt, c_last[cs], cs, codes[cs]], axis=1)
    recs = np.concatenate([term, unterm])
    return recs[np.lexsort((recs[:, 2], recs[:, 4]))]


def test_lines_first_launch_full_size(ctx):
    data = gen.lines_parallel(8_000_000)
    offs, sizes = cj.chunk_lines(data, 1, 1 << 20)
    acc, b = 0, 0
    while b < len(sizes) and acc + sizes[b] <= (1 << 31) - 1:
        acc += sizes[b]
        b += 1
    lo = offs[0]
    part = data[lo:offs[b - 1] + sizes[b - 1]].copy()
    chunks = [(o - lo, s) for o, s in zip(offs[:b], sizes[:b])]
    del data
    assert b > 2000 and part.size > 2_000_000_000
    buf, aoffs = autotune.aligned_layout(part, [c[0] for c in chunks], [c[1] for c in chunks])
    dev = DeviceBuffer(buf.size)
    dev.upload(buf)
    del buf
    cap = part.size // 3 + 256 * b + 64
    max_rec = int(np.count_nonzero(part == 10)) + b
    st, pr = DeviceBuffer(cap * 4), DeviceBuffer(cap * 4)
    rec = DeviceBuffer(max_rec * cj.RECORD_DTYPE.itemsize)
    assert ctx.parse_lines_device(dev.ptr, aoffs, [c[1] for c in chunks], st.ptr, pr.ptr, cap,
                                  rec.ptr, max_rec) == 0
    rc, docs, N, M, over = ctx.batch_wait()
    R = ctx.last_records
    o = oracle.parse_lines(part, chunks)
    assert o.code == oracle.OK and rc == 0 and not over, (o.code, rc, over)
    assert N == o.N, (N, o.N)
    assert np.array_equal(docs[:, 2] - docs[:, 1], o.result_sizes)
    srow = st.download(np.int32, cap)
    idx = np.concatenate([np.arange(b0, e0) for b0, e0 in docs[:, 1:3]])
    got, want = srow[idx], o.buf[1:N + 1]
    assert np.array_equal(got, want), "structural: " + _first_diff(got, want)
    del srow
    prow = pr.download(np.int32, cap)
    got, want = prow[idx], o.buf[N + 2:2 * N + 2]
    assert np.array_equal(got, want), "pair: " + _first_diff(got, want)
    del prow, got
    want = expected_records(part, chunks, o, np.zeros(len(chunks), np.int64))
    assert R == want.shape[0], (R, want.shape)
    g = np.frombuffer(rec.download(np.uint8, R * cj.RECORD_DTYPE.itemsize).tobytes(),
                      dtype=cj.RECORD_DTYPE)
    got = np.stack([g[f].astype(np.int64) for f in
                    ("byte_begin", "byte_end", "struct_first", "struct_last", "chunk", "error")],
                   axis=1)
    if not np.array_equal(got, want):
        bad = np.flatnonzero((got != want).any(axis=1))
        raise AssertionError(f"{bad.size} records differ, first {bad[:4]}: {got[bad[:2]]} "
                             f"vs {want[bad[:2]]}")


def test_server_large_request():
    """One request of the serving mix's 64-256 MB class (200 MiB) with small
    requests around it; every completion equals the oracle's result."""
    big = gen.standard_doc(target_bytes=200 << 20, seed=gen.SEED_BASE + 55)
    small = [gen.standard_doc(target_bytes=t, seed=gen.SEED_BASE + 60 + i)
             for i, t in enumerate((2 << 10, 9 << 10, 300 << 10, 3 << 20))]
    docs = small[:2] + [big] + small[2:]
    out, tickets = {}, {}
    with cj.Server(budget=64 << 30, slots=4, max_batch=4096) as srv:
        for i, d in enumerate(docs):
            tickets[srv.submit(d)] = i
        while len(out) < len(docs):
            got = srv.completions(timeout_ms=20000)
            assert got, f"no completion within 20 s ({len(out)} of {len(docs)} done)"
            for ticket, err, buf, lat, q in got:
                out[tickets[ticket]] = (err, buf)
    for i, d in enumerate(docs):
        want = oracle.parse_standard(d)
        err, buf = out[i]
        assert err == want.code == 0, (i, err, want.code)
        assert buf.size == want.buf.size and np.array_equal(buf, want.buf), \
            f"request {i} ({d.size} B): " + _first_diff(buf, want.buf)
