"""Request server (continuous batching): every request's result is the
oracle's for that document alone, and identical whether the request ran
alone or batched with a mix of other requests (sizes, faults, flags)."""
import threading

import numpy as np
import pytest

import cujson_amd as cj
import oracle
from cujson_amd import gen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mix():
    data, offs, sizes, kinds = gen.many_docs(n_docs=300, max_size=1 << 20, fault_frac=0.2,
                                             deep_frac=0.1, seed=21)
    docs = [data[o:o + s].copy() for o, s in zip(offs, sizes)]
    # dense documents exceed the guessed output region: their batch is re-run
    docs += [np.frombuffer(b"[" + b"1," * 40000 + b"1]", np.uint8).copy(),
             np.frombuffer(b"[" * 3000 + b"]" * 3000, np.uint8).copy()]
    docs += [np.frombuffer(d, np.uint8).copy() for d in
             (b"[]", b"1", b"][", b"[}", b'{"a":"\\\\"}', b"\xc3\xa9", b"[" * 257 + b"]" * 257)]
    flags = [cj.CJX_FLAG_DEPTH_U16 if i % 5 == 3 else 0 for i in range(len(docs))]
    return docs, flags


def run_server(docs, flags, threads=1, reserve=0, **kw):
    """Submits every document (from `threads` threads, interleaved) and
    returns {index: (error, buf)} plus the server statistics."""
    out, tickets = {}, {}
    lock = threading.Lock()
    with cj.Server(**kw) as srv:
        if reserve:
            srv.reserve(reserve)
        def submit(part):
            for i in part:
                t = srv.submit(docs[i], flags[i])
                with lock:
                    tickets[t] = i
        parts = [range(k, len(docs), threads) for k in range(threads)]
        ths = [threading.Thread(target=submit, args=(p,)) for p in parts]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        while len(out) < len(docs):
            got = srv.completions(timeout_ms=2000)
            assert got, f"no completion within 2 s ({len(out)} of {len(docs)} done)"
            for ticket, err, buf, lat, q in got:
                assert lat >= q >= 0
                out[tickets[ticket]] = (err, buf)
        stats = srv.stats()
    return out, stats


def check_oracle(docs, flags, out, label):
    for i, d in enumerate(docs):
        want = oracle.parse_standard(d.tobytes(), flags[i])
        err, buf = out[i]
        assert err == want.code, (label, i, err, want.code)
        if not np.array_equal(buf, want.buf):
            bad = np.flatnonzero(buf != want.buf) if buf.size == want.buf.size else []
            raise AssertionError(f"{label} doc {i} ({d.size} B): cells differ {bad[:6]}")


def test_server_mix_matches_oracle(mix):
    docs, flags = mix
    out, st = run_server(docs, flags, threads=2, budget=1 << 30, slots=3, max_batch=64)
    check_oracle(docs, flags, out, "mix")
    assert st["documents"] == len(docs)
    assert st["batches"] < len(docs)  # requests were batched
    assert st["relaunches"] >= 1      # the dense documents overflowed their guess


def test_server_alone_equals_in_mix(mix):
    docs, flags = mix
    mixed, _ = run_server(docs, flags, threads=2, budget=1 << 30, slots=4, max_batch=4096)
    alone, st = run_server(docs, flags, threads=1, budget=1 << 30, slots=1, max_batch=1)
    assert st["batches"] == len(docs)
    for i in range(len(docs)):
        assert mixed[i][0] == alone[i][0], i
        assert np.array_equal(mixed[i][1], alone[i][1]), i


def test_server_small_budget_runs_large_requests_alone(mix):
    docs, flags = mix
    big = [i for i, d in enumerate(docs) if d.size > 64 << 10][:20]
    sub = [docs[i] for i in big]
    fl = [flags[i] for i in big]
    out, st = run_server(sub, fl, budget=2 << 20, slots=2, max_batch=4096)
    check_oracle(sub, fl, out, "small budget")
    assert st["batches"] == len(sub)


def test_server_matches_plain_parse(mix):
    docs, flags = mix
    out, _ = run_server(docs[:60], flags[:60], budget=1 << 30, slots=2, max_batch=16)
    ctx = cj.Context(0)
    try:
        for i, d in enumerate(docs[:60]):
            r = cj.parse_standard_json(d, flags[i], ctx)
            assert out[i][0] == r.error, i
            if r.error == 0:
                assert np.array_equal(out[i][1], r.buf), i
    finally:
        ctx.close()


def test_server_large_batches_take_the_copy_engine_path():
    """Batches above the kernel-copy limit (16 MiB of input or of results) go
    through the copy engines: large inputs copied from the caller's memory,
    small ones staged, results laid out on the device and copied out."""
    docs = [gen.standard_doc(target_bytes=t, seed=100 + i)
            for i, t in enumerate((3 << 20, 22 << 20, 20 << 10, 6 << 20, 9 << 10, 7 << 20))]
    docs.append(docs[0].copy())
    assert gen.plant_fault(docs[-1], 3, 7)
    docs += [np.frombuffer(b'{"k":[1,2,{"z":null}]}', np.uint8).copy()] * 3
    flags = [0] * len(docs)
    out, st = run_server(docs, flags, budget=16 << 30, slots=2, max_batch=4096)
    check_oracle(docs, flags, out, "large batch")
    assert st["batches"] < len(docs)


def test_server_reserved_buffers_serve_the_mix(mix):
    docs, flags = mix
    out, st = run_server(docs, flags, threads=2, reserve=max(d.size for d in docs),
                         budget=4 << 30, slots=4, max_batch=256)
    check_oracle(docs, flags, out, "reserved")
    with cj.Server(budget=1 << 30, slots=2, max_batch=8) as srv:
        srv.submit(docs[0])
        with pytest.raises(RuntimeError):
            srv.reserve(1 << 20)  # only before the first request
        assert srv.completions(timeout_ms=2000)
