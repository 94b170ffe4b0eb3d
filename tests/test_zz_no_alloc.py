"""cjx_parse_device_noalloc (no allocation: every working buffer in the
library's own device array) against cjx_parse_device, the existing entry
whose kernels it shares, and against the oracle: rows, pair cells, counts and
codes bit for bit, over consecutive calls whose second input is smaller
(the array keeps the previous call's contents), on the null stream and on a
context's stream."""
import numpy as np
import pytest

import cujson_amd as cj
import oracle
from cujson_amd import gen
from cujson_amd.hip import DeviceBuffer, synchronize

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = cj.Context(0)
    yield c
    c.close()


def _raw(doc):
    return doc.tobytes() if isinstance(doc, np.ndarray) else bytes(doc)


def _device_doc(doc):
    d = np.frombuffer(_raw(doc), np.uint8)
    buf = DeviceBuffer(((d.size + 63) // 64) * 64)
    buf.upload(d)
    return buf, d.size


def run_noalloc(dev, n, cap, flags=0, stream=0):
    st, pr, status = DeviceBuffer(cap * 4 + 64), DeviceBuffer(cap * 4 + 64), DeviceBuffer(64)
    st.fill(0x7F)
    pr.fill(0x7F)
    status.fill(0x55)
    synchronize()
    rc = cj.parse_device_noalloc(dev.ptr, n, st.ptr, pr.ptr, cap, status.ptr, flags, stream)
    assert rc == cj.CJX_OK, cj.strerror(rc)
    synchronize()
    s = np.frombuffer(status.download(np.uint8, cj.DEVICE_STATUS_DTYPE.itemsize).tobytes(),
                      cj.DEVICE_STATUS_DTYPE)[0]
    N = int(s["n_structural"])
    rows = min(N, cap)
    tail = st.download(np.uint8, 64, cap * 4), pr.download(np.uint8, 64, cap * 4)
    return (int(s["error"]), N, int(s["n_brackets"]), bool(s["capacity_exceeded"]),
            st.download(np.int32, rows), pr.download(np.int32, rows), tail)


def run_context(ctx, dev, n, cap, flags=0):
    st, pr = DeviceBuffer(cap * 4 + 64), DeviceBuffer(cap * 4 + 64)
    assert ctx.parse_device(dev.ptr, n, st.ptr, pr.ptr, cap, flags) == 0
    rc, N, M, over = ctx.wait()
    rows = min(N, cap)
    return rc, N, M, over, st.download(np.int32, rows), pr.download(np.int32, rows)


def check_same(ctx, doc, flags=0, cap=None, stream=0, label=""):
    dev, n = _device_doc(doc)
    cap = n if cap is None else cap
    a = run_noalloc(dev, n, cap, flags, stream)
    b = run_context(ctx, dev, n, cap, flags)
    assert a[:4] == b[:4], (label, a[:4], b[:4])
    assert np.array_equal(a[4], b[4]) and np.array_equal(a[5], b[5]), label
    # nothing is written past the caller's capacity
    assert np.all(a[6][0] == 0x7F) and np.all(a[6][1] == 0x7F), label
    if not a[3]:
        o = oracle.parse_standard(_raw(doc), flags)
        assert a[0] == o.code and a[1] == o.N, (label, a[0], o.code, a[1], o.N)
        assert np.array_equal(a[4], o.structural[1:o.N + 1]), label
        assert np.array_equal(a[5], o.pair_pos[1:o.N + 1]), label
    return a


def test_two_calls_second_smaller(ctx):
    big = gen.config1()  # about 0.89 MB: tiles, look-back, chunk matching
    assert big.size <= cj.noalloc_max_bytes()
    check_same(ctx, big, label="config1")
    check_same(ctx, big[:5000], label="second, smaller (one wave)")
    check_same(ctx, gen.standard_doc(target_bytes=200 << 10, seed=77), label="third, 200 KiB")


def test_cases_against_existing_entry(ctx):
    base = gen.config1()
    cases = [
        (b"[]", 0), (b'{"a":[1,2,{"b":"\\\\\\""}]}', 0), (bytes(base[:16384]), 0),
        (bytes(base[:16385]), 0), (bytes(base[:65537]), 0),
        (b"[" * 300 + b"1" + b"]" * 300, 0),                    # sort path (depth > 256)
        (b"[" * 300 + b"1" + b"]" * 300, cj.CJX_FLAG_DEPTH_U16),
        (b"[}" * 10, 0), (b'["\xc0\xaf"]', 0), (b'["\xc0\xaf"]', cj.CJX_FLAG_UTF8_STRICT),
        (b'["\xe2\x82"]', 0),
    ]
    for i, (d, fl) in enumerate(cases):
        check_same(ctx, d, fl, label=f"case {i}")
    for kind in range(5):
        d = base.copy()
        assert gen.plant_fault(d, kind, 300 + kind)
        check_same(ctx, d, label=f"fault {kind}")
        check_same(ctx, d, cj.CJX_FLAG_UTF8_STRICT, label=f"fault {kind} strict")


def test_capacity_overflow_and_stream(ctx):
    d = gen.config1()
    a = check_same(ctx, d, cap=1000, label="small capacity")
    assert a[3] and a[0] == cj.CJX_ERR_CAPACITY
    check_same(ctx, d, stream=ctx.stream, label="on the context stream")


def test_too_large_launches_nothing(ctx):
    n = cj.noalloc_max_bytes() + 1
    dev = DeviceBuffer(n + 64)
    st, pr, status = DeviceBuffer(64), DeviceBuffer(64), DeviceBuffer(64)
    status.fill(0x55)
    assert cj.parse_device_noalloc(dev.ptr, n, st.ptr, pr.ptr, n, status.ptr) == \
        cj.CJX_ERR_TOO_LARGE
    synchronize()
    assert np.all(status.download(np.uint8, 32) == 0x55)
    # the next valid call is unaffected
    check_same(ctx, gen.config1()[:3000], label="after a refused call")
