"""Every tokenizer launch shape the autotuner can pick gives the oracle's
output, bit for bit: tiles of 2, 4, 8 and 16 waves (32, 16, 8 and 4 KiB per
wave, so the wave edges inside a tile move) and one-wave documents packed 1,
2, 4 or 8 per workgroup (ragged last workgroups)."""
import random

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


def check_standard(ctx, data, flags=0, label=""):
    data = bytes(data) if not isinstance(data, np.ndarray) else data
    g = cj.parse_standard_json(data, flags, ctx)
    o = oracle.parse_standard(data, flags)
    assert g.error == o.code, (label, g.error, o.code)
    if o.code == oracle.ERR_INPUT:
        return
    assert g.N == o.N, (label, g.N, o.N)
    if not np.array_equal(g.buf, o.buf):
        bad = np.flatnonzero(g.buf != o.buf)
        raise AssertionError(f"{label}: {bad.size} cells differ, first {bad[:8]}")


def test_default_shape(ctx):
    assert ctx.tokenizer_shape() == (4, 4)
    with pytest.raises(ValueError):
        ctx.set_tokenizer_shape(3, 4)
    with pytest.raises(ValueError):
        ctx.set_tokenizer_shape(4, 16)
    assert ctx.tokenizer_shape() == (4, 4)


@pytest.mark.parametrize("tw", autotune.TILE_WAVES)
def test_tile_shapes_edges(ctx, tw):
    ctx.set_tokenizer_shape(tw, 4)
    try:
        wave = 65536 // tw
        base = bytes(gen.config1()[:400000])
        for n in (wave - 1, wave, wave + 1, 2 * wave + 3, 65535, 65536, 65537, 3 * 65536 + wave,
                  393216 + 17):
            check_standard(ctx, base[:n], label=f"tw={tw} prefix n={n}")
            check_standard(ctx, base[:n], oracle.FLAG_UTF8_STRICT, label=f"tw={tw} strict n={n}")
        # backslash runs and strings across every wave edge of the tile
        for edge in sorted({wave, 2 * wave, 65536 - wave, 65536, 65536 + wave}):
            for k in (1, 2, 3, 31, 32, 33):
                for shift in (-1, 0, 1):
                    pre = b'["' + b"x" * (edge - 2 - k + shift)
                    check_standard(ctx, pre + b"\\" * k + b'"]",{"a":[1]}]',
                                   label=f"tw={tw} edge={edge} k={k} s={shift}")
        for k in (wave - 1, wave, wave + 1, 65537):
            check_standard(ctx, b'["' + b"\\" * k + b'"]",[1]]', label=f"tw={tw} allbs k={k}")
        doc = b'[{"a":"' + (b"{[:,]}" * (70000 // 6 + 1))[:70000] + b'"},[1,2,{"b":[]}]]'
        check_standard(ctx, doc, label=f"tw={tw} long string")
        rng = random.Random(tw)
        alphabet = [ord(c) for c in '{}[]:,"\\a \n'] + [0xC3, 0xA9, 0xE2, 0x82, 0xAC, 0xF0]
        for it in range(6):
            data = bytes(rng.choice(alphabet) for _ in range(rng.choice([20000, 70000, 200000])))
            check_standard(ctx, data, label=f"tw={tw} soup {it}")
        for d in (255, 256, 1000):
            check_standard(ctx, b"[" * d + b"1" + b"]" * d, oracle.FLAG_DEPTH_U16,
                           label=f"tw={tw} depth {d}")
    finally:
        ctx.set_tokenizer_shape(4, 4)


@pytest.mark.parametrize("tw", autotune.TILE_WAVES)
def test_tile_shapes_lines_and_large(ctx, twitter_small, tw):
    ctx.set_tokenizer_shape(tw, 4)
    try:
        d = bytes(twitter_small)
        chunks = oracle.chunks_by_count(d, 4)
        inp = cj.cuJSONLinesInput(np.frombuffer(d, np.uint8), len(chunks), len(d),
                                  [c[0] for c in chunks], [c[1] for c in chunks])
        g = cj.parse_json_lines(inp, 0, ctx)
        o = oracle.parse_lines(d, chunks)
        assert g.error == o.code == 0 and g.N == o.N and np.array_equal(g.buf, o.buf)
        # 70 MB from host memory: the tokenizer runs per copied part
        big = gen.config2_parallel(total_bytes=70 << 20)
        check_standard(ctx, big, label=f"tw={tw} chunked host")
    finally:
        ctx.set_tokenizer_shape(4, 4)


@pytest.mark.parametrize("sw", autotune.SMALL_WAVES)
def test_small_shapes_batch(ctx, sw):
    ctx.set_tokenizer_shape(4, sw)
    try:
        data, offs, sizes, kinds = gen.many_docs(n_docs=333, max_size=100 << 10, fault_frac=0.2,
                                                 deep_frac=0.1, seed=50 + sw)
        for flags in (0, oracle.FLAG_UTF8_STRICT):
            r = cj.parse_batch(data, offs.tolist(), sizes.tolist(), flags, ctx)
            assert r.rc == 0
            for i, (o, s) in enumerate(zip(offs, sizes)):
                want = oracle.parse_standard(data[o:o + s].tobytes(), flags)
                err, buf = r.doc(i)
                assert err == want.code, (sw, flags, i, err, want.code)
                assert np.array_equal(buf, want.buf), (sw, flags, i)
        # one document alone in the batch, and one more than a workgroup's share
        for count in (1, sw + 1):
            r = cj.parse_batch(data, offs[:count].tolist(), sizes[:count].tolist(), 0, ctx)
            for i in range(count):
                want = oracle.parse_standard(data[offs[i]:offs[i] + sizes[i]].tobytes())
                err, buf = r.doc(i)
                assert err == want.code and np.array_equal(buf, want.buf), (sw, count, i)
    finally:
        ctx.set_tokenizer_shape(4, 4)


def test_graph_recaptured_on_shape_change(ctx):
    """The HIP graph of a device parse is keyed by the tokenizer shape: a
    replay after a shape change runs the new shape's kernels, same output."""
    d = gen.config1()
    n = d.size
    dev = DeviceBuffer(((n + 63) // 64) * 64)
    dev.upload(d)
    cap = n + 64
    st, pr = DeviceBuffer(cap * 4), DeviceBuffer(cap * 4)
    o = oracle.parse_standard(d)
    ref = None
    try:
        for tw in (4, 8, 2, 16, 4):
            ctx.set_tokenizer_shape(tw, 4)
            for _ in range(2):
                assert ctx.parse_device(dev.ptr, n, st.ptr, pr.ptr, cap, cj.CJX_FLAG_GRAPH) == 0
                rc, N, M, over = ctx.wait()
                assert rc == 0 and N == o.N and not over
                got = (st.download(np.int32, N), pr.download(np.int32, N))
                if ref is None:
                    ref = got
                assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), tw
        assert np.array_equal(ref[0], o.structural[1:N + 1])
        assert np.array_equal(ref[1], o.pair_pos[1:N + 1])
    finally:
        ctx.set_tokenizer_shape(4, 4)


def test_autotune_picks_a_measured_shape(ctx):
    d = gen.config1()
    n = d.size
    dev = DeviceBuffer(((n + 63) // 64) * 64)
    dev.upload(d)
    cap = n + 64
    st, pr = DeviceBuffer(cap * 4), DeviceBuffer(cap * 4)

    def run():
        assert ctx.parse_device(dev.ptr, n, st.ptr, pr.ptr, cap) == 0

    try:
        best, table = autotune.tune_tokenizer_shape(ctx, run, "test-config1", reps=3,
                                                    use_cache=False)
        assert f"{best[0]}x{best[1]}" in table
        assert set(table) >= {f"{t}x4" for t in autotune.TILE_WAVES}
        assert ctx.tokenizer_shape() == best
        assert table[f"{best[0]}x{best[1]}"] == min(table.values())
    finally:
        ctx.set_tokenizer_shape(4, 4)
