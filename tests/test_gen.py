"""The synthetic workload generator: determinism and validity (no GPU)."""
import hashlib
import json

import numpy as np

import oracle
from cujson_amd import gen


def test_config1_deterministic_and_valid():
    a = gen.config1()
    b = gen.config1()
    assert np.array_equal(a, b)
    assert 800_000 < a.size < 1_200_000
    doc = json.loads(bytes(a))
    assert len(doc) == 2000
    assert len(doc[0]) == 12
    assert doc[0]["user"]["meta"]["tags"]["k"] is not None  # nested depth 3
    assert int((a >= 0x80).sum()) > 0  # multi-byte UTF-8 present
    assert b"\\u" in bytes(a[:100000]) and b'\\"' in bytes(a[:100000])


def test_config2_shape_at_small_scale():
    d = gen.config2(target_bytes=24 << 20)
    assert (24 << 20) - 8192 < d.size <= 24 << 20
    r = oracle.parse_standard(d)
    assert r.code == 0
    # the deep branch reaches depth 64 + the top-level array
    ch = d[r.structural[1:r.N + 1] - 1]
    depth = np.cumsum(np.where((ch == ord("[")) | (ch == ord("{")), 1,
                               np.where((ch == ord("]")) | (ch == ord("}")), -1, 0)))
    assert depth.max() >= 65
    # at least 10% of 1 MiB chunk boundaries fall inside a string
    bs = oracle.bitmaps(d)
    instr = bs["instr"]
    inside = 0
    nchunks = d.size >> 20
    for k in range(1, nchunks + 1):
        p = (k << 20) - 1
        if p < d.size:
            inside += (int(instr[p >> 5]) >> (p & 31)) & 1
    assert inside >= 0.10 * nchunks


def test_lines_generator():
    d = gen.lines(500)
    recs = bytes(d).split(b"\n")
    assert recs[-1] == b""
    assert len(recs) == 501
    sizes = [len(r) + 1 for r in recs[:-1]]
    for r in recs[:-1]:
        json.loads(r)
    assert 200 <= np.median(sizes) <= 800


def test_fault_planting():
    base = gen.config1()
    for kind, code in [(2, oracle.ERR_UTF8), (3, oracle.ERR_STRUCTURE)]:
        d = base.copy()
        assert gen.plant_fault(d, kind, 42)
        assert oracle.parse_standard(d).code == code
    d = base.copy()
    assert gen.plant_fault(d, 0, 42)  # overlong: not flagged by the reference rule
    assert oracle.parse_standard(d).code == oracle.OK
    assert oracle.utf8_check(d, oracle.FLAG_UTF8_STRICT) == 1
