"""Deterministic synthetic workloads (the five benchmark configurations).

Thin ctypes front end over ``libcjx_gen.so`` (built from ``csrc/cjx_gen.c``
by ``cujson_amd/build.py``).  Seeds follow ``0xC0FFEE00 + config_id``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
GEN_PATH = os.path.join(_HERE, "libcjx_gen.so")
SEED_BASE = 0xC0FFEE00

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(GEN_PATH):
            raise RuntimeError(f"{GEN_PATH} is missing: build with `python -m cujson_amd.build`")
        L = ctypes.CDLL(GEN_PATH)
        u8p = ctypes.c_void_p
        L.cjx_gen_standard.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                       ctypes.c_int, ctypes.c_uint64, ctypes.c_int, u8p,
                                       ctypes.c_size_t]
        L.cjx_gen_standard.restype = ctypes.c_size_t
        L.cjx_gen_record.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, u8p,
                                     ctypes.c_size_t]
        L.cjx_gen_record.restype = ctypes.c_size_t
        L.cjx_gen_lines.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_uint64, ctypes.c_uint64, u8p,
                                    ctypes.c_size_t]
        L.cjx_gen_lines.restype = ctypes.c_size_t
        L.cjx_gen_standard_segment.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                               ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64,
                                               ctypes.c_int, u8p]
        L.cjx_gen_standard_segment.restype = ctypes.c_size_t
        L.cjx_plant_fault.argtypes = [ctypes.c_uint64, ctypes.c_int, u8p, ctypes.c_size_t]
        L.cjx_plant_fault.restype = ctypes.c_int
        _lib = L
    return _lib


def standard_doc(n_objects=0, target_bytes=0, deep_depth=0, chunk_bytes=0, boundary_every=0,
                 seed=SEED_BASE + 1, out=None) -> np.ndarray:
    """Top-level array of post-like objects (configs 1 and 2)."""
    cap = target_bytes + (1 << 20) if target_bytes else max(n_objects, 1) * 4096 + (1 << 20)
    if out is None:
        out = np.empty(cap, dtype=np.uint8)
    n = lib().cjx_gen_standard(seed, n_objects, target_bytes, deep_depth, chunk_bytes,
                               boundary_every, out.ctypes.data, out.size)
    if n == 0:
        raise RuntimeError("generator buffer too small")
    return out[:n]


def config1(seed=SEED_BASE + 1) -> np.ndarray:
    """~1 MB: 2,000 post-like objects in a top-level array."""
    return standard_doc(n_objects=2000, seed=seed)


def config2(target_bytes=2_000_000_000, seed=SEED_BASE + 2, out=None) -> np.ndarray:
    """~2 GB document with a depth-64 branch; every 8th 1 MiB boundary inside a string."""
    return standard_doc(target_bytes=target_bytes, deep_depth=64, chunk_bytes=1 << 20,
                        boundary_every=8, seed=seed, out=out)


def config2_parallel(total_bytes=2_000_000_000, seed=SEED_BASE + 2, segments=64,
                     threads=None) -> np.ndarray:
    """Config-2 document of exactly total_bytes built from fixed-size segments
    generated in parallel ('[' + segments + ']', segments space-padded)."""
    from concurrent.futures import ThreadPoolExecutor
    body = total_bytes - 2
    seg = body // segments
    sizes = [seg] * segments
    sizes[-1] += body - seg * segments
    out = np.empty(total_bytes, dtype=np.uint8)
    out[0] = ord("[")
    out[-1] = ord("]")
    starts = np.cumsum([1] + sizes[:-1]).tolist()
    L = lib()

    def work(k):
        p = out.ctypes.data + starts[k]
        n = L.cjx_gen_standard_segment(seed, k, starts[k], sizes[k], 64, 1 << 20, 8, p)
        if n != sizes[k]:
            raise RuntimeError("segment generation failed")

    threads = threads or min(16, os.cpu_count() or 4)
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(work, range(segments)))
    return out


def lines(n_records, median=400.0, sigma=1.2, max_size=64 * 1024, seed=SEED_BASE + 3):
    """JSON Lines records with log-normal sizes (config 3 shape)."""
    cap = int(n_records * max(median * 8, 2048)) + (1 << 20)
    out = np.empty(cap, dtype=np.uint8)
    n = lib().cjx_gen_lines(seed, n_records, median, sigma, 200, max_size, out.ctypes.data, cap)
    if n == 0:
        raise RuntimeError("generator buffer too small")
    return out[:n]


def plant_fault(doc: np.ndarray, kind: int, seed: int) -> bool:
    """Plant one fault in place (0 overlong, 1 surrogate, 2 truncated UTF-8,
    3 unbalanced bracket, 4 unterminated string)."""
    return bool(lib().cjx_plant_fault(seed, kind, doc.ctypes.data, doc.size))


def many_docs(n_docs=50_000, alpha=1.1, min_size=200, max_size=16 << 20, fault_frac=0.10,
              deep_frac=0.05, max_depth=512, align=16, seed=SEED_BASE + 4):
    """Batch of independent standard-JSON documents (config 4 shape).

    Sizes are Pareto(alpha) from min_size, clipped to max_size.  A deep_frac
    share of documents is wrapped in alternating arrays/objects of depth
    uniform in [65, max_depth]; a fault_frac share carries exactly one fault
    (plant_fault kinds 0-4 in turn).  Documents start at multiples of `align`
    in the returned buffer (gaps are zero bytes).

    Returns (data, offsets, sizes, kinds) with kinds[i] = fault kind or -1.
    """
    rng = np.random.default_rng(seed)
    u = rng.random(n_docs)
    sizes_t = np.minimum(min_size / np.power(1.0 - u, 1.0 / alpha), max_size).astype(np.int64)
    deep = rng.random(n_docs) < deep_frac
    depths = rng.integers(65, max_depth + 1, n_docs)
    faulty = rng.random(n_docs) < fault_frac
    seeds = rng.integers(0, 2 ** 62, n_docs)
    L = lib()
    docs, kinds = [], []
    scratch = np.empty(max_size + (1 << 20), dtype=np.uint8)
    k_fault = 0
    for i in range(n_docs):
        t = int(sizes_t[i])
        if t < 8192:
            n = L.cjx_gen_record(int(seeds[i]), t, int(seeds[i] % 6), scratch.ctypes.data,
                                 scratch.size)
            d = scratch[:n - 1].copy()  # drop the record's '\n'
        else:
            n = L.cjx_gen_standard(int(seeds[i]), 0, t, 0, 0, 0, scratch.ctypes.data, scratch.size)
            d = scratch[:n].copy()
        if deep[i]:
            k = int(depths[i])
            pre = b"".join(b"[" if j % 2 == 0 else b'{"d":' for j in range(k))
            post = b"".join(b"]" if j % 2 == 0 else b"}" for j in reversed(range(k)))
            d = np.concatenate([np.frombuffer(pre, np.uint8), d, np.frombuffer(post, np.uint8)])
        kind = -1
        if faulty[i]:
            kind = k_fault % 5
            k_fault += 1
            if not plant_fault(d, kind, int(seeds[i]) ^ 0x5A5A):
                kind = -1
        docs.append(d)
        kinds.append(kind)
    sizes = np.array([d.size for d in docs], dtype=np.int64)
    offs = np.zeros(n_docs, dtype=np.int64)
    pos = 0
    for i in range(n_docs):
        offs[i] = pos
        pos += -(-int(sizes[i]) // align) * align
    data = np.zeros(pos + 64, dtype=np.uint8)
    for i in range(n_docs):
        data[offs[i]:offs[i] + sizes[i]] = docs[i]
    return data, offs, sizes, np.array(kinds, dtype=np.int64)


def lines_parallel(n_records, median=400.0, sigma=1.2, max_size=64 * 1024,
                   seed=SEED_BASE + 3, blocks=64, threads=None) -> np.ndarray:
    """lines() for large record counts: `blocks` independent record streams
    (seeds seed + k) generated on several threads and concatenated."""
    from concurrent.futures import ThreadPoolExecutor
    per = [n_records // blocks + (1 if k < n_records % blocks else 0) for k in range(blocks)]
    L = lib()

    def work(k):
        cap = int(per[k] * max(median * 8, 2048)) + (1 << 20)
        out = np.empty(cap, dtype=np.uint8)
        n = L.cjx_gen_lines(seed + 1000 + k, per[k], median, sigma, 200, max_size,
                            out.ctypes.data, cap)
        if n == 0 and per[k]:
            raise RuntimeError("generator buffer too small")
        return out[:n]

    threads = threads or min(16, os.cpu_count() or 4)
    with ThreadPoolExecutor(threads) as ex:
        parts = list(ex.map(work, range(blocks)))
    return np.concatenate(parts)
