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
