"""CPU oracle for cuJSON's three-phase parse.

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` -- never by the product package
``cujson_amd``.  ``liboracle.so`` is built from ``oracle.c`` by
``oracle/Makefile``; ``_ref/libsimdjson_ref.so`` is simdjson 3.2.1 compiled
from the reference's vendored sources.

The Python functions below add the host-side steps of the reference around
the C core:

* ``parse_standard``  parse_standard_json, cujson/parse_standard_json.cu:1584-1687
* ``parse_lines``     parse_json_lines,    cujson/parse_json_lines.cu:1092-1264
* ``chunks_by_count`` / ``chunks_by_bytes`` the JSONL chunking policies of
  cujson/load_file.cu:38-107 and :109-180.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

OK, ERR_INPUT, ERR_UTF8, ERR_STRUCTURE = 0, 1, 2, 3
FLAG_UTF8_STRICT, FLAG_DEPTH_U16, FLAG_JSONL = 0x1, 0x2, 0x4

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run make -C oracle)")
        L = ctypes.CDLL(path)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        i32p = ctypes.POINTER(ctypes.c_int32)
        u64p = ctypes.POINTER(ctypes.c_uint64)
        L.orc_parse_unit.argtypes = [u8p, ctypes.c_size_t, ctypes.c_uint, ctypes.c_uint64,
                                     ctypes.c_uint64, u32p, i32p, u64p, u64p,
                                     ctypes.POINTER(ctypes.c_int)]
        L.orc_parse_unit.restype = ctypes.c_int
        L.orc_utf8_check.argtypes = [u8p, ctypes.c_size_t, ctypes.c_uint]
        L.orc_utf8_check.restype = ctypes.c_int
        L.orc_utf8_valid_strict.argtypes = [u8p, ctypes.c_size_t]
        L.orc_utf8_valid_strict.restype = ctypes.c_int
        L.orc_bitmaps.argtypes = [u8p, ctypes.c_size_t, ctypes.c_uint] + [u32p] * 8
        L.orc_bitmaps.restype = None
        L.orc_pair.argtypes = [u8p, u32p, ctypes.c_uint64, ctypes.c_uint, ctypes.c_uint64, i32p,
                               u32p, u32p]
        L.orc_pair.restype = ctypes.c_int
        for nm in ("orc_vcmpeq4", "orc_vcmpltu4", "orc_vcmpgeu4", "orc_vcmpgtu4",
                   "orc_vsubus4", "orc_vsubss4"):
            f = getattr(L, nm)
            f.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
            f.restype = ctypes.c_uint32
        _lib = L
    return _lib


def ref_lib():
    """simdjson 3.2.1 built from the reference's vendored sources, or None."""
    global _ref
    if _ref is None:
        path = os.path.join(_HERE, "_ref", "libsimdjson_ref.so")
        if not os.path.exists(path):
            return None
        L = ctypes.CDLL(path)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        L.sjr_stage1.argtypes = [u8p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint32),
                                 ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        L.sjr_stage1.restype = ctypes.c_int
        L.sjr_validate_utf8.argtypes = [u8p, ctypes.c_size_t]
        L.sjr_validate_utf8.restype = ctypes.c_int
        L.sjr_iterate_ms.argtypes = [u8p, ctypes.c_size_t, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_int)]
        L.sjr_iterate_ms.restype = ctypes.c_double
        L.sjr_parse_many_ms.argtypes = [u8p, ctypes.c_size_t, ctypes.c_size_t,
                                        ctypes.POINTER(ctypes.c_size_t)]
        L.sjr_parse_many_ms.restype = ctypes.c_double
        szp = ctypes.POINTER(ctypes.c_size_t)
        L.sjr_iterate_docs_ms.argtypes = [u8p, szp, szp, ctypes.c_size_t, ctypes.c_int, szp]
        L.sjr_iterate_docs_ms.restype = ctypes.c_double
        _ref = L
    return _ref


def _u8(buf):
    a = np.frombuffer(buf, dtype=np.uint8) if not isinstance(buf, np.ndarray) else buf
    a = np.ascontiguousarray(a, dtype=np.uint8)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


@dataclass
class UnitResult:
    code: int
    structural: np.ndarray  # uint32[N]: 1-based positions (+ pos_base)
    pair: np.ndarray        # int32[N]: pair cells, -1 where never written
    M: int
    utf8_err: int


def parse_unit(data, flags=0, pos_base=0, struct_base=0) -> UnitResult:
    a, p = _u8(data)
    n = a.size
    cap = max(n, 1)
    st = np.zeros(cap, dtype=np.uint32)
    pr = np.zeros(cap, dtype=np.int32)
    N = ctypes.c_uint64()
    M = ctypes.c_uint64()
    ue = ctypes.c_int()
    code = lib().orc_parse_unit(p, n, flags, pos_base, struct_base,
                                st.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                pr.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                ctypes.byref(N), ctypes.byref(M), ctypes.byref(ue))
    return UnitResult(code, st[: N.value].copy(), pr[: N.value].copy(), M.value, ue.value)


def bitmaps(data, flags=0):
    a, p = _u8(data)
    W = (a.size + 31) // 32
    outs = [np.zeros(max(W, 1), dtype=np.uint32) for _ in range(8)]
    lib().orc_bitmaps(p, a.size, flags,
                      *[o.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)) for o in outs])
    names = ["bs", "quote", "op", "oc", "rq", "instr", "st", "ocm"]
    return {k: o[:W] for k, o in zip(names, outs)}


def utf8_check(data, flags=0) -> int:
    a, p = _u8(data)
    return lib().orc_utf8_check(p, a.size, flags)


def pair_phase(oc_char, oc_idx, N, flags=0, struct_base=0):
    """Phase 3 alone: returns (err, pair[N], keys[M], perm[M])."""
    oc_char = np.ascontiguousarray(oc_char, dtype=np.uint8)
    oc_idx = np.ascontiguousarray(oc_idx, dtype=np.uint32)
    M = oc_char.size
    pair = np.full(max(N, 1), -1, dtype=np.int32)
    keys = np.zeros(max(M, 1), dtype=np.uint32)
    perm = np.zeros(max(M, 1), dtype=np.uint32)
    err = lib().orc_pair(oc_char.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                         oc_idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), M, flags,
                         struct_base, pair.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                         keys.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                         perm.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    return err, pair[:N], keys[:M], perm[:M]


@dataclass
class ParseResult:
    """Host layout of cuJSONResult (cujson/cujson_types.h:22-33).

    ``buf`` is the single int32 buffer of 2(N+2) cells; ``structural`` is
    buf[0:N+2] and ``pair_pos`` is buf[N+1:2N+4] (PSJ:1659-1680)."""
    code: int
    buf: np.ndarray
    N: int
    chunk_count: int = 1
    result_sizes: list = field(default_factory=list)
    result_sizes_prefix: list = field(default_factory=list)

    @property
    def structural(self):
        return self.buf[: self.N + 2]

    @property
    def pair_pos(self):
        return self.buf[self.N + 1:]

    @property
    def total_result_size(self):
        return self.N + 2

    @property
    def file_size(self):
        return self.N + 2


def assemble(structural_rows, pair_rows):
    """Host layout assembly, PSJ:1659-1680 / PJL:1070-1089 with S12 sentinels."""
    st = np.concatenate(structural_rows) if structural_rows else np.zeros(0, np.uint32)
    pr = np.concatenate(pair_rows) if pair_rows else np.zeros(0, np.int32)
    N = st.size
    buf = np.full(2 * (N + 2), -1, dtype=np.int32)
    buf[0] = 0
    buf[1:N + 1] = st.view(np.int32)
    buf[N + 1] = N + 1
    buf[N + 2:2 * N + 2] = pr
    return buf, N


def parse_standard(data, flags=0) -> ParseResult:
    """parse_standard_json (PSJ:1584-1687) on the CPU."""
    a, _ = _u8(data)
    if a.size == 0:
        return ParseResult(ERR_INPUT, np.zeros(0, np.int32), 0, 0)
    r = parse_unit(a, flags & ~FLAG_JSONL)
    buf, N = assemble([r.structural], [r.pair])
    return ParseResult(r.code, buf, N, 1, [N], [N])


def chunks_by_count(data, chunk_count):
    """loadJSONLines_chunkCount line split, cujson/load_file.cu:71-103 -> [(start, size)]."""
    a, _ = _u8(data)
    nl = np.flatnonzero(a == 0x0A)
    offs = [0] + list((nl + 1).tolist())
    if offs[-1] < a.size:
        offs.append(a.size)
    total = len(offs) - 1
    per, extra = divmod(total, chunk_count)
    out, li = [], 0
    for i in range(chunk_count):
        k = per + (1 if i < extra else 0)
        out.append((offs[li], offs[li + k] - offs[li]))
        li += k
    return out


def chunks_by_bytes(data, chunk_bytes):
    """loadJSONLines_chunkSizeBytes greedy split, cujson/load_file.cu:137-176."""
    a, _ = _u8(data)
    nl = np.flatnonzero(a == 0x0A)
    offs = [0] + list((nl + 1).tolist())
    if offs[-1] < a.size:
        offs.append(a.size)
    out = []
    cstart, cur = 0, 0
    for i in range(1, len(offs)):
        ls, le = offs[i - 1], offs[i]
        if le - cstart > chunk_bytes:
            out.append((cstart, cur - cstart))
            cstart = ls
        cur = le
    if cstart < a.size:
        out.append((cstart, a.size - cstart))
    return out


def parse_lines(data, chunks, flags=0) -> ParseResult:
    """parse_json_lines (PJL:1092-1264) on the CPU; chunks = [(start, size)].

    Chunks run in order with lastChunkIndex / lastStructuralIndex carried
    (PJL:1222-1223); the first failing chunk decides the code."""
    a, _ = _u8(data)
    if a.size == 0 or len(chunks) == 0:
        return ParseResult(ERR_INPUT, np.zeros(0, np.int32), 0, 0)
    rows_s, rows_p, sizes, prefix = [], [], [], []
    last_chunk, last_struct, total = 0, 0, 0
    for start, size in chunks:
        if size == 0 or size > a.size:
            return ParseResult(ERR_INPUT, np.zeros(0, np.int32), 0, 0)
        r = parse_unit(a[start:start + size], flags | FLAG_JSONL, last_chunk, last_struct)
        if r.code != OK:
            return ParseResult(r.code, np.zeros(0, np.int32), 0, len(chunks))
        rows_s.append(r.structural)
        rows_p.append(r.pair)
        N = r.structural.size
        total += N
        sizes.append(N)
        prefix.append(total)
        last_struct += N
        last_chunk += size
    buf, N = assemble(rows_s, rows_p)
    return ParseResult(OK, buf, N, len(chunks), sizes, prefix)


def simdjson_op_indices(data):
    """simdjson stage-1 structural indices filtered to { } [ ] : , as 1-based
    positions, or None when the reference build is absent.  Returns
    (error_code, positions)."""
    L = ref_lib()
    if L is None:
        return None
    a, p = _u8(data)
    cap = a.size + 16
    out = np.zeros(cap, dtype=np.uint32)
    n = ctypes.c_size_t()
    e = L.sjr_stage1(p, a.size, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), cap,
                     ctypes.byref(n))
    idx = out[: n.value]
    idx = idx[idx < a.size]
    ch = a[idx]
    keep = np.isin(ch, np.frombuffer(b"{}[]:,", dtype=np.uint8))
    return e, (idx[keep] + 1).astype(np.uint32)
