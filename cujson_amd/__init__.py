"""cujson_amd - MI355X-native three-phase JSON parse (UTF-8 validation,
tokenization, nesting-structure recognition) behind cuJSON's interface.

Python view of the C ABI in ``include/cujson_amd.h`` (``libcujson_amd.so``,
built in-tree by ``cujson_amd/build.py``).  The names mirror the reference's
C++ API so code written against cuJSON reads the same:

=============================================  ==========================================
reference (cujson/)                             here
=============================================  ==========================================
``loadJSON`` load_file.h:9                      :func:`loadJSON`
``loadJSONLines_chunkCount`` load_file.h:10     :func:`loadJSONLines_chunkCount`
``loadJSONLines_chunkSizeBytes`` :11            :func:`loadJSONLines_chunkSizeBytes`
``loadJSONLines_chunkSizeMegaBytes`` :12        :func:`loadJSONLines_chunkSizeMegaBytes`
``parse_standard_json`` parse_standard_json.h:9 :func:`parse_standard_json`
``parse_json_lines`` parse_json_lines.h:9       :func:`parse_json_lines`
``cuJSONResult`` cujson_types.h:22-33           :class:`cuJSONResult`
=============================================  ==========================================

Error behaviour: where the reference returns an empty ``cuJSONResult{}``
(bad arguments) these functions return a result with ``structural is None``;
where the reference calls ``exit(0)`` (invalid UTF-8, invalid structure) they
return a result with ``structural is None`` and ``error`` set, instead of
ending the process.  There is no CPU fallback: every parse runs the HIP
kernels, and importing the package fails if the library is not built.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcujson_amd.so")

CJX_OK = 0
CJX_ERR_INPUT = 1
CJX_ERR_UTF8 = 2
CJX_ERR_STRUCTURE = 3
CJX_ERR_ALLOC = 4
CJX_ERR_DEVICE = 5
CJX_ERR_CAPACITY = 6
CJX_ERR_TOO_LARGE = 7

CJX_FLAG_UTF8_STRICT = 0x1
CJX_FLAG_DEPTH_U16 = 0x2
CJX_FLAG_JSONL = 0x4
CJX_FLAG_GRAPH = 0x8
CJX_FLAG_RECORDS = 0x10

# Every symbol include/cujson_amd.h declares.
ABI_SYMBOLS = (
    "cjx_context_create", "cjx_context_destroy", "cjx_load", "cjx_free_input",
    "cjx_chunk_lines", "cjx_parse_standard", "cjx_parse_lines", "cjx_free_result",
    "cjx_strerror", "cjx_parse_device", "cjx_device_wait", "cjx_context_stream",
    "cjx_version", "cjx_context_set_timing", "cjx_context_timings",
    "cjx_context_set_tokenizer_shape", "cjx_context_tokenizer_shape",
    "cjx_parse_batch", "cjx_free_batch_result", "cjx_parse_batch_device", "cjx_batch_wait",
    "cjx_parse_lines_device", "cjx_server_create", "cjx_server_submit", "cjx_server_wait",
    "cjx_server_release", "cjx_server_reserve", "cjx_server_get_stats", "cjx_server_destroy",
    "cjx_scan_lines_device", "cjx_chunk_lines_device", "cjx_load_device",
    "cjx_values_device", "cjx_values", "cjx_noalloc_max_bytes", "cjx_parse_device_noalloc",
)


class _Result(ctypes.Structure):
    _fields_ = [
        ("buf", ctypes.POINTER(ctypes.c_int32)),
        ("structural", ctypes.POINTER(ctypes.c_int32)),
        ("pair_pos", ctypes.POINTER(ctypes.c_int32)),
        ("n_structural", ctypes.c_int64),
        ("n_brackets", ctypes.c_int64),
        ("chunk_count", ctypes.c_int32),
        ("result_sizes", ctypes.POINTER(ctypes.c_int32)),
        ("result_sizes_prefix", ctypes.POINTER(ctypes.c_int32)),
        ("total_result_size", ctypes.c_int32),
        ("file_size", ctypes.c_int32),
        ("error", ctypes.c_int32),
        ("error_chunk", ctypes.c_int32),
        ("device_ms", ctypes.c_double),
        ("n_records", ctypes.c_int64),
        ("records", ctypes.c_void_p),
    ]


# cjx_record: one JSON Lines record (CJX_FLAG_RECORDS)
# Typed value table (cjx_value): types and flags.
CJX_VALUE_NONE, CJX_VALUE_KEY, CJX_VALUE_STRING, CJX_VALUE_INT, CJX_VALUE_FLOAT = 0, 1, 2, 3, 4
CJX_VALUE_TRUE, CJX_VALUE_FALSE, CJX_VALUE_NULL, CJX_VALUE_INVALID = 5, 6, 7, 8
CJX_VALUE_ESCAPED, CJX_VALUE_BIG_INT, CJX_VALUE_UNPAIRED, CJX_VALUE_APPROX = 0x1, 0x2, 0x4, 0x8
VALUE_DTYPE = np.dtype([("v", np.int64), ("row", np.uint32), ("text_off", np.uint32),
                        ("text_len", np.uint32), ("type", np.uint8), ("flags", np.uint8),
                        ("reserved", np.uint16)])

RECORD_DTYPE = np.dtype([("byte_begin", np.int64), ("byte_end", np.int64),
                         ("struct_first", np.int32), ("struct_last", np.int32),
                         ("chunk", np.int32), ("error", np.int32)])


class _Input(ctypes.Structure):
    _fields_ = [("data", ctypes.POINTER(ctypes.c_uint8)), ("size", ctypes.c_size_t)]


class _Timings(ctypes.Structure):
    _fields_ = [("tokenize_ms", ctypes.c_float), ("phase3_ms", ctypes.c_float),
                ("total_ms", ctypes.c_float)]


class _DeviceStatus(ctypes.Structure):
    _fields_ = [("n_structural", ctypes.c_int64), ("n_brackets", ctypes.c_int64),
                ("error", ctypes.c_int32), ("capacity_exceeded", ctypes.c_int32),
                ("n_records", ctypes.c_int64)]


class _DocResult(ctypes.Structure):
    _fields_ = [("byte_base", ctypes.c_int64), ("struct_begin", ctypes.c_int64),
                ("struct_end", ctypes.c_int64), ("error", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class _BatchResult(ctypes.Structure):
    _fields_ = [("structural", ctypes.POINTER(ctypes.c_int32)),
                ("pair_pos", ctypes.POINTER(ctypes.c_int32)),
                ("n_structural", ctypes.c_int64), ("n_docs", ctypes.c_int64),
                ("docs", ctypes.POINTER(_DocResult)), ("error", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("device_ms", ctypes.c_double)]


class _Completion(ctypes.Structure):
    _fields_ = [("ticket", ctypes.c_uint64), ("buf", ctypes.POINTER(ctypes.c_int32)),
                ("owner", ctypes.c_void_p), ("n_structural", ctypes.c_int64),
                ("error", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("latency_ms", ctypes.c_double), ("queue_ms", ctypes.c_double)]


class _ServerStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in ("batches", "documents", "bytes",
                                                "max_inflight_bytes", "relaunches", "pending",
                                                "region_bytes")]


_lib = None


def lib():
    """The loaded C ABI library.  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m cujson_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    u8p = ctypes.POINTER(ctypes.c_uint8)
    szp = ctypes.POINTER(ctypes.c_size_t)
    L.cjx_context_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    L.cjx_context_create.restype = ctypes.c_int
    L.cjx_context_destroy.argtypes = [vp]
    L.cjx_context_destroy.restype = None
    L.cjx_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(_Input)]
    L.cjx_load.restype = ctypes.c_int
    L.cjx_free_input.argtypes = [ctypes.POINTER(_Input)]
    L.cjx_free_input.restype = None
    L.cjx_chunk_lines.argtypes = [u8p, sz, ctypes.c_int, sz, szp, szp, sz, szp]
    L.cjx_chunk_lines.restype = ctypes.c_int
    L.cjx_parse_standard.argtypes = [vp, u8p, sz, ctypes.c_uint, ctypes.POINTER(_Result)]
    L.cjx_parse_standard.restype = ctypes.c_int
    L.cjx_parse_lines.argtypes = [vp, u8p, sz, szp, szp, sz, ctypes.c_uint,
                                  ctypes.POINTER(_Result)]
    L.cjx_parse_lines.restype = ctypes.c_int
    L.cjx_free_result.argtypes = [ctypes.POINTER(_Result)]
    L.cjx_free_result.restype = None
    L.cjx_strerror.argtypes = [ctypes.c_int]
    L.cjx_strerror.restype = ctypes.c_char_p
    L.cjx_parse_device.argtypes = [vp, vp, sz, ctypes.c_uint, vp, vp, sz]
    L.cjx_parse_device.restype = ctypes.c_int
    L.cjx_device_wait.argtypes = [vp, ctypes.POINTER(_DeviceStatus)]
    L.cjx_device_wait.restype = ctypes.c_int
    L.cjx_context_stream.argtypes = [vp]
    L.cjx_context_stream.restype = vp
    L.cjx_context_set_timing.argtypes = [vp, ctypes.c_int]
    L.cjx_context_set_timing.restype = ctypes.c_int
    L.cjx_context_timings.argtypes = [vp, ctypes.POINTER(_Timings), ctypes.c_int]
    L.cjx_context_timings.restype = ctypes.c_int
    L.cjx_context_set_tokenizer_shape.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.cjx_context_set_tokenizer_shape.restype = ctypes.c_int
    L.cjx_context_tokenizer_shape.argtypes = [vp, ctypes.POINTER(ctypes.c_int),
                                              ctypes.POINTER(ctypes.c_int)]
    L.cjx_context_tokenizer_shape.restype = ctypes.c_int
    L.cjx_version.argtypes = []
    L.cjx_version.restype = ctypes.c_char_p
    L.cjx_parse_batch.argtypes = [vp, u8p, szp, szp, sz, ctypes.c_uint,
                                  ctypes.POINTER(_BatchResult)]
    L.cjx_parse_batch.restype = ctypes.c_int
    L.cjx_free_batch_result.argtypes = [ctypes.POINTER(_BatchResult)]
    L.cjx_free_batch_result.restype = None
    L.cjx_parse_batch_device.argtypes = [vp, vp, szp, szp, sz, ctypes.c_uint, vp, vp, sz]
    L.cjx_parse_batch_device.restype = ctypes.c_int
    L.cjx_batch_wait.argtypes = [vp, ctypes.POINTER(_DocResult), sz,
                                 ctypes.POINTER(_DeviceStatus)]
    L.cjx_batch_wait.restype = ctypes.c_int
    L.cjx_parse_lines_device.argtypes = [vp, vp, szp, szp, sz, ctypes.c_uint, vp, vp, sz, vp, sz]
    L.cjx_parse_lines_device.restype = ctypes.c_int
    L.cjx_scan_lines_device.argtypes = [vp, vp, sz, vp, sz, szp]
    L.cjx_scan_lines_device.restype = ctypes.c_int
    L.cjx_chunk_lines_device.argtypes = [vp, vp, sz, ctypes.c_int, sz, szp, szp, sz, szp]
    L.cjx_chunk_lines_device.restype = ctypes.c_int
    L.cjx_load_device.argtypes = [vp, ctypes.c_char_p, vp, sz, szp]
    L.cjx_load_device.restype = ctypes.c_int
    L.cjx_values_device.argtypes = [vp, vp, sz, vp, sz, vp, vp, szp]
    L.cjx_values_device.restype = ctypes.c_int
    L.cjx_values.argtypes = [vp, u8p, sz, ctypes.POINTER(_Result), vp, vp]
    L.cjx_values.restype = ctypes.c_int
    L.cjx_server_create.argtypes = [ctypes.c_int, sz, ctypes.c_int, sz, ctypes.POINTER(vp)]
    L.cjx_server_create.restype = ctypes.c_int
    L.cjx_server_reserve.argtypes = [vp, sz]
    L.cjx_server_reserve.restype = ctypes.c_int
    L.cjx_server_submit.argtypes = [vp, vp, sz, ctypes.c_uint, ctypes.POINTER(ctypes.c_uint64)]
    L.cjx_server_submit.restype = ctypes.c_int
    L.cjx_server_wait.argtypes = [vp, ctypes.POINTER(_Completion), sz, ctypes.POINTER(sz),
                                  ctypes.c_int]
    L.cjx_server_wait.restype = ctypes.c_int
    L.cjx_server_release.argtypes = [vp, ctypes.POINTER(_Completion)]
    L.cjx_server_release.restype = None
    L.cjx_server_get_stats.argtypes = [vp, ctypes.POINTER(_ServerStats)]
    L.cjx_server_get_stats.restype = ctypes.c_int
    L.cjx_server_destroy.argtypes = [vp]
    L.cjx_server_destroy.restype = None
    L.cjx_noalloc_max_bytes.argtypes = []
    L.cjx_noalloc_max_bytes.restype = sz
    L.cjx_parse_device_noalloc.argtypes = [vp, sz, ctypes.c_uint, vp, vp, sz, vp, vp]
    L.cjx_parse_device_noalloc.restype = ctypes.c_int
    _lib = L
    return L


def strerror(code: int) -> str:
    return lib().cjx_strerror(code).decode()


# cjx_device_status as a numpy record (device status of the no-allocation entry)
DEVICE_STATUS_DTYPE = np.dtype([("n_structural", "<i8"), ("n_brackets", "<i8"), ("error", "<i4"),
                                ("capacity_exceeded", "<i4"), ("n_records", "<i8")])


def noalloc_max_bytes() -> int:
    """Largest input cjx_parse_device_noalloc accepts."""
    return int(lib().cjx_noalloc_max_bytes())


def parse_device_noalloc(d_input: int, n: int, d_structural: int, d_pair: int, cap: int,
                         d_status: int, flags: int = 0, stream: int = 0) -> int:
    """Enqueue cjx_parse_device_noalloc on `stream` (0: the null stream): the
    device-resident parse with no allocation; the status (DEVICE_STATUS_DTYPE)
    lands in device memory at d_status once the stream reaches it."""
    return lib().cjx_parse_device_noalloc(ctypes.c_void_p(d_input), n, flags,
                                          ctypes.c_void_p(d_structural), ctypes.c_void_p(d_pair),
                                          cap, ctypes.c_void_p(d_status),
                                          ctypes.c_void_p(stream or None))


class Context:
    """One HIP stream plus the device buffers of the parse (cjx_context)."""

    def __init__(self, device: int = 0):
        self._h = ctypes.c_void_p()
        rc = lib().cjx_context_create(device, ctypes.byref(self._h))
        if rc != CJX_OK:
            raise RuntimeError(f"cjx_context_create failed: {strerror(rc)}")

    @property
    def handle(self):
        return self._h

    @property
    def stream(self) -> int:
        return lib().cjx_context_stream(self._h) or 0

    def close(self):
        if self._h:
            lib().cjx_context_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- device-resident form ------------------------------------------------
    def parse_device(self, d_input: int, n: int, d_structural: int, d_pair: int, cap: int,
                     flags: int = 0) -> int:
        """Enqueue a parse of a device-resident document (pointers are ints)."""
        return lib().cjx_parse_device(self._h, ctypes.c_void_p(d_input), n, flags,
                                      ctypes.c_void_p(d_structural), ctypes.c_void_p(d_pair), cap)

    def set_timing(self, on: bool = True):
        lib().cjx_context_set_timing(self._h, 1 if on else 0)

    def set_tokenizer_shape(self, tile_waves: int = 4, small_waves: int = 4):
        """Tokenizer launch shape of later parses: wave64s per 64 KiB tile
        (2, 4, 8, 16) and one-wave documents per workgroup (1, 2, 4, 8)."""
        rc = lib().cjx_context_set_tokenizer_shape(self._h, int(tile_waves), int(small_waves))
        if rc:
            raise ValueError(f"invalid tokenizer shape ({tile_waves}, {small_waves})")

    def tokenizer_shape(self):
        """(tile_waves, small_waves) of later parses."""
        t, sm = ctypes.c_int(), ctypes.c_int()
        if lib().cjx_context_tokenizer_shape(self._h, ctypes.byref(t), ctypes.byref(sm)):
            raise RuntimeError("no context")
        return t.value, sm.value

    def timings(self, max_entries: int = 1):
        """[(tokenize_ms, phase3_ms, total_ms)] of the most recent parses, oldest
        first (timing enabled); the default returns the last parse only."""
        arr = (_Timings * max_entries)()
        k = lib().cjx_context_timings(self._h, arr, max_entries)
        if k < 0:
            raise RuntimeError("timings unavailable (timing not enabled?)")
        out = [(arr[i].tokenize_ms, arr[i].phase3_ms, arr[i].total_ms) for i in range(k)]
        return out if max_entries > 1 else out[0]

    def wait(self):
        st = _DeviceStatus()
        rc = lib().cjx_device_wait(self._h, ctypes.byref(st))
        return rc, st.n_structural, st.n_brackets, bool(st.capacity_exceeded)

    def parse_batch_device(self, d_input: int, doc_off, doc_size, d_structural: int,
                           d_pair: int, cap: int, flags: int = 0) -> int:
        """Enqueue a batch of device-resident documents (documents at unaligned
        offsets are copied into an aligned layout first)."""
        offs = np.ascontiguousarray(doc_off, dtype=np.uint64)
        sizes = np.ascontiguousarray(doc_size, dtype=np.uint64)
        k = offs.size
        self._batch_n = k
        szp = ctypes.POINTER(ctypes.c_size_t)
        return lib().cjx_parse_batch_device(self._h, ctypes.c_void_p(d_input),
                                            offs.ctypes.data_as(szp), sizes.ctypes.data_as(szp),
                                            k, flags, ctypes.c_void_p(d_structural),
                                            ctypes.c_void_p(d_pair), cap)

    def parse_lines_device(self, d_input: int, chunk_off, chunk_size, d_structural: int,
                           d_pair: int, cap: int, d_records: int = 0, max_records: int = 0,
                           flags: int = 0) -> int:
        """Enqueue device-resident JSON Lines chunks (chunks at unaligned offsets
        are copied into an aligned layout first); with d_records the per-record
        table goes there (CJX_FLAG_RECORDS is set)."""
        offs = np.ascontiguousarray(chunk_off, dtype=np.uint64)
        sizes = np.ascontiguousarray(chunk_size, dtype=np.uint64)
        self._batch_n = offs.size
        szp = ctypes.POINTER(ctypes.c_size_t)
        if d_records:
            flags |= CJX_FLAG_RECORDS
        return lib().cjx_parse_lines_device(self._h, ctypes.c_void_p(d_input),
                                            offs.ctypes.data_as(szp), sizes.ctypes.data_as(szp),
                                            offs.size, flags, ctypes.c_void_p(d_structural),
                                            ctypes.c_void_p(d_pair), cap,
                                            ctypes.c_void_p(d_records), max_records)

    def scan_lines_device(self, d_input: int, n: int, d_newlines: int, cap: int):
        """(rc, count): offsets of every '\\n' of device text into d_newlines
        (uint64[cap])."""
        k = ctypes.c_size_t()
        rc = lib().cjx_scan_lines_device(self._h, ctypes.c_void_p(d_input), n,
                                         ctypes.c_void_p(d_newlines), cap, ctypes.byref(k))
        return rc, k.value

    def chunk_lines_device(self, d_input: int, n: int, mode: int, param: int):
        """The chunks of chunk_lines(mode, param) for device text: (offsets, sizes)."""
        k = ctypes.c_size_t()
        rc = lib().cjx_chunk_lines_device(self._h, ctypes.c_void_p(d_input), n, mode, param, None,
                                          None, 0, ctypes.byref(k))
        if rc not in (CJX_OK, CJX_ERR_CAPACITY):
            raise ValueError(strerror(rc))
        cnt = k.value
        offs = np.zeros(max(cnt, 1), np.uint64)
        sizes = np.zeros(max(cnt, 1), np.uint64)
        szp = ctypes.POINTER(ctypes.c_size_t)
        rc = lib().cjx_chunk_lines_device(self._h, ctypes.c_void_p(d_input), n, mode, param,
                                          offs.ctypes.data_as(szp), sizes.ctypes.data_as(szp), cnt,
                                          ctypes.byref(k))
        if rc != CJX_OK:
            raise ValueError(strerror(rc))
        return offs[:cnt].astype(np.int64), sizes[:cnt].astype(np.int64)

    def load_device(self, path: str, d_buf: int, cap: int):
        """(rc, size): the file at path read into device memory d_buf[0..cap)."""
        k = ctypes.c_size_t()
        rc = lib().cjx_load_device(self._h, os.fsencode(path), ctypes.c_void_p(d_buf), cap,
                                   ctypes.byref(k))
        return rc, k.value

    def values_device(self, d_input: int, n: int, d_structural: int, n_structural: int,
                      d_values: int, d_strings: int = 0):
        """(rc, n_exact): the typed value table (VALUE_DTYPE, n_structural + 1
        entries) of device text and its structural row into d_values; unescaped
        strings into d_strings (n bytes) when given."""
        k = ctypes.c_size_t()
        rc = lib().cjx_values_device(self._h, ctypes.c_void_p(d_input), n,
                                     ctypes.c_void_p(d_structural), n_structural,
                                     ctypes.c_void_p(d_values), ctypes.c_void_p(d_strings),
                                     ctypes.byref(k))
        return rc, k.value

    def values(self, data, result: "cuJSONResult", strings: bool = True):
        """(table, strings) for a parse result of `data`: table is a VALUE_DTYPE
        array of N + 1 entries; strings (bytes of len(data)) holds the unescaped
        keys and strings at their text_off."""
        a, p = _as_u8(data)
        N = int(result._r.n_structural)
        table = np.zeros(N + 1, VALUE_DTYPE)
        sbuf = np.zeros(a.size if strings else 1, np.uint8)
        rc = lib().cjx_values(self._h, p, a.size, ctypes.byref(result._r),
                              table.ctypes.data_as(ctypes.c_void_p),
                              sbuf.ctypes.data_as(ctypes.c_void_p) if strings else None)
        if rc != CJX_OK:
            raise RuntimeError(f"cjx_values: {strerror(rc)}")
        return table, (sbuf.tobytes() if strings else None)

    def batch_wait(self):
        """(rc, docs, N, M, overflow): docs is an int64 array [k, 4] of
        (byte_base, struct_begin, struct_end, error)."""
        k = getattr(self, "_batch_n", 0)
        arr = np.zeros((max(k, 1), 4), dtype=np.int64)  # the cjx_doc_result layout
        st = _DeviceStatus()
        rc = lib().cjx_batch_wait(self._h, arr.ctypes.data_as(ctypes.POINTER(_DocResult)), k,
                                  ctypes.byref(st))
        docs = arr[:k].copy()
        docs[:, 3] = docs[:, 3] & 0xFFFFFFFF  # error (int32) then the reserved field
        self.last_records = int(st.n_records)
        return rc, docs, st.n_structural, st.n_brackets, bool(st.capacity_exceeded)


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


@dataclass
class cuJSONInput:
    """cuJSONInput (cujson/cujson_types.h:9-12); ``data`` is a numpy uint8 view."""
    data: Optional[np.ndarray]
    size: int
    _pinned: object = None


@dataclass
class cuJSONLinesInput:
    """cuJSONLinesInput (cujson/cujson_types.h:14-20); chunks as (offset, size)."""
    data: Optional[np.ndarray]
    chunkCount: int
    size: int
    chunks: List[int] = field(default_factory=list)       # chunk start offsets
    chunksSize: List[int] = field(default_factory=list)
    _pinned: object = None


class cuJSONResult:
    """cuJSONResult (cujson/cujson_types.h:22-33) over the pinned result buffer.

    ``structural`` and ``pair_pos`` are int32 numpy views into one buffer of
    2(N+2) cells with ``pair_pos = buf + N + 1`` exactly as in the reference."""

    def __init__(self, r: _Result):
        self._r = r
        self.error = int(r.error)
        self.error_chunk = int(r.error_chunk)
        self.device_ms = float(r.device_ms)
        self.chunkCount = int(r.chunk_count)
        self.bufferSize = 0
        self.depth = 0
        self.inputJSON = None
        N = int(r.n_structural)
        self.n_brackets = int(r.n_brackets)
        if r.buf:
            cells = 2 * (N + 2)
            self.buf = np.ctypeslib.as_array(r.buf, shape=(cells,))
            self.structural = self.buf[: N + 2]
            self.pair_pos = self.buf[N + 1:]
            self.totalResultSize = int(r.total_result_size)
            self.fileSize = int(r.file_size)
            k = self.chunkCount
            self.resultSizes = list(np.ctypeslib.as_array(r.result_sizes, shape=(k,))) if k else []
            self.resultSizesPrefix = (list(np.ctypeslib.as_array(r.result_sizes_prefix, shape=(k,)))
                                      if k else [])
        else:
            self.buf = None
            self.structural = None
            self.pair_pos = None
            self.totalResultSize = 0
            self.fileSize = 0
            self.resultSizes = []
            self.resultSizesPrefix = []
        self.N = N
        nrec = int(r.n_records)
        if r.records and nrec:
            raw = (ctypes.c_char * (nrec * RECORD_DTYPE.itemsize)).from_address(r.records)
            self.records = np.frombuffer(raw, dtype=RECORD_DTYPE).copy()
        else:
            self.records = np.zeros(0, dtype=RECORD_DTYPE)

    def free(self):
        if self._r is not None:
            lib().cjx_free_result(ctypes.byref(self._r))
            self._r = None
            self.buf = self.structural = self.pair_pos = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _as_u8(data):
    if isinstance(data, cuJSONInput) or isinstance(data, cuJSONLinesInput):
        data = data.data
    if data is None:
        return None, None
    a = np.frombuffer(data, dtype=np.uint8) if not isinstance(data, np.ndarray) else data
    a = np.ascontiguousarray(a, dtype=np.uint8)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def loadJSON(filePath: str) -> cuJSONInput:
    """load_file.cu:7-36: file -> pinned host buffer ({None, 0} on failure)."""
    inp = _Input()
    rc = lib().cjx_load(filePath.encode(), ctypes.byref(inp))
    if rc != CJX_OK or not inp.data:
        return cuJSONInput(None, 0)
    arr = np.ctypeslib.as_array(inp.data, shape=(max(inp.size, 1),))[: inp.size]
    return cuJSONInput(arr, int(inp.size), _PinnedInput(inp))


class _PinnedInput:
    def __init__(self, inp: _Input):
        self.inp = inp

    def __del__(self):
        try:
            lib().cjx_free_input(ctypes.byref(self.inp))
        except Exception:
            pass


def chunk_lines(data, mode: int, param: int):
    """Line-aligned chunking (load_file.cu:71-103 count mode 0, :137-176 bytes mode 1)."""
    a, p = _as_u8(data)
    n = ctypes.c_size_t()
    lib().cjx_chunk_lines(p, a.size, mode, param, None, None, 0, ctypes.byref(n))
    k = n.value
    offs = (ctypes.c_size_t * max(k, 1))()
    sizes = (ctypes.c_size_t * max(k, 1))()
    rc = lib().cjx_chunk_lines(p, a.size, mode, param, offs, sizes, k, ctypes.byref(n))
    if rc != CJX_OK:
        raise ValueError(strerror(rc))
    return [int(offs[i]) for i in range(k)], [int(sizes[i]) for i in range(k)]


def _lines_input(inp: cuJSONInput, mode: int, param: int) -> cuJSONLinesInput:
    if inp.data is None:
        return cuJSONLinesInput(None, 0, 0)
    offs, sizes = chunk_lines(inp.data, mode, param)
    return cuJSONLinesInput(inp.data, len(offs), inp.size, offs, sizes, inp._pinned)


def loadJSONLines_chunkCount(filePath: str, chunkCount: int = 1) -> cuJSONLinesInput:
    return _lines_input(loadJSON(filePath), 0, chunkCount)


def loadJSONLines_chunkSizeBytes(filePath: str, chunkSizeBytes: int) -> cuJSONLinesInput:
    return _lines_input(loadJSON(filePath), 1, chunkSizeBytes)


def loadJSONLines_chunkSizeMegaBytes(filePath: str, chunkSizeMegaBytes: int) -> cuJSONLinesInput:
    return _lines_input(loadJSON(filePath), 1, chunkSizeMegaBytes * 1024 * 1024)


def parse_standard_json(input, flags: int = 0, ctx: Optional[Context] = None) -> cuJSONResult:
    """parse_standard_json (parse_standard_json.cu:1584-1687) on the GPU."""
    a, p = _as_u8(input)
    ctx = ctx or default_context()
    r = _Result()
    n = 0 if a is None else a.size
    lib().cjx_parse_standard(ctx.handle, p, n, flags, ctypes.byref(r))
    return cuJSONResult(r)


def parse_json_lines(input: cuJSONLinesInput, flags: int = 0,
                     ctx: Optional[Context] = None) -> cuJSONResult:
    """parse_json_lines (parse_json_lines.cu:1092-1264) on the GPU."""
    a, p = _as_u8(input.data)
    ctx = ctx or default_context()
    k = len(input.chunks)
    offs = (ctypes.c_size_t * max(k, 1))(*input.chunks)
    sizes = (ctypes.c_size_t * max(k, 1))(*input.chunksSize)
    r = _Result()
    n = 0 if a is None else a.size
    nch = min(input.chunkCount, k)
    if input.chunkCount > k:  # metadata smaller than chunkCount, parse_json_lines.cu:1118-1121
        nch = 0
    lib().cjx_parse_lines(ctx.handle, p, n, offs, sizes, nch, flags, ctypes.byref(r))
    return cuJSONResult(r)


class BatchResult:
    """Result of parse_batch: packed rows of all documents plus a per-document
    table.  ``doc(i)`` rebuilds the cuJSONResult buffer parse_standard_json
    would return for document i (2(N_i+2) cells, sentinels, -1 cells)."""

    def __init__(self, r: _BatchResult, rc: int):
        self._r = r
        self.rc = rc
        self.error = int(r.error)
        self.N = int(r.n_structural)
        self.n_docs = int(r.n_docs)
        self.device_ms = float(r.device_ms)
        if rc == CJX_OK and r.structural:
            self.structural = np.ctypeslib.as_array(r.structural, shape=(max(self.N, 1),))[:self.N]
            self.pair_pos = np.ctypeslib.as_array(r.pair_pos, shape=(max(self.N, 1),))[:self.N]
            raw = np.ctypeslib.as_array(ctypes.cast(r.docs, ctypes.POINTER(ctypes.c_int64)),
                                        shape=(self.n_docs * 4,)).reshape(-1, 4).copy()
            raw[:, 3] = raw[:, 3] & 0xFFFFFFFF  # error (int32) then the reserved field
            self.docs = raw
        else:
            self.structural = self.pair_pos = None
            self.docs = np.zeros((0, 4), np.int64)

    def doc(self, i: int):
        """(error, buf) of document i in the reference's host layout."""
        _, b, e, err = self.docs[i]
        n = int(e - b)
        buf = np.empty(2 * (n + 2), dtype=np.int32)
        buf[0] = 0
        buf[1:n + 1] = self.structural[b:e]
        buf[n + 1] = n + 1
        buf[n + 2:2 * n + 2] = self.pair_pos[b:e]
        buf[2 * n + 2:] = -1
        return int(err), buf

    def free(self):
        if self._r is not None:
            lib().cjx_free_batch_result(ctypes.byref(self._r))
            self._r = None
            self.structural = self.pair_pos = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def parse_batch(data, doc_off, doc_size, flags: int = 0,
                ctx: Optional[Context] = None) -> BatchResult:
    """A batch of independent standard-JSON documents in one launch sequence
    (documents are data[doc_off[i] : doc_off[i] + doc_size[i]])."""
    a, p = _as_u8(data)
    ctx = ctx or default_context()
    offs = np.ascontiguousarray(doc_off, dtype=np.uint64)
    sizes = np.ascontiguousarray(doc_size, dtype=np.uint64)
    szp = ctypes.POINTER(ctypes.c_size_t)
    r = _BatchResult()
    rc = lib().cjx_parse_batch(ctx.handle, p, offs.ctypes.data_as(szp), sizes.ctypes.data_as(szp),
                               offs.size, flags, ctypes.byref(r))
    return BatchResult(r, rc)


class Server:
    """Request server with continuous batching (cjx_server_* in the C ABI).

    submit() queues one standard-JSON document and returns its ticket;
    completions() returns finished requests as (ticket, error, buf,
    latency_ms, queue_ms) with buf the reference host layout (a numpy copy).
    The submitted bytes are kept referenced until their completion."""

    def __init__(self, device: int = 0, budget: int = 8 << 30, slots: int = 4,
                 max_batch: int = 4096):
        h = ctypes.c_void_p()
        rc = lib().cjx_server_create(device, budget, slots, max_batch, ctypes.byref(h))
        if rc:
            raise RuntimeError(f"cjx_server_create: {strerror(rc)}")
        self.handle = h
        self._live = {}

    def reserve(self, max_request_bytes: int) -> None:
        """Pre-sizes every slot (only before the first submit)."""
        rc = lib().cjx_server_reserve(self.handle, max_request_bytes)
        if rc:
            raise RuntimeError(f"cjx_server_reserve: {strerror(rc)}")

    def submit(self, data, flags: int = 0) -> int:
        a = np.ascontiguousarray(np.frombuffer(data, dtype=np.uint8)
                                 if isinstance(data, (bytes, bytearray)) else data,
                                 dtype=np.uint8)
        t = ctypes.c_uint64()
        rc = lib().cjx_server_submit(self.handle, a.ctypes.data, a.size, flags, ctypes.byref(t))
        if rc:
            raise RuntimeError(f"cjx_server_submit: {strerror(rc)}")
        self._live[t.value] = a
        return t.value

    def completions(self, max_n: int = 1024, timeout_ms: int = 100):
        arr = (_Completion * max_n)()
        n = ctypes.c_size_t()
        L = lib()
        L.cjx_server_wait(self.handle, arr, max_n, ctypes.byref(n), timeout_ms)
        out = []
        for i in range(n.value):
            c = arr[i]
            buf = None
            if c.buf:
                cells = 2 * (int(c.n_structural) + 2)
                buf = np.ctypeslib.as_array(c.buf, shape=(cells,)).copy()
            out.append((int(c.ticket), int(c.error), buf, float(c.latency_ms),
                        float(c.queue_ms)))
            L.cjx_server_release(self.handle, ctypes.byref(c))
            self._live.pop(int(c.ticket), None)
        return out

    def stats(self) -> dict:
        st = _ServerStats()
        lib().cjx_server_get_stats(self.handle, ctypes.byref(st))
        return {k: int(getattr(st, k)) for k, _ in _ServerStats._fields_}

    def close(self):
        if self.handle:
            lib().cjx_server_destroy(self.handle)
            self.handle = None
            self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
