"""Autotuner: JSON Lines chunk size and the tokenizer's launch shape.

The launch shape of a JSON Lines parse is fixed by the chunk size: every
chunk is one parse unit.  Small chunks mean many units (and more tail waste
per unit); large chunks mean fewer, larger units whose phase-3 work runs per
unit.  `tune_lines_chunk` times the candidate chunk sizes on a sample of the
workload, device-resident, and returns the fastest.

The tokenizer's per-thread work and workgroup size are picked at run time
from the library's built instantiations (`Context.set_tokenizer_shape`):
a 64 KiB tile is 2, 4, 8 or 16 wave64s (1024, 512, 256 or 128 input bytes
per lane), and documents of at most 16 KiB take one wave each, 1, 2, 4 or 8
per workgroup.  `tune_tokenizer_shape` times the candidates on the caller's
own parse call and leaves the fastest set on the context.

Results are cached in a JSON file (``$CJX_TUNE_CACHE`` or
``~/.cache/cujson_amd/tune.json``) keyed by the workload signature the
caller gives (name it for the device and the input shape), so a later run
reuses them.
"""
from __future__ import annotations

import json
import os
from typing import Callable, Dict, Sequence, Tuple

import numpy as np

CANDIDATES = (256 << 10, 1 << 20, 4 << 20, 16 << 20, 64 << 20)
TILE_WAVES = (2, 4, 8, 16)   # wave64s per 64 KiB tokenizer tile
SMALL_WAVES = (1, 2, 4, 8)   # one-wave documents per workgroup


def _cache_path() -> str:
    return os.environ.get("CJX_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "cujson_amd",
                                       "tune.json"))


def _load_cache() -> Dict[str, dict]:
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(c: Dict[str, dict]) -> None:
    p = _cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        with open(p, "w") as f:
            json.dump(c, f, indent=1)
    except OSError:
        pass


def aligned_layout(data: np.ndarray, offs: Sequence[int], sizes: Sequence[int]):
    """Places chunks at 16-byte aligned offsets of one buffer (the device
    entry points require it).  Returns (buffer, new_offsets)."""
    new = np.empty(len(offs), dtype=np.int64)
    pos = 0
    for i, s in enumerate(sizes):
        new[i] = pos
        pos += -(-int(s) // 16) * 16
    buf = np.zeros(pos + 64, dtype=np.uint8)
    for i, (o, s) in enumerate(zip(offs, sizes)):
        buf[new[i]:new[i] + s] = data[o:o + s]
    return buf, new


def time_lines(ctx, data: np.ndarray, chunk_bytes: int, reps: int = 3) -> float:
    """Median device time (ms) of one device-resident JSON Lines parse of
    `data` cut into line-aligned chunks of at most chunk_bytes."""
    import cujson_amd as cj
    from cujson_amd.hip import DeviceBuffer
    offs, sizes = cj.chunk_lines(data, 1, chunk_bytes)
    if any(s == 0 for s in sizes):
        raise ValueError("a line is longer than the chunk size")
    buf, aoffs = aligned_layout(data, offs, sizes)
    dev = DeviceBuffer(buf.size)
    dev.upload(buf)
    cap = data.size // 3 + 256 * len(sizes) + 64
    st, pr = DeviceBuffer(cap * 4), DeviceBuffer(cap * 4)
    rc = ctx.parse_lines_device(dev.ptr, aoffs, sizes, st.ptr, pr.ptr, cap)
    rc2 = ctx.batch_wait()[0]
    if rc or rc2:
        raise RuntimeError(f"tuning parse failed: {rc} {rc2}")
    ctx.set_timing(True)
    for _ in range(reps):
        ctx.parse_lines_device(dev.ptr, aoffs, sizes, st.ptr, pr.ptr, cap)
    t = [x[2] for x in ctx.timings(reps)]
    ctx.set_timing(False)
    ctx.batch_wait()
    return float(np.median(t))


def tune_lines_chunk(ctx, sample: np.ndarray, candidates: Sequence[int] = CANDIDATES,
                     signature: str = "", use_cache: bool = True) -> Tuple[int, Dict[int, float]]:
    """Chunk size (bytes) with the lowest device time on `sample`, and the
    measured table {chunk_bytes: ms}.  Candidates larger than the sample are
    skipped."""
    key = f"lines-chunk:{signature}:{sample.size}"
    cache = _load_cache() if use_cache else {}
    if key in cache:
        e = cache[key]
        return int(e["best"]), {int(k): v for k, v in e["table"].items()}
    table: Dict[int, float] = {}
    for cb in candidates:
        if cb > sample.size and table:
            continue
        try:
            table[int(cb)] = time_lines(ctx, sample, int(cb))
        except ValueError:
            continue
    best = min(table, key=table.get)
    if use_cache:
        cache[key] = {"best": best, "table": {str(k): v for k, v in table.items()}}
        _store_cache(cache)
    return best, table


def time_device(ctx, run: Callable[[], None], reps: int = 5) -> float:
    """Median device time (ms, from the parse's reset kernel to its last
    phase-3 kernel) of `reps` calls of run(), a parse on `ctx` that waits for
    nothing.  One untimed call first (a HIP graph is captured there)."""
    reps = max(2, min(int(reps), 256))
    run()
    ctx.set_timing(True)
    for _ in range(reps):
        run()
    t = [x[2] for x in ctx.timings(reps)]
    ctx.set_timing(False)
    return float(np.median(t))


def tune_tokenizer_shape(ctx, run: Callable[[], None], signature: str,
                         tiles: bool = True, small: bool = False,
                         tile_candidates: Sequence[int] = TILE_WAVES,
                         small_candidates: Sequence[int] = SMALL_WAVES,
                         reps: int = 5, use_cache: bool = True, rounds: int = 3
                         ) -> Tuple[Tuple[int, int], Dict[str, float]]:
    """Fastest tokenizer shape (tile_waves, small_waves) for the parse that
    run() performs, and the measured table {"tile_waves x small_waves": ms}.
    The tile shape is searched first with the small shape at its default,
    then the small shape with the best tile shape (`tiles` / `small` say
    which kinds of unit the workload has).  The candidates of a search are
    timed in `rounds` interleaved passes (each pass: two untimed calls after
    the switch, then `reps` timed ones), and a candidate's time is the median
    of its passes, so that a drift of the device over the search, or the
    first calls after a switch, favour no candidate.  The winner is left set
    on ctx."""
    key = f"tok-shape:{signature}"
    cache = _load_cache() if use_cache else {}
    if key in cache:
        e = cache[key]
        best = (int(e["best"][0]), int(e["best"][1]))
        ctx.set_tokenizer_shape(*best)
        return best, dict(e["table"])
    table: Dict[str, float] = {}

    def search(shapes):
        runs: Dict[Tuple[int, int], list] = {sh: [] for sh in shapes}
        for _ in range(max(1, int(rounds))):
            for sh in shapes:
                ctx.set_tokenizer_shape(*sh)
                run()
                runs[sh].append(time_device(ctx, run, reps))
        for sh, v in runs.items():
            table[f"{sh[0]}x{sh[1]}"] = round(float(np.median(v)), 4)
        return min(shapes, key=lambda sh: float(np.median(runs[sh])))

    best_t, best_s = ctx.tokenizer_shape()
    if tiles:
        shapes = [(best_t, best_s)] + [(tw, best_s) for tw in tile_candidates if tw != best_t]
        best_t, best_s = search(shapes)
    if small:
        shapes = [(best_t, best_s)] + [(best_t, sw) for sw in small_candidates if sw != best_s]
        best_t, best_s = search(shapes)
    ctx.set_tokenizer_shape(best_t, best_s)
    if use_cache:
        cache[key] = {"best": [best_t, best_s], "table": table}
        _store_cache(cache)
    return (best_t, best_s), table
