"""Builds libcujson_amd.so (HIP kernels + C ABI) in-tree for gfx950.

Usage: python -m cujson_amd.build   (or __graft_entry__.build())
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcujson_amd.so")
ARCH = os.environ.get("CJX_OFFLOAD_ARCH", "gfx950")

SOURCES = ["cjx_tokenize.hip", "cjx_utf8.hip", "cjx_kernels.hip", "cjx_match.hip", "cjx_records.hip",
           "cjx_lines.hip", "cjx_values.hip", "cjx_noalloc.hip", "cjx_api.cpp", "cjx_chunk.cpp",
           "cjx_server.cpp"]
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed fp32 VALU ops are disabled for every HIP object of this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "unknown target feature" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    return out


def build_library(verbose: bool = False) -> str:
    from concurrent.futures import ThreadPoolExecutor
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
        cmd = ["hipcc", *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-c",
               os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd.insert(1, "-x")
            cmd.insert(2, "hip")
        cmds.append(cmd)
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=min(len(cmds), 8)) as pool:
        for out in pool.map(_run, cmds):
            if verbose and out:
                print(out)
    tmp = LIB + ".tmp"
    _run(["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


GEN_LIB = os.path.join(HERE, "libcjx_gen.so")


def build_generator() -> str:
    tmp = GEN_LIB + ".tmp"
    _run(["gcc", "-O2", "-std=c11", "-Wall", "-fPIC", "-shared", "-o", tmp,
          os.path.join(CSRC, "cjx_gen.c"), "-lm"])
    os.replace(tmp, GEN_LIB)
    return GEN_LIB


WRAPPER_CHECK = os.path.join(HERE, "cjx_wrapper_check")


def build_wrapper_check() -> str:
    """The C++ program that drives the reference-shaped headers (include/cujson/)."""
    tmp = WRAPPER_CHECK + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-o", tmp,
          os.path.join(ROOT, "tests", "cpp", "wrapper_check.cpp"), "-L", HERE, "-lcujson_amd",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, WRAPPER_CHECK)
    return WRAPPER_CHECK


QUERY_CHECK = os.path.join(HERE, "cjx_query_check")


def build_query_check() -> str:
    """The C++ program that runs queries through include/cujson/query_iterator.h."""
    tmp = QUERY_CHECK + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-o", tmp,
          os.path.join(ROOT, "tests", "cpp", "query_check.cpp"), "-L", HERE, "-lcujson_amd",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, QUERY_CHECK)
    return QUERY_CHECK


LOADGEN_LIB = os.path.join(HERE, "libcjx_loadgen.so")


def build_loadgen() -> str:
    """The serving load driver (bench tool) over the library and the generator."""
    tmp = LOADGEN_LIB + ".tmp"
    _run(["hipcc", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", os.path.join(ROOT, "include"),
          "-o", tmp, os.path.join(CSRC, "cjx_loadgen.cpp"), "-L", HERE, "-lcujson_amd",
          "-lcjx_gen", "-Wl,-rpath,$ORIGIN", "-lpthread"])
    os.replace(tmp, LOADGEN_LIB)
    return LOADGEN_LIB


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"], check=True)


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_generator())
    print(build_wrapper_check())
    print(build_query_check())
    print(build_loadgen())
    build_oracle()
