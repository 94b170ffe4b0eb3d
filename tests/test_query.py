"""Known-answer queries through the iterators of include/cujson/query_iterator.h
(the reference's cuJSONIterator and cuJSONLinesIterator interfaces): paths
drawn at random from Python's own parse of the same documents, each answer
compared with the value Python finds there.

Each query set runs twice: over the oracle's parse result on the CPU (the
driver's --result form, which pins the iterators themselves) and, marked gpu,
over this build's own parse through the reference-shaped C++ API."""
import json
import os
import random
import subprocess

import numpy as np
import pytest

import oracle
from cujson_amd import gen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "cujson_amd", "cjx_query_check")
LARGE = os.path.join(ROOT, "tests", "golden", "twitter_sample_large_record.json")
SMALL = os.path.join(ROOT, "tests", "golden", "twitter_sample_small_records.json")
ON_GPU = [pytest.param("cpu"), pytest.param("gpu", marks=pytest.mark.gpu)]


def matches(raw: bytes, want) -> bool:
    s = raw.decode("utf-8")
    if isinstance(want, str):
        return json.loads('"' + s + '"') == want
    return json.loads(s) == want


def plain_keys(d):
    """Keys a query line can carry (no spaces, quotes or escapes)."""
    return [k for k in d if k.isascii() and k.isprintable() and " " not in k
            and '"' not in k and "\\" not in k]


def run(tmp_path, where, path, lines, kind="std", chunks=1):
    """Runs query lines through the driver; `where` = "gpu" parses with this
    build, "cpu" hands the driver the oracle's result."""
    qf = str(tmp_path / "queries.txt")
    with open(qf, "w") as f:
        f.write("\n".join(lines) + "\n")
    if where == "gpu":
        args = [BIN, path, qf] if kind == "std" else [BIN, "--lines", str(chunks), path, qf]
    else:
        with open(path, "rb") as f:
            data = f.read()
        o = (oracle.parse_standard(data) if kind == "std"
             else oracle.parse_lines(data, oracle.chunks_by_count(data, chunks)))
        assert o.code == oracle.OK
        rf = str(tmp_path / "result.i32")
        o.buf.astype(np.int32).tofile(rf)
        args = [BIN, "--result", kind, rf, path, qf]
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.returncode, r.stderr[-500:])
    return [ln for ln in r.stdout.splitlines() if ln.startswith(("ok ", "miss ", "scan "))]


def check_answers(queries, out):
    assert len(out) == len(queries)
    for (steps, want), line in zip(queries, out):
        assert line.startswith("ok "), (steps, line)
        raw = bytes.fromhex(line[3:])
        assert matches(raw, want), (steps, raw[:200], repr(want)[:200])


# ---- standard JSON ---------------------------------------------------------


def walk(doc, rng, max_steps):
    """A path from the root ("i:0" enters it, as in the reference's mains)
    and the value the iterator's getValue() gives at its end.  An array
    element that is itself an array leaves the iterator on its '[', where
    getValue() gives the first element (the reference's behaviour)."""
    steps, cur, last_index = ["i:0"], doc, False
    for _ in range(rng.randint(1, max_steps)):
        if isinstance(cur, list) and cur:
            j = rng.randrange(len(cur))
            steps.append(f"i:{j}")
            cur, last_index = cur[j], True
        elif isinstance(cur, dict) and cur:
            keys = plain_keys(cur)
            if not keys:
                break
            k = rng.choice(keys)
            steps.append(f"k:{k}")
            cur, last_index = cur[k], False
        else:
            break
    if len(steps) == 1:
        return None
    if last_index and isinstance(cur, list):
        if not cur:
            return None
        cur = cur[0]
    return steps, cur


def check_doc(tmp_path, where, path, n_queries, seed):
    with open(path, "rb") as f:
        doc = json.loads(f.read())
    rng = random.Random(seed)
    queries = []
    while len(queries) < n_queries:
        q = walk(doc, rng, 6)
        if q is not None:
            queries.append(q)
    check_answers(queries, run(tmp_path, where, path, [" ".join(s) for s, _ in queries]))


def test_query_program_is_built():
    assert os.path.exists(BIN), "cjx_query_check is not built (run __graft_entry__.build())"
    assert subprocess.run([BIN], capture_output=True).returncode == 2


@pytest.mark.parametrize("where", ON_GPU)
def test_queries_twitter_large_record(tmp_path, where):
    check_doc(tmp_path, where, LARGE, 400, 1)


@pytest.mark.parametrize("where", ON_GPU)
def test_queries_generated_document(tmp_path, where):
    p = str(tmp_path / "config1.json")
    gen.config1().tofile(p)
    check_doc(tmp_path, where, p, 400, 2)


@pytest.mark.parametrize("where", ON_GPU)
def test_queries_deep_document(tmp_path, where):
    p = str(tmp_path / "deep.json")
    gen.standard_doc(n_objects=300, deep_depth=40, seed=5).tofile(p)
    check_doc(tmp_path, where, p, 300, 3)


# ---- JSON Lines ------------------------------------------------------------
# The records are the elements of a virtual top-level array: element e is the
# text between the e-th and (e+1)-th newline, so blank lines are empty
# elements.  cuJSONLinesIterator.gotoArrayIndex stops on the separator before
# the element it reaches (the '[' for element 0); when element 0 of an array
# is itself an array, a further index step walks the outer array again (the
# reference's behaviour), so a path ends there.


def line_elements(data: bytes):
    out = []
    for text in data.split(b"\n"):
        try:
            out.append(json.loads(text))
        except ValueError:
            out.append(None)  # blank line: an empty element
    return out


def walk_lines(elems, rng, max_steps):
    recs = [e for e, v in enumerate(elems) if isinstance(v, dict) and v]
    r = rng.choice(recs)
    steps, cur = [], elems[r]
    if rng.random() < 0.3 and r > 0:  # reach it as a sibling of an earlier element
        back = rng.randint(1, min(r, 5))
        steps += [f"i:{r - back}", f"s:{back}"]
    else:
        steps.append(f"i:{r}")
    for _ in range(rng.randint(0, max_steps)):
        if isinstance(cur, list) and cur:
            j = rng.randrange(len(cur))
            steps.append(f"i:{j}")
            cur = cur[j]
            if j == 0 and isinstance(cur, list):
                break
        elif isinstance(cur, dict) and cur:
            keys = plain_keys(cur)
            if not keys:
                break
            k = rng.choice(keys)
            steps.append(f"k:{k}")
            cur = cur[k]
        else:
            break
    return steps, cur


def check_lines_doc(tmp_path, where, path, n_queries, seed, chunks):
    with open(path, "rb") as f:
        elems = line_elements(f.read())
    rng = random.Random(seed)
    queries = [walk_lines(elems, rng, 5) for _ in range(n_queries)]
    out = run(tmp_path, where, path, [" ".join(s) for s, _ in queries], "lines", chunks)
    check_answers(queries, out)


def value_text(v):
    """The text getValue() gives for a scalar: a string without its quotes,
    true / false / null as written."""
    return v if isinstance(v, str) else json.dumps(v)


def check_scan(tmp_path, where, path, chunks, key, value, key2):
    """The loop of the reference's real-world use case: every record whose
    top-level `key` reads `value` contributes its top-level `key2`."""
    with open(path, "rb") as f:
        elems = line_elements(f.read())
    want = [e[key2] for e in elems if isinstance(e, dict) and key in e and key2 in e
            and not isinstance(e[key], (dict, list)) and value_text(e[key]) == value]
    out = run(tmp_path, where, path, [f"scan {key} {value} {key2} {len(elems)}"], "lines", chunks)
    assert len(out) == 1 and out[0].startswith("scan "), out
    parts = out[0].split(" ")
    got = bytes.fromhex(parts[2]).split(b"\n")[:-1] if len(parts) > 2 and parts[2] else []
    assert int(parts[1]) == len(want) == len(got), (parts[1], len(want), len(got))
    for raw, w in zip(got, want):
        assert matches(raw, w), (raw[:200], repr(w)[:200])
    return len(want)


@pytest.mark.parametrize("where", ON_GPU)
@pytest.mark.parametrize("chunks", [1, 3])
def test_lines_queries_twitter(tmp_path, where, chunks):
    check_lines_doc(tmp_path, where, SMALL, 300, 10 + chunks, chunks)


@pytest.mark.parametrize("where", ON_GPU)
def test_lines_scan_twitter(tmp_path, where):
    assert check_scan(tmp_path, where, SMALL, 2, "lang", "en", "id") > 0


@pytest.mark.parametrize("where", ON_GPU)
def test_lines_queries_generated(tmp_path, where):
    p = str(tmp_path / "records.jsonl")
    gen.lines(3000, seed=17).tofile(p)
    check_lines_doc(tmp_path, where, p, 400, 21, 4)
    assert check_scan(tmp_path, where, p, 4, "flag", "true", "id") > 0
