"""Known-answer queries through the iterator of include/cujson/query_iterator.h
(the reference's cuJSONIterator interface) on this build's parse results: paths
drawn at random from Python's own parse of the same documents, each answer
compared with the value Python finds there."""
import json
import os
import random
import subprocess

import pytest

from cujson_amd import gen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "cujson_amd", "cjx_query_check")
LARGE = os.path.join(ROOT, "tests", "golden", "twitter_sample_large_record.json")


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
            keys = [k for k in cur if k.isascii() and k.isprintable() and " " not in k
                    and '"' not in k and "\\" not in k]
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


def matches(raw: bytes, want) -> bool:
    s = raw.decode("utf-8")
    if isinstance(want, str):
        return json.loads('"' + s + '"') == want
    return json.loads(s) == want


def run_queries(path, queries):
    qf = path + ".queries"
    with open(qf, "w") as f:
        for steps, _ in queries:
            f.write(" ".join(steps) + "\n")
    r = subprocess.run([BIN, path, qf], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.returncode, r.stderr[-500:])
    return [ln for ln in r.stdout.splitlines() if ln.startswith(("ok ", "miss "))]


def check_doc(path, n_queries, seed):
    with open(path, "rb") as f:
        doc = json.loads(f.read())
    rng = random.Random(seed)
    queries = []
    while len(queries) < n_queries:
        q = walk(doc, rng, 6)
        if q is not None:
            queries.append(q)
    out = run_queries(path, queries)
    assert len(out) == len(queries)
    for (steps, want), line in zip(queries, out):
        assert line.startswith("ok "), (steps, line)
        raw = bytes.fromhex(line[3:])
        assert matches(raw, want), (steps, raw[:200], repr(want)[:200])


def test_query_program_is_built():
    assert os.path.exists(BIN), "cjx_query_check is not built (run __graft_entry__.build())"
    assert subprocess.run([BIN], capture_output=True).returncode == 2


@pytest.mark.gpu
def test_queries_twitter_large_record():
    check_doc(LARGE, 400, 1)


@pytest.mark.gpu
def test_queries_generated_document(tmp_path):
    p = str(tmp_path / "config1.json")
    gen.config1().tofile(p)
    check_doc(p, 400, 2)


@pytest.mark.gpu
def test_queries_deep_document(tmp_path):
    p = str(tmp_path / "deep.json")
    d = gen.standard_doc(n_objects=300, deep_depth=40, seed=5)
    d.tofile(p)
    check_doc(p, 300, 3)
