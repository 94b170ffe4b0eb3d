"""Scalar restatements used to pin the oracle (test helpers, pure Python).

These are formulated differently from the oracle's word-level bit
arithmetic: a byte-at-a-time state machine for escapes and strings, a stack
for bracket matching, a byte-level form of the UTF-8 word rule, and Python's
stable ``sorted`` for the depth sort.  They are only meant for small inputs.
"""
from __future__ import annotations

OPEN = (0x5B, 0x7B)
CLOSE = (0x5D, 0x7D)
BRACKETS = OPEN + CLOSE


def op_set(jsonl: bool):
    s = {0x5B, 0x5D, 0x7B, 0x7D, 0x3A, 0x2C}
    if jsonl:
        s.add(0x0A)
    return s


def lex(data: bytes, jsonl: bool = False):
    """1-based positions of structural characters outside strings.

    A quote toggles the string state unless the run of backslashes right
    before it (counted from the start, in or out of strings) is odd."""
    ops = op_set(jsonl)
    out = []
    in_str = False
    run = 0
    for p, c in enumerate(data):
        if c == 0x22 and run % 2 == 0:
            in_str = not in_str
        elif not in_str and c in ops:
            out.append(p + 1)
        run = run + 1 if c == 0x5C else 0
    return out


def stack_pairs(data: bytes, positions):
    """Bracket matching with a stack over the structural positions.

    Returns (ok, pair) where pair[t] is the index t' of the closer matching
    the opener at t (0-based structural indices), or None if mismatched."""
    pair = {}
    stack = []
    for t, p in enumerate(positions):
        c = data[p - 1]
        if c in OPEN:
            stack.append((t, c))
        elif c in CLOSE:
            if not stack:
                return False, None
            t0, c0 = stack.pop()
            if (c0 ^ c) != 0x06:
                return False, None
            pair[t0] = t
    if stack:
        return False, None
    return True, pair


def max_depth(data: bytes, positions):
    d = m = 0
    for p in positions:
        c = data[p - 1]
        if c in OPEN:
            d += 1
            m = max(m, d)
        elif c in CLOSE:
            d -= 1
    return m


def utf8_word_rule(data: bytes) -> bool:
    """Byte-level form of the reference's per-word check over zero-padded
    32-bit words; True if the unit is flagged."""
    n = len(data)
    if not any(b >= 0x80 for b in data):
        return False
    W4 = (n + 3) // 4
    padded = bytes(data) + b"\x00" * (4 * W4 - n)

    def cont(b):
        return 0x80 <= b <= 0xBF

    for j in range(W4):
        word = padded[4 * j:4 * j + 4]
        prev = padded[4 * j - 4:4 * j] if j else b"\x00\x00\x00\x00"
        if all(b < 0x80 for b in word):
            if prev[3] >= 0xC0 or prev[2] >= 0xE0 or prev[1] >= 0xF0:
                return True
            continue
        seq = prev + word
        for i in range(4):
            b = seq[4 + i]
            p1, p2, p3 = seq[3 + i], seq[2 + i], seq[1 + i]
            if (cont(p1) and cont(b)) != (p2 >= 0xE0 or p3 >= 0xF0):
                return True
    return False


def pair_by_sorted_groups(chars, idx, u16=False, base=0):
    """The group-of-four pairing over a stable sort by depth key.

    Returns (err, writes) with writes = {idx_of_first: idx_of_second+base+1}."""
    mask = 0xFFFF if u16 else 0xFF
    depth = 0
    keys = []
    for c in chars:
        if c in OPEN:
            depth = (depth + 1) & mask
            keys.append((depth - 1) & mask)
        else:
            depth = (depth - 1) & mask
            keys.append(depth)
    order = sorted(range(len(chars)), key=lambda m: keys[m])
    M = len(chars)
    err = False
    writes = {}
    groups = (M + 3) // 4

    def sc(x):  # signed char
        return x - 256 if x >= 128 else x

    for g in range(groups):
        k = 4 * g
        rem = M - k
        last = g == groups - 1
        if last and rem in (1, 3):
            err = True
            continue
        for a in (k, k + 2):
            if a + 1 >= M or (last and rem == 2 and a == k + 2):
                continue
            c0, c1 = sc(chars[order[a]]), sc(chars[order[a + 1]])
            if (c0 ^ c1) != 6:
                err = True
            if not last and c0 > c1:
                err = True
            writes[idx[order[a]]] = idx[order[a + 1]] + base + 1
    return err, writes


def scalar_parse(data: bytes, jsonl=False, u16=False):
    """(code, structural positions, pair row with -1 defaults) by the scalar forms."""
    if len(data) == 0:
        return 1, [], []
    pos = lex(data, jsonl)
    chars, idx = [], []
    for t, p in enumerate(pos):
        c = data[p - 1]
        if c in BRACKETS:
            chars.append(c)
            idx.append(t)
    err, writes = pair_by_sorted_groups(chars, idx, u16)
    row = [-1] * len(pos)
    for t, v in writes.items():
        row[t] = v
    code = 2 if utf8_word_rule(data) else (3 if err else 0)
    return code, pos, row
