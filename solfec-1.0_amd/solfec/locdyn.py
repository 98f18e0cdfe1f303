"""Reader and comparison of the LOCDYN_Dump text format (ldy.c:892-1002).

One line per row: ``KIND (point) (master, slave) [base 9] [W_ii 9] (B) (V) (U)
(R)`` followed by `` => KIND (point) (master, slave) [base 9] [W_ij 9]`` for
every neighbour.  ``read`` returns the rows sorted by (kind, point, pair,
base, W_ii) with their neighbours sorted the same way, so that two dumps of one
system written in different row orders compare entry by entry, as the
reference's ``scripts/locdyn_compare`` does.
"""
import re

import numpy as np

_NUM = r"[-+0-9.eEinfa]+"
_VEC = re.compile(r"\(([^()]*)\)")
_MAT = re.compile(r"\[([^\[\]]*)\]")


def _floats(text):
    return tuple(float(x) for x in text.split(","))


def _head(text):
    """kind, point, pair, base, W of 'KIND (p) (m, s) [9] [9]' and the rest of the text."""
    kind, rest = text.strip().split(" ", 1)
    vecs = _VEC.findall(rest)
    mats = _MAT.findall(rest)
    point = _floats(vecs[0])
    pair = tuple(int(x) for x in vecs[1].split(","))
    return kind, point, pair, _floats(mats[0]), _floats(mats[1]), vecs[2:]


def read(path):
    """Returns (rows, neighbours, BVUR): rows = [(kind, point, pair, base, W)] sorted, the
    neighbour lists in the same order, and (B, V, U, R) per row."""
    rows, offd, bvur = [], [], []
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            if not line:
                continue
            parts = line.split(" => ")
            kind, point, pair, base, W, rest = _head(parts[0])
            rows.append((kind, point, pair, base, W))
            bvur.append(tuple(_floats(v) for v in rest[:4]))
            adj = []
            for p in parts[1:]:
                k, pt, pr, bs, w, _ = _head(p)
                adj.append((k, pt, pr, bs, w))
            offd.append(sorted(adj))
    perm = sorted(range(len(rows)), key=lambda k: rows[k])
    return [rows[i] for i in perm], [offd[i] for i in perm], [bvur[i] for i in perm]


def compare(path1, path2, rtol=0.0, atol=0.0):
    """True when two dumps hold the same rows, neighbours and B, V, U, R, entry by entry
    within |a - b| <= atol + rtol |b|."""
    d1, o1, v1 = read(path1)
    d2, o2, v2 = read(path2)
    if len(d1) != len(d2):
        return False

    def close(a, b):
        if a[0] != b[0] or a[2] != b[2]:
            return False
        return all(np.allclose(x, y, rtol=rtol, atol=atol) for x, y in zip(a[1:], b[1:]) if not isinstance(x, (str, int)))

    for i in range(len(d1)):
        if not close(d1[i], d2[i]) or len(o1[i]) != len(o2[i]):
            return False
        if not all(close(a, b) for a, b in zip(o1[i], o2[i])):
            return False
        if not all(np.allclose(a, b, rtol=rtol, atol=atol) for a, b in zip(v1[i], v2[i])):
            return False
    return True
