"""Pins the oracle colouring (or_colour.c): the result is a proper colouring,
it is first-fit in descending priority order (every vertex of colour c has a
higher-priority neighbour of each colour below c, and none of its own colour),
and it equals a brute-force sequential greedy written here from the priority
definition; colour-major order lists colour classes in index order."""
import numpy as np
import pytest

import oracle


def random_graph(rng, n, deg):
    edges = set()
    for _ in range(n * deg // 2):
        a, b = rng.integers(0, n, 2)
        if a != b:
            edges.add((int(a), int(b)))
            edges.add((int(b), int(a)))
    rows = [[] for _ in range(n)]
    for a, b in sorted(edges):
        rows[a].append(b)
    rowptr = np.zeros(n + 1, np.int32)
    rowptr[1:] = np.cumsum([len(r) for r in rows])
    col = np.array([x for r in rows for x in r], np.int32)
    return rowptr, col, rows


def priority(rowptr, v):
    return ((int(rowptr[v + 1] - rowptr[v]) << 32) | int(oracle.lib().OR_Hash32(v)), v)


@pytest.mark.parametrize("n,deg,seed", [(1, 0, 0), (10, 2, 1), (300, 6, 2), (2000, 12, 3)])
def test_colouring_properties(n, deg, seed):
    rng = np.random.default_rng(seed)
    rowptr, col, rows = random_graph(rng, n, deg)
    nc, c = oracle.colour(rowptr, col)
    assert nc == (c.max() + 1 if n else 0)
    for v in range(n):
        nb = rows[v]
        assert all(c[u] != c[v] for u in nb)
        higher = {c[u] for u in nb if priority(rowptr, u) > priority(rowptr, v)}
        assert set(range(c[v])) <= higher
    # brute-force greedy in descending priority
    ref = -np.ones(n, int)
    for v in sorted(range(n), key=lambda v: priority(rowptr, v), reverse=True):
        used = {ref[u] for u in rows[v] if ref[u] >= 0}
        k = 0
        while k in used:
            k += 1
        ref[v] = k
    assert np.array_equal(ref, c)
    order, colptr = oracle.colour_order(c, max(nc, 1))
    if n:
        assert np.array_equal(np.sort(order), np.arange(n))
        for k in range(nc):
            cls = order[colptr[k]:colptr[k + 1]]
            assert np.all(c[cls] == k) and np.all(np.diff(cls) > 0)


def test_colour_count_bounded_by_degree():
    rng = np.random.default_rng(5)
    rowptr, col, rows = random_graph(rng, 3000, 10)
    nc, c = oracle.colour(rowptr, col)
    assert nc <= max(len(r) for r in rows) + 1


def test_hash_is_a_bijection_sample():
    h = {oracle.lib().OR_Hash32(v) for v in range(100000)}
    assert len(h) == 100000
