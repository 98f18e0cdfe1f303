"""The LOCDYN_Dump text format (ldy.c:892-1002) and the MatrixMarket export of
W (ldy.c:1005-1052) written by the host layer from an oracle-assembled system:
read back with solfec.locdyn (the reading rules of the reference's
scripts/locdyn_compare) and with scipy.io.mmread, and compared with the
arrays they were written from."""
import ctypes as C

import numpy as np
import pytest
import scipy.io

import oracle
import solfec
from solfec import locdyn
from solfec._lib import ptr, solfec_lib
from scenes import lattice


def _world_points(b, c):
    n = len(c["gap"])
    x, x0, rot = b["x"].reshape(-1, 3), b["x0"].reshape(-1, 3), b["rot"].reshape(-1, 3, 3).transpose(0, 2, 1)
    m = c["master"]
    X = c["mpnt"].reshape(n, 3)
    return np.ascontiguousarray(np.stack([x[m[i]] + rot[m[i]] @ (X[i] - x0[m[i]]) for i in range(n)]))


@pytest.fixture(scope="module")
def system(tmp_path_factory):
    s = lattice(125, seed=6)
    b, c = s.bodies(), s.contacts()
    O = oracle.assemble(oracle.bodies_from(b), oracle.contacts_from(c), 1e-3)
    n = O["n"]
    rng = np.random.default_rng(1)
    R, U = rng.normal(size=3 * n), rng.normal(size=3 * n)
    point = _world_points(b, c)
    d = tmp_path_factory.mktemp("dump")
    lib = solfec_lib()
    f = np.ascontiguousarray
    args = [f(point), f(c["master"], np.int32), f(c["slave"], np.int32), f(c["base"]), f(O["Wd"]), f(O["B"]),
            f(O["V"]), f(U), f(R), f(O["rowptr"], np.int32), f(O["col"], np.int32), f(O["W"])]
    assert lib.SOLFEC_Locdyn_Write(str(d / "ldy.txt").encode(), n, *[ptr(a) for a in args]) == 0
    assert lib.SOLFEC_W_MatrixMarket_Write(str(d / "W.mtx").encode(), n, ptr(args[4]), ptr(args[9]), ptr(args[10]),
                                           ptr(args[11])) == 0
    return O, c, point, R, U, d


def test_dump_lines_and_values(system):
    O, c, point, R, U, d = system
    n = O["n"]
    lines = open(d / "ldy.txt").read().splitlines()
    assert len(lines) == n
    # list order: the newest contact first (ldy.c:505-508)
    assert lines[0].startswith("CONTACT (%.6g, %.6g, %.6g) (%d, %d)" % (*point[n - 1], c["master"][n - 1],
                                                                         c["slave"][n - 1]))
    rows, offd, bvur = locdyn.read(d / "ldy.txt")
    # map rows back to contacts by (point, pair) and check every value at the printed precision
    index = {(tuple(np.round(point[i], 12)), (int(c["master"][i]), int(c["slave"][i]))): i for i in range(n)}
    for (kind, pt, pair, base, W), adj, (B, V, Uo, Ro) in zip(rows, offd, bvur):
        assert kind == "CONTACT"
        i = min(index, key=lambda k: np.abs(np.array(k[0]) - pt).max() + (k[1] != pair) * 1e9)
        i = index[i]
        np.testing.assert_allclose(W, O["Wd"][9 * i:9 * i + 9], rtol=1e-5, atol=1e-12)
        np.testing.assert_allclose(B, O["B"][3 * i:3 * i + 3], rtol=1e-5, atol=1e-12)
        np.testing.assert_allclose(V, O["V"][3 * i:3 * i + 3], rtol=1e-5, atol=1e-12)
        np.testing.assert_allclose(Ro, R[3 * i:3 * i + 3], rtol=1e-5, atol=1e-12)
        np.testing.assert_allclose(Uo, U[3 * i:3 * i + 3], rtol=1e-5, atol=1e-12)
        # neighbours: one entry per neighbouring contact, blocks through both bodies added
        blocks = {}
        for q in range(O["rowptr"][i], O["rowptr"][i + 1]):
            j = int(O["col"][q])
            blocks[j] = blocks.get(j, 0.0) + O["W"][9 * q:9 * q + 9]
        assert len(adj) == len(blocks)
        for j, Wij in blocks.items():
            hit = [a for a in adj if a[2] == (int(c["master"][j]), int(c["slave"][j]))
                   and np.allclose(a[1], point[j], rtol=1e-5, atol=1e-9)]
            assert len(hit) == 1
            np.testing.assert_allclose(hit[0][4], Wij, rtol=1e-5, atol=1e-12)
    assert locdyn.compare(d / "ldy.txt", d / "ldy.txt")


def test_matrix_market_is_W(system):
    O, c, point, R, U, d = system
    n = O["n"]
    M = scipy.io.mmread(str(d / "W.mtx")).toarray()
    assert M.shape == (3 * n, 3 * n)
    dense = np.zeros((3 * n, 3 * n))
    pos = lambda i: 3 * (n - 1 - i)   # list order
    for i in range(n):
        dense[pos(i):pos(i) + 3, pos(i):pos(i) + 3] = O["Wd"][9 * i:9 * i + 9].reshape(3, 3).T
        for q in range(O["rowptr"][i], O["rowptr"][i + 1]):
            j = int(O["col"][q])
            dense[pos(i):pos(i) + 3, pos(j):pos(j) + 3] += O["W"][9 * q:9 * q + 9].reshape(3, 3).T
    np.testing.assert_allclose(M, dense, rtol=1e-14, atol=1e-18)
    np.testing.assert_allclose(M, M.T, atol=1e-14 * np.abs(M).max())
