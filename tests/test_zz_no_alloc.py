"""SFGS_Colour_Static, the colouring entry point that allocates nothing (its
scratch is a file-scope device array), against SFGS_Colour on the same graphs:
bit-identical colours and colour counts on two calls in a row, the second with
a smaller graph; sizes beyond the stated maximum and graphs that are not valid
CSR are refused before any colouring round runs.  The test itself owns the
device buffers of the inputs and the output (hipMalloc through ctypes)."""
import ctypes as C

import numpy as np
import pytest

import solfec
from solfec._lib import sfgs

pytestmark = pytest.mark.gpu

MAX = 1 << 20
H2D, D2H = 1, 2


def _hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


def random_graph(n, seed, deg=12):
    """Symmetric graph without self loops: each vertex links to `deg` random
    others (duplicates removed), so degrees vary and priorities tie-break."""
    rng = np.random.default_rng(seed)
    a = np.repeat(np.arange(n), deg)
    b = rng.integers(0, n, n * deg)
    keep = a != b
    a, b = a[keep], b[keep]
    e = np.unique(np.concatenate([a * n + b, b * n + a]))
    src, dst = (e // n).astype(np.int32), (e % n).astype(np.int32)
    rowptr = np.zeros(n + 1, np.int32)
    np.add.at(rowptr, src + 1, 1)
    return np.cumsum(rowptr).astype(np.int32), dst


class DeviceGraph:
    def __init__(self, hip, rowptr, col):
        self.hip, self.bufs = hip, []
        self.rowptr = self._put(rowptr)
        self.col = self._put(col if len(col) else np.zeros(1, np.int32))
        self.out = self._alloc(4 * max(len(rowptr) - 1, 1))

    def _alloc(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), nbytes) == 0
        self.bufs.append(p)
        return p

    def _put(self, a):
        a = np.ascontiguousarray(a, np.int32)
        p = self._alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def colours(self, n):
        out = np.zeros(n, np.int32)
        assert self.hip.hipMemcpy(out.ctypes.data, self.out, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)


def test_colour_static_matches_colour_on_two_calls():
    hip, lib = _hip(), sfgs()
    dev = solfec.Device(0)
    try:
        for n, seed in ((200000, 1), (5000, 2)):   # the second call smaller than the first
            print(f"[no-alloc] n={n} seed={seed}", flush=True)
            rowptr, col = random_graph(n, seed)
            nc_ref, ref = dev.colour(rowptr, col)
            g = DeviceGraph(hip, rowptr, col)
            try:
                nc = lib.SFGS_Colour_Static(dev.ctx, n, len(col), g.rowptr, g.col, g.out)
                assert nc == nc_ref, (nc, nc_ref)
                np.testing.assert_array_equal(g.colours(n), ref)
            finally:
                g.free()
    finally:
        dev.close()


def test_colour_static_refuses_bad_sizes_and_graphs():
    hip, lib = _hip(), sfgs()
    dev = solfec.Device(0)
    try:
        rowptr, col = random_graph(1000, 3)
        g = DeviceGraph(hip, rowptr, col)
        try:
            assert lib.SFGS_Colour_Static(dev.ctx, MAX + 1, len(col), g.rowptr, g.col, g.out) == -2
            assert lib.SFGS_Colour_Static(dev.ctx, 1000, len(col) + 5, g.rowptr, g.col, g.out) == -2
            assert lib.SFGS_Colour_Static(dev.ctx, 0, 0, None, None, None) == 0
        finally:
            g.free()
        bad = col.copy()
        bad[7] = 1000                                   # a column outside [0, n)
        g = DeviceGraph(hip, rowptr, bad)
        try:
            assert lib.SFGS_Colour_Static(dev.ctx, 1000, len(bad), g.rowptr, g.col, g.out) == -2
            assert b"not valid CSR" in lib.SFGS_Last_Error(dev.ctx)
        finally:
            g.free()
        # a valid graph after the refusals still colours like SFGS_Colour
        nc_ref, ref = dev.colour(rowptr, col)
        g = DeviceGraph(hip, rowptr, col)
        try:
            assert lib.SFGS_Colour_Static(dev.ctx, 1000, len(col), g.rowptr, g.col, g.out) == nc_ref
            np.testing.assert_array_equal(g.colours(1000), ref)
        finally:
            g.free()
    finally:
        dev.close()
