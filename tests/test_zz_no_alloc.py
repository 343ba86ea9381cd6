"""The entry point that allocates nothing, fluid_remove_mean_noalloc, through the C ABI:
on caller-owned device memory, its result equals fluid_remove_mean's (the existing entry
point, host arrays) bit for bit on two calls in a row, the second on a smaller grid (its
static scratch keeps the first call's contents); a grid without device tables and a grid
beyond the scratch's 6,291,456 cells are refused with an error, not a fault."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc_remove_mean(ggb, grid, f):
    L = ggb.lib()
    n = grid.n
    d = L.ggbk_malloc(4 * n)  # the caller's buffer
    assert d
    try:
        a = np.ascontiguousarray(f, np.float32)
        assert L.ggbk_h2d(d, a.ctypes.data, 4 * n) == 0
        rc = L.fluid_remove_mean_noalloc(C.byref(grid.c), d)
        L.ggb_sync()
        out = np.empty(n, np.float32)
        assert L.ggbk_d2h(out.ctypes.data, d, 4 * n) == 0
    finally:
        L.ggbk_free(d)
    return rc, out


def test_remove_mean_noalloc_matches_remove_mean(ggb, gpu):
    rng = np.random.default_rng(11)
    for N in (48, 20):  # the second call on a smaller grid
        g = ggb.Grid(N)
        f = (3.0 + rng.standard_normal(g.n)).astype(np.float32)
        ref = g.remove_mean(f)  # also puts the grid's tables on the device
        rc, out = _noalloc_remove_mean(ggb, g, f)
        assert rc == 0
        assert np.array_equal(out.view(np.uint32), ref.view(np.uint32)), N
        assert abs(float(np.sum(g.area.astype(np.float64) * out))) < 1e-3 * float(np.sum(g.area * np.abs(out)))


def test_remove_mean_noalloc_refuses(ggb, gpu):
    L = ggb.lib()
    g = ggb.Grid(8)  # tables never uploaded
    rc, _ = _noalloc_remove_mean(ggb, g, np.ones(g.n, np.float32))
    assert rc == -1
    big = ggb.CCubeSphere()
    C.memmove(C.byref(big), C.byref(g.c), C.sizeof(ggb.CCubeSphere))
    big.ncells = 6291457
    assert L.fluid_remove_mean_noalloc(C.byref(big), None) == -2
