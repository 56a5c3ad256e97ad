"""The allocation-free norm entry point, xg_grid_norm_diff_noalloc, against the
allocating one, xg_grid_norm_diff, through the C ABI: equal results on two
calls in a row, the second on a smaller input with a list class, with the
allocation-free one called first (before the allocating one has ever run in
the process)."""
import ctypes
from ctypes import byref, c_double, c_int64

import numpy as np
import pytest

from xgrid_amd import _lib

pytestmark = pytest.mark.gpu


def _call(fn, a, b, klass):
    l2, linf, cells = c_double(), c_double(), c_int64()
    rc = fn(a, 0, _lib.T_F64, b, 0, _lib.T_F64, klass, byref(l2), byref(linf), byref(cells))
    assert rc == 0, _lib.last_error()
    return l2.value, linf.value, cells.value


def test_noalloc_norm_equals_allocating_norm(double):
    lib = _lib.load()
    rng = np.random.default_rng(31)
    cases = [((700, 900), 0, None), ((130, 70), 1, "list")]
    for shape, klass, kind in cases:
        print(f"no-alloc norm: shape {shape} class {klass}", flush=True)
        a = double.Grid(shape, float)
        b = double.Grid(shape, float)
        av, bv = rng.standard_normal(shape), rng.standard_normal(shape)
        a.now[...] = av
        b.now[...] = bv
        mask = np.zeros(shape, np.int32)
        if kind == "list":
            mask[::7, ::5] = 1                       # a sparse class: a list of cells
        a.boundary[...] = mask
        ha, hb = a._sync_for_launch(), b._sync_for_launch()
        free = _call(lib.xg_grid_norm_diff_noalloc, ha, hb, klass)
        alloc = _call(lib.xg_grid_norm_diff, ha, hb, klass)
        assert free == alloc
        sel = mask == klass
        d = (av - bv)[sel]
        assert free[2] == sel.sum()
        assert free[1] == np.abs(d).max()
        assert abs(free[0] - np.sqrt((d * d).sum())) <= 1e-12 * free[0]
        again = _call(lib.xg_grid_norm_diff_noalloc, ha, hb, klass)
        assert again == free
