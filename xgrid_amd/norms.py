"""FP64 norms over one boundary class of device grids.

No counterpart in the reference: its DSL has no reductions (a grid load is
only allowed inside a stencil store, ``xgrid/lang/generator.py:62-65``), so a
convergence test or an error report needs the field on the host.  Here they
run on the device, reading FP32 or FP64 storage and accumulating in FP64:

* ``diff_norm(a, b)``       - L2 and Linf of ``a - b`` (``b`` may be None,
  or a grid of the other precision, e.g. an FP64 twin of an FP32 field);
* ``poisson_residual(p, rhs, coef)`` - L2 and Linf of
  ``sum_a coef[a] (p[+e_a] - 2 p + p[-e_a]) - rhs_scale * rhs``.

Both reduce deterministically (fixed partition, fixed order) and synchronise
the backend stream.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .grid import Grid


def _type_code(g: Grid) -> int:
    if g.numpy_dtype == np.float32:
        return _lib.T_F32
    if g.numpy_dtype == np.float64:
        return _lib.T_F64
    raise TypeError(f"norms need a floating-point grid, not {g.numpy_dtype}")


def diff_norm(a: Grid, b: Grid | None = None, klass: int = 0, level_a: int = 0, level_b: int = 0) -> dict:
    """{'l2', 'linf', 'cells'} of a[level_a] - b[level_b] over class `klass` of a."""
    ha = a._sync_for_launch()
    hb = b._sync_for_launch() if b is not None else None
    l2, linf, cells = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
    _lib.check(_lib.load().xg_grid_norm_diff(ha, level_a, _type_code(a), hb, level_b,
                                             _type_code(b) if b is not None else 0, klass, ctypes.byref(l2),
                                             ctypes.byref(linf), ctypes.byref(cells)), "xg_grid_norm_diff")
    return {"l2": l2.value, "linf": linf.value, "cells": cells.value}


def poisson_residual(p: Grid, rhs: Grid, coef, rhs_scale: float = 1.0, klass: int = 0, level_p: int = 0,
                     level_rhs: int = 0) -> dict:
    """{'l2', 'linf'} of the Poisson residual over class `klass` of p (one coefficient per axis)."""
    if len(coef) != p.dimension:
        raise ValueError(f"need {p.dimension} coefficients, got {len(coef)}")
    hp = p._sync_for_launch()
    hr = rhs._sync_for_launch()
    c = (ctypes.c_double * 3)(*([float(x) for x in coef] + [0.0] * (3 - len(coef))))
    l2, linf = ctypes.c_double(), ctypes.c_double()
    _lib.check(_lib.load().xg_grid_poisson_residual(hp, level_p, _type_code(p), hr, level_rhs, _type_code(rhs), klass,
                                                    c, float(rhs_scale), ctypes.byref(l2), ctypes.byref(linf)),
               "xg_grid_poisson_residual")
    return {"l2": l2.value, "linf": linf.value}
