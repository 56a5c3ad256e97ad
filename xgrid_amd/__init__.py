"""xgrid on MI355X: the xgrid stencil DSL with a HIP/gfx950 execution backend.

Public names match the reference package (``xgrid/__init__.py:52-53``):
``kernel, function, init, ptr, grid, boundary, c, external, Grid, shape,
dimension, tick``.
"""
from __future__ import annotations

import struct
from typing import Any

from .config import init, get_config
from .grid import Grid
from .lang import boundary, c
from .operator import external, function, kernel, precompile
from .types import GridType as _GridType
from .types import Integer, Void, grid, ptr


def _dimension_typecheck(args):
    if not isinstance(args[0], _GridType):
        raise Exception(f"Incompatible dimension to type '{args[0]}'")
    return Integer(struct.calcsize("i"))


def _shape_typecheck(args):
    if not isinstance(args[0], _GridType):
        raise Exception(f"Incompatible shape to type '{args[0]}'")
    if not isinstance(args[1], Integer):
        raise Exception(f"Incompatible shape dimension '{args[1]}'")
    return Integer(struct.calcsize("i"))


def _tick_typecheck(args):
    if not isinstance(args[0], _GridType):
        raise Exception(f"Incompatible tick to type '{args[0]}'")
    return Void()


@external(typecheck_override=_dimension_typecheck)
def dimension(grid: Any) -> int:
    ...


@external(typecheck_override=_shape_typecheck)
def shape(grid: Any, dimension: int) -> int:
    ...


@external(typecheck_override=_tick_typecheck)
def tick(grid: Any) -> None:
    ...


dimension.intrinsic = "dimension"
shape.intrinsic = "shape"
tick.intrinsic = "tick"

__all__ = ["kernel", "function", "init", "ptr", "grid", "boundary", "c", "external", "Grid", "shape",
           "dimension", "tick"]
