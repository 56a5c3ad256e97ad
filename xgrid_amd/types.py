"""Type lattice of the DSL and the ``grid[T, D]`` / ``ptr[T]`` annotations.

Restates the reference's typing package (``xgrid/util/typing/__init__.py:6-29``,
``value.py:40-113``, ``reference.py:14-65``, ``annotation.py:25-68``): the
same classes of values (bool, 1/2/4/8-byte int, 4/8-byte float, dataclass
structures), the same reference types (pointer, grid), exact-equality typing
and the same annotation rules (``int`` is int32, ``float`` follows
``precision``).  Layout helpers for ctypes, numpy and C/HIP source text live
here so the emitters and the Grid agree on one definition.
"""
from __future__ import annotations

import ctypes
import dataclasses
import struct
import threading
import typing
from dataclasses import dataclass, fields, is_dataclass
from typing import Any

import numpy as np

from .config import get_config


class DSLTypeError(TypeError):
    pass


@dataclass(eq=True, frozen=True)
class BaseType:
    @property
    def ctype(self):  # pragma: no cover - overridden
        raise DSLTypeError(f"type {self!r} has no C representation")


@dataclass(eq=True, frozen=True)
class Void(BaseType):
    def __repr__(self) -> str:
        return "Void"


@dataclass(eq=False, frozen=True)
class Ignore(BaseType):
    """``typing.Any``: compares equal to every type (``typing/__init__.py:27-29``)."""

    def __eq__(self, other):
        return True

    def __hash__(self):
        return 0

    def __repr__(self) -> str:
        return "Any"


@dataclass(eq=True, frozen=True)
class Value(BaseType):
    pass


@dataclass(eq=True, frozen=True)
class Boolean(Value):
    @property
    def ctype(self):
        return ctypes.c_bool

    @property
    def cname(self) -> str:
        return "bool"

    @property
    def abbr(self) -> str:
        return "b"

    @property
    def nbytes(self) -> int:
        return 1

    def __repr__(self) -> str:
        return "Boolean"


@dataclass(eq=True, frozen=True)
class Number(Value):
    width_bytes: int = 4

    @property
    def width_bits(self) -> int:
        return self.width_bytes * 8

    @property
    def nbytes(self) -> int:
        return self.width_bytes


_INT_CTYPES = {1: ctypes.c_int8, 2: ctypes.c_int16, 4: ctypes.c_int32, 8: ctypes.c_int64}


@dataclass(eq=True, frozen=True)
class Integer(Number):
    def __post_init__(self):
        if self.width_bytes not in _INT_CTYPES:
            raise DSLTypeError(f"unsupported integer width {self.width_bytes}")

    @property
    def ctype(self):
        return _INT_CTYPES[self.width_bytes]

    @property
    def cname(self) -> str:
        return f"int{self.width_bits}_t"

    @property
    def abbr(self) -> str:
        return f"i{self.width_bits}"

    def __repr__(self) -> str:
        return f"Integer({self.width_bits})"


@dataclass(eq=True, frozen=True)
class Floating(Number):
    def __post_init__(self):
        if self.width_bytes not in (4, 8):
            raise DSLTypeError(f"unsupported float width {self.width_bytes}")

    @property
    def ctype(self):
        return ctypes.c_float if self.width_bytes == 4 else ctypes.c_double

    @property
    def cname(self) -> str:
        return "float" if self.width_bytes == 4 else "double"

    @property
    def abbr(self) -> str:
        return "f" if self.width_bytes == 4 else "d"

    def __repr__(self) -> str:
        return f"Floating({self.width_bits})"


_STRUCT_CTYPES: dict = {}
_STRUCT_LOCK = threading.Lock()     # one ctypes type per structure, whichever thread asks first


@dataclass(eq=True, frozen=True)
class Structure(Value):
    """A dataclass passed by value; C layout is the natural (padded) one."""

    dataclass: type = None
    name: str = ""
    elements: tuple = ()

    @property
    def elements_map(self) -> dict:
        return dict(self.elements)

    @property
    def ctype(self):
        key = (self.dataclass, self.name, self.elements)
        ct = _STRUCT_CTYPES.get(key)
        if ct is None:
            fields_ = [(n, t.ctype) for n, t in self.elements]
            with _STRUCT_LOCK:
                ct = _STRUCT_CTYPES.get(key)
                if ct is None:
                    ct = _STRUCT_CTYPES[key] = type(f"st{self.name}", (ctypes.Structure,), {"_fields_": fields_})
        return ct

    @property
    def cname(self) -> str:
        return f"struct {self.name}"

    @property
    def abbr(self) -> str:
        return f"st{self.name}"

    @property
    def nbytes(self) -> int:
        return ctypes.sizeof(self.ctype)

    def to_c(self, value):
        return self.ctype(*dataclasses.astuple(value))

    def from_c(self, cvalue):
        return self.dataclass(*[getattr(cvalue, n) for n, _ in self.elements])

    def __repr__(self) -> str:
        return self.name


@dataclass(eq=True, frozen=True)
class Reference(BaseType):
    pass


@dataclass(eq=True, frozen=True)
class Pointer(Reference):
    element: Value = None

    @property
    def ctype(self):
        return ctypes.POINTER(self.element.ctype)

    def __repr__(self) -> str:
        return f"Pointer of {self.element!r}"


@dataclass(eq=True, frozen=True)
class GridType(Reference):
    element: Value = None
    dimension: int = 1

    def __repr__(self) -> str:
        return f"Grid({self.dimension}) of {self.element!r}"


# ---------------------------------------------------------------------------
# annotations: xgrid.grid[T, D] and xgrid.ptr[T]
# ---------------------------------------------------------------------------
class _Annotation:
    __slots__ = ("origin", "args")

    def __init__(self, origin, args):
        self.origin = origin
        self.args = args

    def __repr__(self) -> str:
        return f"{self.origin.__name__}[{', '.join(map(repr, self.args))}]"

    def __eq__(self, other):
        return isinstance(other, _Annotation) and self.origin is other.origin and self.args == other.args

    def __hash__(self):
        return hash((self.origin, self.args))


class grid:
    """``grid[T, D]``: a D-dimensional grid of T (``annotation.py:25-28``)."""

    def __class_getitem__(cls, item):
        if not isinstance(item, tuple):
            item = (item,)
        return _Annotation(grid, item)


class ptr:
    """``ptr[T]``: a scalar passed by address (``annotation.py:21-22``)."""

    def __class_getitem__(cls, item):
        if not isinstance(item, tuple):
            item = (item,)
        return _Annotation(ptr, item)


def _resolve(annotation, glbs):
    if isinstance(annotation, str):
        if glbs is not None and annotation in glbs:
            return glbs[annotation]
        builtins = {"int": int, "float": float, "bool": bool, "None": None}
        if annotation in builtins:
            return builtins[annotation]
        raise DSLTypeError(f"unresolved annotation '{annotation}'")
    return annotation


def parse_annotation(annotation, glbs: dict | None = None) -> BaseType | None:
    """Map a Python annotation to a DSL type, or None when it has none.

    Same rules as ``xgrid/util/typing/annotation.py:31-68``.
    """
    annotation = _resolve(annotation, glbs)
    if annotation is None or annotation is type(None):
        return Void()
    if annotation is Any:
        return Ignore()
    if annotation is int:
        return Integer(struct.calcsize("i"))
    if annotation is float:
        return Floating(get_config().fsize)
    if annotation is bool:
        return Boolean()
    if is_dataclass(annotation) and isinstance(annotation, type):
        import sys
        mod_globals = vars(sys.modules.get(annotation.__module__, None) or object())
        hints = {}
        try:
            hints = typing.get_type_hints(annotation)
        except Exception:
            hints = {}
        elements = []
        for f in fields(annotation):
            t = parse_annotation(hints.get(f.name, f.type), mod_globals)
            if not isinstance(t, Value):
                return None
            elements.append((f.name, t))
        return Structure(annotation, annotation.__name__, tuple(elements))
    if isinstance(annotation, _Annotation):
        args = annotation.args
        if annotation.origin is ptr and len(args) == 1:
            v = parse_annotation(args[0], glbs)
            return Pointer(v) if isinstance(v, Value) else None
        if annotation.origin is grid and len(args) == 2 and type(args[1]) is int:
            v = parse_annotation(args[0], glbs)
            return GridType(v, args[1]) if isinstance(v, Value) else None
        return None
    return None


def numpy_dtype(t: Value):
    """numpy dtype of a grid element (``xgrid/xgrid/__init__.py:10-18``).

    Structures use the aligned (C) layout so host arrays and device code agree.
    """
    if isinstance(t, Boolean):
        return np.dtype(np.bool_)
    if isinstance(t, Integer):
        return np.dtype({1: np.int8, 2: np.int16, 4: np.int32, 8: np.int64}[t.width_bytes])
    if isinstance(t, Floating):
        return np.dtype(np.float32 if t.width_bytes == 4 else np.float64)
    if isinstance(t, Structure):
        return np.dtype([(n, numpy_dtype(e)) for n, e in t.elements], align=True)
    raise DSLTypeError(f"no numpy dtype for {t!r}")


def element_nbytes(t: Value) -> int:
    return numpy_dtype(t).itemsize
