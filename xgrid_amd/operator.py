"""``@kernel``, ``@function`` and ``@external`` operators.

Same public contract as ``xgrid/lang/operator.py:13-88``: calling a kernel
builds it on first use, then extends and ticks every ``Grid`` argument to the
kernel's time depth (``operator.py:36-39``) and runs it; a ``@function`` called
from Python runs the Python function; an external cannot be called from
Python.  The build produces a device kernel (``backend.build_kernel``) in place
of the reference's JIT-compiled C (``Generator(self).result``).
"""
from __future__ import annotations

import re
from typing import Any, Callable

from .types import BaseType

CustomTypecheck = Callable[[list], BaseType]


def _c_ident(text: str) -> str:
    s = re.sub(r"[^0-9A-Za-z_]", "_", text)
    return s if s and not s[0].isdigit() else "_" + s


class Operator:
    def __init__(self, func, mode: str, name: str | None = None, includes: list | None = None,
                 self_type: BaseType | None = None, typecheck_override: CustomTypecheck | None = None,
                 tick: bool = True, macro: list | None = None, intrinsic: str | None = None):
        self.func = func
        self.mode = mode
        self.name = func.__name__ if name is None else name
        self.includes = [] if includes is None else list(includes)
        self.macro = [] if macro is None else list(macro)
        self.native = None
        self.depth = None
        self.self_type = self_type
        self.typecheck_override = typecheck_override
        self.tick = tick
        self.intrinsic = intrinsic
        self._ir = None
        self.__doc__ = getattr(func, "__doc__", None)

    def __call__(self, *args: Any) -> Any:
        if self.mode == "kernel":
            if self.native is None:
                from .backend import build_kernel
                self.native, self.depth = build_kernel(self)
            from .grid import Grid
            for arg in args:
                if isinstance(arg, Grid):
                    arg._op_invoke(self.depth, self.tick)
            return self.native(*args)
        if self.mode == "function":
            return self.func(*args)
        raise RuntimeError(f"Invalid call to non-kernel or non-function ({self.mode}) operator '{self.name}'")

    @property
    def ir(self):
        if self._ir is None:
            from .frontend import parse
            self._ir = parse(self.func, self.name, self.mode, self.self_type)
            for inc in self._ir.includes:
                if inc not in self.includes:
                    self.includes.append(inc)
        return self._ir

    @property
    def c_name(self) -> str:
        if self.self_type is not None:
            return _c_ident(f"{self.self_type.name}_{self.name}")
        return _c_ident(self.name)

    @property
    def src(self) -> str:
        """Generated device source and host stub of a kernel."""
        from .backend import lower_kernel
        low, _ = lower_kernel(self)
        return low.device_src + "\n/* ---- host stub ---- */\n" + low.stub_src

    def signature_args(self) -> list:
        return list(self.ir.args)

    def signature_ret(self):
        return self.ir.ret

    @property
    def signature(self):
        return self.ir.args, self.ir.ret

    def __repr__(self) -> str:
        return f"<xgrid {self.mode} '{self.name}'>"


def kernel(*, name: str | None = None, includes: list | None = None, tick: bool = True,
           macro: list | None = None):
    def aux(func):
        return Operator(func, "kernel", name, includes, tick=tick, macro=macro)
    return aux


def function(*, method: bool = False, name: str | None = None, includes: list | None = None,
             macro: list | None = None):
    if method:
        def aux_method(func):
            setattr(func, "__xgrid_method", (name, includes, macro))
            return func
        return aux_method

    def aux(func):
        return Operator(func, "function", name, includes, macro=macro)
    return aux


def external(*, name: str | None = None, includes: list | None = None,
             typecheck_override: CustomTypecheck):
    def aux(func):
        return Operator(func, "external", name, includes, typecheck_override=typecheck_override)
    return aux


def method_operator(func, self_type) -> Operator:
    """The Operator of a ``@function(method=True)`` bound to one structure type."""
    cache = func.__dict__.setdefault("__xgrid_method_ops", {})
    op = cache.get(self_type)
    if op is None:
        name, includes, macro = getattr(func, "__xgrid_method")
        op = Operator(func, "function", name, includes, self_type=self_type, macro=macro)
        cache[self_type] = op
    return op
