"""``@kernel``, ``@function`` and ``@external`` operators.

Same public contract as ``xgrid/lang/operator.py:13-88``: calling a kernel
builds it on first use, then extends and ticks every ``Grid`` argument to the
kernel's time depth (``operator.py:36-39``) and runs it; a ``@function`` called
from Python runs the Python function; an external cannot be called from
Python.  The build produces a device kernel (``backend.build_kernel``) in place
of the reference's JIT-compiled C (``Generator(self).result``).
"""
from __future__ import annotations

import dataclasses
import re
import threading
from typing import Any, Callable

from .types import BaseType

# A background compile parses every operator it reaches into a table of its own
# (this thread-local one), never into Operator._ir: what the first call parses is
# decided then, under the configuration and globals in force at that time.
_private = threading.local()
_Grid = None                     # xgrid_amd.grid.Grid, imported on the first kernel call
_ir_lock = threading.RLock()     # a parse parses the operators it calls

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
        self._precompiled = None
        self.__doc__ = getattr(func, "__doc__", None)

    def __call__(self, *args: Any) -> Any:
        if self.mode == "kernel":
            global _Grid
            if self.native is None:
                from .backend import build_kernel
                if self._precompiled is not None:
                    self._precompiled.wait()      # its code objects are then in the cache
                self.native, self.depth = build_kernel(self)
            if _Grid is None:
                from .grid import Grid as _Grid
            depth, tick = self.depth, self.tick
            for arg in args:
                if isinstance(arg, _Grid):
                    arg._op_invoke(depth, tick)
            return self.native(*args)
        if self.mode == "function":
            return self.func(*args)
        raise RuntimeError(f"Invalid call to non-kernel or non-function ({self.mode}) operator '{self.name}'")

    @property
    def ir(self):
        from .frontend import parse
        table = getattr(_private, "table", None)
        if table is not None:                    # a background compile: its own parse
            f = table.get(self)
            if f is None:
                f = table[self] = parse(self.func, self.name, self.mode, self.self_type)
            return f
        if self._ir is None:
            with _ir_lock:
                if self._ir is None:
                    f = parse(self.func, self.name, self.mode, self.self_type)
                    for inc in f.includes:
                        if inc not in self.includes:
                            self.includes.append(inc)
                    self._ir = f
        return self._ir

    def precompile(self, wait: bool = False) -> "Precompile":
        """Compile this kernel's device code and host stub now, in a background
        thread, under the configuration in force now (no GPU is touched).  The first
        call waits for it and then finds both in the on-disk cache; if the call's own
        parse produces different code (the configuration or a folded global changed
        in between), it compiles that instead.  Returns the handle (waited on when
        `wait`)."""
        if self.mode != "kernel":
            raise RuntimeError(f"only kernels are compiled; '{self.name}' is a {self.mode}")
        if self._precompiled is None or (self._precompiled.done() and self._precompiled.error is not None):
            self._precompiled = Precompile(self)
        if wait:
            self._precompiled.wait()
        return self._precompiled

    @property
    def c_name(self) -> str:
        if self.self_type is not None:
            return _c_ident(f"{self.self_type.name}_{self.name}")
        return _c_ident(self.name)

    @property
    def src(self) -> str:
        """Generated device source (plain launches, then the batched variants) and host
        stub of a kernel."""
        from .backend import lower_kernel
        low, _ = lower_kernel(self)
        return (low.device_src + "\n/* ---- batched variants ---- */\n" + low.batch_src +
                "\n/* ---- host stub ---- */\n" + low.stub_src)

    def signature_args(self) -> list:
        return list(self.ir.args)

    def signature_ret(self):
        return self.ir.ret

    @property
    def signature(self):
        return self.ir.args, self.ir.ret

    def __repr__(self) -> str:
        return f"<xgrid {self.mode} '{self.name}'>"


class Precompile:
    """A kernel's device code and host stub compiled by a background (daemon)
    thread into the on-disk cache.  Errors are kept, not raised: the first call
    builds the kernel itself and reports them there."""

    def __init__(self, op: Operator):
        from .config import get_config
        self.op = op
        self.cfg = dataclasses.replace(get_config())
        self.error = None
        self.seconds = None
        self._done = threading.Event()
        threading.Thread(target=self._run, name=f"xgrid-precompile-{op.name}", daemon=True).start()

    def _run(self):
        import time
        from .backend import lower_kernel
        from .compilers import compile_device, compile_host
        from .config import thread_config
        t0 = time.perf_counter()
        try:
            _private.table = {}
            with thread_config(self.cfg):
                low, _ = lower_kernel(self.op)
                compile_device(low.device_src, self.cfg.cacheroot, self.cfg.fp_contract)
                compile_host(low.stub_src, self.cfg.cacheroot, self.cfg.cc, self.cfg.opt_level)
        except Exception as e:          # reported by the first call's own build
            self.error = e
        finally:
            _private.table = None
            self.seconds = time.perf_counter() - t0
            self._done.set()

    def done(self) -> bool:
        return self._done.is_set()

    def wait(self, timeout: float | None = None) -> bool:
        return self._done.wait(timeout)


def precompile(*ops, wait: bool = False) -> list:
    """Start background compiles of several kernels; returns their handles."""
    handles = [op.precompile() for op in ops]
    if wait:
        for h in handles:
            h.wait()
    return handles


def kernel(*, name: str | None = None, includes: list | None = None, tick: bool = True,
           macro: list | None = None):
    """The kernel decorator.  When the configuration is already initialised (and its
    `precompile` option is on), the kernel's compile starts in the background at
    once, so that its first call does not wait for hipRTC (xgrid/util/ffi.py:59-94
    compiles on the first call)."""
    def aux(func):
        from .config import config_ready, get_config
        op = Operator(func, "kernel", name, includes, tick=tick, macro=macro)
        if config_ready() and get_config().precompile:
            op.precompile()
        return op
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
