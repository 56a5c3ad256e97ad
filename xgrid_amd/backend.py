"""Build and call a kernel on the MI355X backend.

``build_kernel(op)`` returns ``(native, depth)`` - the pair the reference's
``Generator(op).result`` returns (``xgrid/lang/generator.py:106-108``) - where
``native`` is a callable with the reference's calling convention
(``xgrid/util/ffi.py:30-34``: positional arguments, arity checked, scalars and
dataclasses converted to C, the C return value converted back).  Underneath:
hipRTC code object + host stub + ``xg_kernel`` launch plan.  There is no CPU
fallback: without the library or a GPU this raises.
"""
from __future__ import annotations

import ctypes
import weakref
from ctypes import c_void_p

from . import _lib
from .analysis import analyze
from .compilers import compile_device, compile_host, load_host
from .config import get_config
from .emit import Emitter, Lowered
from .grid import Grid
from .types import Boolean, Floating, GridType, Integer, Pointer, Structure, Void

_MODE = {"inplace": _lib.MODE_INPLACE, "pingpong": _lib.MODE_PINGPONG, "twophase": _lib.MODE_TWOPHASE}


class KernelError(RuntimeError):
    pass


def lower_kernel(op, block_shape: tuple | None = None) -> Lowered:
    cfg = get_config()
    info = analyze(op.ir)
    if not cfg.parallel and any(st.implicit for st in info.stmts):
        raise KernelError("Unable to perform implicit operation while parallel is off")
    return Emitter(info, cfg.overstep, cfg.comment, macro=op.macro, block_shape=block_shape).lower(), info


def _ctype_of(t):
    if isinstance(t, GridType):
        return c_void_p
    if isinstance(t, Pointer):
        return ctypes.POINTER(t.element.ctype)
    return t.ctype


def packed_args_type(arg_types):
    """ctypes mirror of the stub's ``struct xg_args`` (one field per kernel argument)."""
    fields_ = [(f"a{i}", _ctype_of(t)) for i, (_, t) in enumerate(arg_types)] or [("xg_unused", ctypes.c_int32)]
    return type("xg_args", (ctypes.Structure,), {"_fields_": fields_})


def _release(lib, plan, modules):
    """Destroy a kernel's own launch plan (its graph execs, events, flag) and unload
    its code objects (`modules`: a list the kernel appends the batched module to
    when it loads it).  Plans made by new_plan() for batch instances belong to their
    batch, which holds the kernel and destroys them first."""
    try:
        if plan:
            lib.xg_kernel_destroy(plan)
        for module in modules:
            if module:
                lib.xg_module_unload(module)
    except Exception:
        pass


class HipKernel:
    """A compiled kernel: code object, stub entry point and launch plan.

    close() (or the garbage collector, through a finalizer) destroys the plan and
    unloads the code object.  The finalizer does not run at interpreter exit: the
    process's end releases both, with no GPU calls during shutdown."""

    def __init__(self, op, block_shape: tuple | None = None):
        cfg = get_config()
        self.cfg = cfg                   # options are those in force when the kernel was built
        self.op = op
        self.lib = _lib.ensure_device(cfg.device)
        low, info = lower_kernel(op, block_shape)
        self.block_shape = tuple(block_shape) if block_shape is not None else low.block_shape
        self.lowered = low
        self.info = info
        self.depth = info.depth
        self.code_object = compile_device(low.device_src, cfg.cacheroot, cfg.fp_contract)
        self.stub_path = compile_host(low.stub_src, cfg.cacheroot, cfg.cc, cfg.opt_level)
        mod = c_void_p()
        _lib.check(self.lib.xg_module_load(self.code_object.encode(), ctypes.byref(mod)), "xg_module_load")
        self.module = mod.value
        self.module_b = None             # the batched variants: loaded by ensure_batch()
        self._modules = [self.module]
        self.kernel = None
        self._finalizer = weakref.finalize(self, _release, self.lib, None, self._modules)
        self._finalizer.atexit = False
        fn = self._function_of(self.module)

        grid_index = {n: k for k, (_, n, _) in enumerate(low.grid_args)}
        self.grid_index = grid_index
        descs = (_lib.StmtDesc * max(1, len(low.stmts)))()
        overstep = {"none": 0, "limit": 1, "wrap": 2}[cfg.overstep]
        self.param_bytes = []
        for i, m in enumerate(low.stmts):
            st = m.stmt
            d = descs[i]
            d.store_arg = grid_index[st.grid]
            d.klass = st.klass
            d.mode = _MODE[st.mode]
            d.dims = st.dimension
            d.elem_bytes = st.gtype.element.nbytes if hasattr(st.gtype.element, "nbytes") else 1
            d.nrefs = len(m.refs)
            for r, (g, lvl) in enumerate(m.refs):
                d.ref_arg[r] = grid_index[g]
                d.ref_level[r] = lvl
            so = list(st.store_offsets) + [0] * (3 - len(st.store_offsets))
            for a in range(3):
                d.store_off[a] = so[a]
            haz = st.hazard_offsets if st.mode == "inplace" else []
            if len(haz) > _lib.XG_MAX_HAZ:
                haz = []                       # too many to check per cell: always compute-then-store
                d.mode = _lib.MODE_TWOPHASE
            d.nhazard = len(haz)
            for h, off in enumerate(haz):
                off3 = list(off) + [0] * (3 - len(off))
                for a in range(3):
                    d.hazard[h][a] = off3[a]
            d.param_bytes = self._param_size(m)
            d.overstep = overstep
            d.group = m.group
            d.group_pos = m.group_pos
            d.fn_main = fn(m.kernel)
            d.fn_pingpong = fn(m.kernel_pp)
            d.fn_phase2 = fn(m.phase2)
            d.fn_pair = fn(m.pair) if m.pair else None
            d.pair_fwd = int(m.pair_fwd)
            d.tile3_cols = m.tile3_cols
        groups = (_lib.GroupDesc * max(1, len(low.groups)))()
        for i, g in enumerate(low.groups):
            groups[i].first = g.first
            groups[i].count = g.count
            groups[i].fn = fn(g.kernel)
        # loop plans: one per blockable loop (ids as in the stub), then the small-grid
        # tiles of those that have one
        alts = [L.alt for L in low.loops if L.alt is not None]
        loops = (_lib.LoopDesc * max(1, len(low.loops) + len(alts)))()
        nalt = 0
        for i, L in enumerate(list(low.loops) + alts):
            ld = loops[i]
            ld.nstmts = len(L.stmts)
            for j, m in enumerate(L.stmts):
                ld.sids[j] = m.stmt.sid
            ld.grid_arg = grid_index[L.grid]
            ld.core0, ld.core1 = L.core
            ld.halo = L.halo
            ld.tmax = L.tmax
            ld.threads = L.threads
            ld.fn = fn(L.kernel)
            ld.fn_check = fn(L.kernel + "_chk")
            ld.alt = -1
            if i < len(low.loops) and L.alt is not None:
                ld.alt = len(low.loops) + nalt
                ld.alt_max_cells = L.alt_max_cells
                nalt += 1
        self._nloops = len(low.loops) + len(alts)
        self._descs, self._groups, self._loops = descs, groups, loops
        self.kernel = self.new_plan()
        self._finalizer.detach()
        self._finalizer = weakref.finalize(self, _release, self.lib, self.kernel, self._modules)
        self._finalizer.atexit = False
        self.set_options()
        self.api = self.lib.xg_api_table()
        self.stub = load_host(self.stub_path)
        self.entry = self.stub.xg_entry
        self.arg_types = list(op.ir.args)
        self.entry.argtypes = [c_void_p, c_void_p] + [_ctype_of(t) for _, t in self.arg_types]
        ret = op.ir.ret
        self.ret = ret
        self.entry.restype = None if isinstance(ret, Void) else ret.ctype
        self.written = [grid_index[g] for g in info.written]
        self.entry_packed = ctypes.cast(self.stub.xg_entry_packed, c_void_p).value
        self.stub.xg_args_size.restype = ctypes.c_size_t
        self.Args = packed_args_type(self.arg_types)
        if ctypes.sizeof(self.Args) != self.stub.xg_args_size():
            raise KernelError("packed argument layout differs between the stub and ctypes")

    def _function_of(self, module):
        def fn(name):
            if name is None:
                return None
            out = c_void_p()
            _lib.check(self.lib.xg_module_function(module, name.encode(), ctypes.byref(out)), f"kernel {name}")
            return out.value
        return fn

    def ensure_batch(self) -> None:
        """Compile (or find in the cache) and load the batched variants of this kernel's
        launches, and enter them in the descriptors: plans made by new_plan() after this
        can run in a batch.  Plain calls never need them, so they are built only when a
        batch of this kernel is made."""
        if self.module_b is not None:
            return
        cfg = self.cfg
        path = compile_device(self.lowered.batch_src, cfg.cacheroot, cfg.fp_contract)
        mod = c_void_p()
        _lib.check(self.lib.xg_module_load(path.encode(), ctypes.byref(mod)), "xg_module_load")
        self.module_b = mod.value
        self._modules.append(self.module_b)
        fn = self._function_of(self.module_b)
        low = self.lowered
        for i, m in enumerate(low.stmts):
            d = self._descs[i]
            d.fn_main_b = fn(m.kernel + "_b")
            d.fn_pingpong_b = fn(m.kernel_pp + "b") if m.kernel_pp else None
            d.fn_phase2_b = fn(m.phase2 + "_b")
        for i, g in enumerate(low.groups):
            self._groups[i].fn_b = fn(g.kernel + "_b")
        for i, L in enumerate(list(low.loops) + [L.alt for L in low.loops if L.alt is not None]):
            self._loops[i].fn_b = fn(L.kernel + "_b") if L.batched else None
            self._loops[i].fn_check_b = fn(L.kernel + "_chkb") if L.batched else None

    def close(self) -> None:
        """Destroy the launch plan and unload the code objects now.  Batch plans made
        with new_plan() must have been destroyed already (Batch.close)."""
        self._finalizer()
        self.kernel = None
        self.module = None
        self.module_b = None

    def new_plan(self) -> int:
        """A fresh launch plan (xg_kernel) over this kernel's code: one per batch instance."""
        low = self.lowered
        k = c_void_p()
        _lib.check(self.lib.xg_kernel_create(len(low.stmts), self._descs, len(low.groups), self._groups,
                                             self._nloops, self._loops, len(low.grid_args), ctypes.byref(k)),
                   "xg_kernel_create")
        return k.value

    def _param_size(self, m) -> int:
        fields_ = [(v.name, (v.type.element if isinstance(v.type, Pointer) else v.type).ctype) for v in m.params]
        fields_ += [(h.name, ctypes.c_double if h.ctype == "double" else ctypes.c_float) for h in m.hoisted]
        if not fields_:
            fields_ = [("xg_unused", ctypes.c_int32)]
        return ctypes.sizeof(type("P", (ctypes.Structure,), {"_fields_": fields_}))

    def set_options(self, graphs: bool | None = None, fuse: bool | None = None, blocked: bool | None = None,
                    plan: int | None = None):
        cfg = self.cfg
        g = cfg.graphs and self.lowered.graphs_ok if graphs is None else graphs
        f = cfg.fuse if fuse is None else fuse
        b = cfg.blocked if blocked is None else blocked
        _lib.check(self.lib.xg_kernel_set_options(self.kernel if plan is None else plan, int(g), int(f), int(b)))

    def loop_plan(self, lid: int, cells: int) -> int:
        """The loop plan plain calls use for loop `lid` on a grid of `cells` cells
        (the loop's own, or its small-grid tile's)."""
        low = self.lowered
        L = low.loops[lid]
        if L.alt is None or cells > L.alt_max_cells:
            return lid
        return len(low.loops) + sum(1 for M in low.loops[:lid] if M.alt is not None)

    def loop_info(self, lid: int = 0) -> dict:
        n, it, st = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
        _lib.check(self.lib.xg_kernel_loop_info(self.kernel, lid, ctypes.byref(n), ctypes.byref(it), ctypes.byref(st)))
        return {"launches": n.value, "iterations": it.value, "state": st.value}

    def time_loop(self, lid: int, trip: int, reps: int = 50) -> float:
        """ms per launch of loop `lid`'s blocked kernel, running `trip` iterations the
        way a call does (chunks of at most tmax), back to back."""
        ms = ctypes.c_double()
        _lib.check(self.lib.xg_kernel_time_loop(self.kernel, lid, trip, reps, ctypes.byref(ms)))
        return ms.value

    def counters(self) -> dict:
        vals = [ctypes.c_int64() for _ in range(4)]
        _lib.check(self.lib.xg_kernel_counters(self.kernel, *[ctypes.byref(v) for v in vals]))
        return dict(zip(("launches", "fused", "graph_replays", "graph_captures"), (v.value for v in vals)))

    def statement_counts(self, reset: bool = False) -> list:
        n = len(self.lowered.stmts)
        arr = (ctypes.c_int64 * max(1, n))()
        _lib.check(self.lib.xg_kernel_stmt_counts(self.kernel, arr, n, int(reset)))
        return list(arr)[:n]

    def algorithmic_bytes(self, counts: list, grids: list) -> int:
        """Sum over executed statements of cells(class) x element bytes x (arrays loaded + 1)."""
        total = 0
        for m, c in zip(self.lowered.stmts, counts):
            if not c:
                continue
            st = m.stmt
            g = grids[self.grid_index[st.grid]]
            cells = g.class_info(st.klass)["count"]
            loaded = len([r for r in m.refs if r[1] >= 0])
            total += c * cells * g.numpy_dtype.itemsize * (loaded + 1)
        return total

    def time_isolated(self, sid: int, reps: int = 100) -> float:
        """ms per launch of statement `sid`, launched back to back with its last parameters."""
        ms = ctypes.c_double()
        _lib.check(self.lib.xg_kernel_time_isolated(self.kernel, sid, reps, ctypes.byref(ms)))
        return ms.value

    def time_statement(self, sid: int, enable: bool = True):
        _lib.check(self.lib.xg_kernel_time_stmt(self.kernel, sid, int(enable)))

    def statement_time(self, sid: int):
        ms, n = ctypes.c_double(), ctypes.c_int64()
        _lib.check(self.lib.xg_kernel_stmt_time(self.kernel, sid, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def convert_args(self, args) -> tuple:
        """Check arity and types (``xgrid/util/ffi.py:30-34``); returns (C values, Grid args in order)."""
        if len(args) != len(self.arg_types):
            raise TypeError(f"this function takes {len(self.arg_types)} argument ({len(args)} given)")
        cargs = []
        grids = []
        for (name, t), a in zip(self.arg_types, args):
            if isinstance(t, GridType):
                if not isinstance(a, Grid):
                    raise TypeError(f"argument '{name}' expects a Grid")
                if a._typed_for is not t:         # a grid's element and rank never change
                    if a.dimension != t.dimension or a.element != t.element:
                        raise TypeError(f"argument '{name}' expects {t!r}, got a grid of {a.element!r} with "
                                        f"{a.dimension} dimension(s)")
                    a._typed_for = t
                cargs.append(a._sync_for_launch())
                grids.append(a)
            elif isinstance(t, Structure):
                cargs.append(t.to_c(a))
            elif isinstance(t, Pointer):
                if isinstance(a, ctypes._Pointer):
                    cargs.append(a)
                else:
                    cargs.append(ctypes.pointer(a))
            else:
                cargs.append(a)
        return cargs, grids

    def __call__(self, *args):
        cargs, grids = self.convert_args(args)
        result = self.entry(self.kernel, self.api, *cargs)
        rc = self.lib.xg_kernel_status(self.kernel)
        for i in self.written:
            grids[i]._device_wrote()
        if rc != 0:
            raise _lib.BackendError(f"kernel '{self.op.name}' failed: {_lib.last_error()} (code {rc})")
        if isinstance(self.ret, Structure):
            return self.ret.from_c(result)
        return result


def build_kernel(op, block_shape: tuple | None = None):
    k = HipKernel(op, block_shape)
    return k, k.depth
