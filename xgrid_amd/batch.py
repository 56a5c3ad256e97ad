"""Batched execution of many independent instances of one ``@xgrid.kernel``.

No counterpart in the reference, which runs one kernel call at a time
(``xgrid/lang/operator.py:29-44``).  A :class:`Batch` holds instances of one
kernel - each with its own grids (of any size) and its own scalar and
structure arguments - and advances all of them by one kernel call per step:

* every instance's grids are extended to the kernel's depth when added and
  ticked before each call, exactly as ``Operator.__call__`` does;
* the runtime runs each instance's host stub with its statements recorded,
  then launches each statement once for all instances (a device table maps
  every workgroup tile to its instance), chains small boundary statements per
  instance in one workgroup, and runs blockable loops temporally blocked for
  all instances in one launch per chunk;
* repeated steps replay as one hipGraph.

After ``step(n)`` every instance's grids hold exactly what ``n`` plain calls
``kernel(*args)`` would have left (the device code per cell is the same).
Instances may be added and removed between steps (the serving loop does).
"""
from __future__ import annotations

import ctypes

from . import _lib
from .backend import build_kernel
from .grid import Grid


class BatchError(RuntimeError):
    pass


class _Member:
    __slots__ = ("id", "plan", "grids", "args", "packed")

    def __init__(self, id_, plan, grids, args, packed):
        self.id = id_
        self.plan = plan
        self.grids = grids
        self.args = args
        self.packed = packed


class Batch:
    """Instances of one kernel advanced together; see the module docstring."""

    def __init__(self, op, graphs: bool = True, max_blocks: int = 0, native=None):
        """`native`: a build of `op` to run (e.g. with a tuned tile shape); default op's own."""
        if op.mode != "kernel":
            raise BatchError(f"'{op.name}' is not a kernel")
        if native is None:
            if op.native is None:
                op.native, op.depth = build_kernel(op)
            native = op.native
        self.op = op
        self.native = native
        native.ensure_batch()            # the batched kernels, before any instance's plan is made
        self.depth = native.depth
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self.lib.xg_batch_create(self.native.entry_packed, int(bool(op.tick)), ctypes.byref(h)),
                   "xg_batch_create")
        self.handle = h.value
        self.members: dict = {}
        self.set_options(graphs, max_blocks)

    # ------------------------------------------------------------ membership
    def add(self, *args) -> int:
        """Add an instance with the kernel's arguments; returns its id."""
        cargs, grids = self.native.convert_args(args)
        owned = {id(g) for m in self.members.values() for g in m.grids}
        for g in grids:
            if id(g) in owned:
                raise BatchError("a grid may belong to one instance of a batch only")
        depth = self.depth
        handles = []
        for g in grids:
            g._extend_time(depth)
            handles.append(g._sync_for_launch())
        # grid arguments were converted before the depth change: convert again
        cargs, grids = self.native.convert_args(args)
        packed = self.native.Args(*cargs)
        plan = self.native.new_plan()
        self.native.set_options(plan=plan)
        ticks = (ctypes.c_void_p * max(1, len(handles)))(*handles)
        mid = ctypes.c_int64()
        rc = self.lib.xg_batch_add(self.handle, plan, ctypes.byref(packed), ctypes.sizeof(packed), ticks,
                                   len(handles), ctypes.byref(mid))
        if rc != 0:
            self.lib.xg_kernel_destroy(plan)
            _lib.check(rc, "xg_batch_add")
        self.members[mid.value] = _Member(mid.value, plan, grids, args, packed)
        return mid.value

    def remove(self, member: int) -> None:
        m = self.members.pop(member)
        _lib.check(self.lib.xg_batch_remove(self.handle, member), "xg_batch_remove")
        _lib.check(self.lib.xg_kernel_destroy(m.plan), "xg_kernel_destroy")

    def __len__(self) -> int:
        return len(self.members)

    def grids(self, member: int) -> list:
        return list(self.members[member].grids)

    # ------------------------------------------------------------ running
    def set_options(self, graphs: bool = True, max_blocks: int = 0) -> None:
        _lib.check(self.lib.xg_batch_set_options(self.handle, int(bool(graphs) and self.native.lowered.graphs_ok),
                                                 int(max_blocks)), "xg_batch_set_options")

    def step(self, steps: int = 1) -> None:
        """Advance every instance by `steps` kernel calls (asynchronous on the device)."""
        if steps <= 0 or not self.members:
            return
        depth = self.depth
        for m in self.members.values():
            for g in m.grids:
                if g.time_levels != depth:
                    raise BatchError("a grid's time levels changed outside the batch")
                g._sync_for_launch()
        _lib.check(self.lib.xg_batch_step(self.handle, int(steps)), f"batch of '{self.op.name}'")
        for m in self.members.values():
            for g in m.grids:
                if self.op.tick and len(g._levels) > 1:
                    r = steps % len(g._levels)
                    for _ in range(r):
                        g._levels.insert(0, g._levels.pop())
                for lvl in g._levels:
                    lvl.host_valid = False
                    lvl.dev_valid = True

    def synchronize(self) -> None:
        _lib.check(self.lib.xg_synchronize(), "xg_synchronize")

    def counters(self) -> dict:
        vals = [ctypes.c_int64() for _ in range(5)]
        _lib.check(self.lib.xg_batch_counters(self.handle, *[ctypes.byref(v) for v in vals]))
        return dict(zip(("launches", "graph_replays", "graph_captures", "steps", "validations"),
                        (v.value for v in vals)))

    def close(self) -> None:
        if self.handle is None:
            return
        lib = self.lib
        lib.xg_batch_destroy(self.handle)
        for m in self.members.values():
            lib.xg_kernel_destroy(m.plan)
        self.members.clear()
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
