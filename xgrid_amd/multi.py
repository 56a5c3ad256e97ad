"""Many independent cavity and 2-D diffusion solves on one GPU.

``MultiSolver`` keeps one :class:`~xgrid_amd.batch.Batch` per kernel (the
lid-driven cavity and 2-D diffusion) in one precision, adds and removes
instances described by :class:`~xgrid_amd.workloads.InstanceSpec`, and
advances them together.  ``plain_run`` runs one instance the drop-in way
(plain kernel calls), which batched results must equal bit for bit.

No counterpart in the reference (one kernel call at a time).
"""
from __future__ import annotations

from . import workloads as W
from .backend import HipKernel, build_kernel
from .batch import Batch
from .grid import Grid
from .precision import precision_scope


class MultiSolver:
    KINDS = ("cavity", "diffusion")

    def __init__(self, precision: str = "double", graphs: bool = True, max_blocks: int = 0):
        self.precision = precision
        self.ops = {}
        with precision_scope(precision):
            for kind, make in (("cavity", W.make_cavity), ("diffusion", W.make_diffusion_2d)):
                op = make()
                op.native, op.depth = build_kernel(op)
                self.ops[kind] = op
        self.graphs, self.max_blocks = graphs, max_blocks
        self.batches = {k: Batch(op, graphs=graphs, max_blocks=max_blocks) for k, op in self.ops.items()}
        self.members = {}            # handle -> (kind, member id, grids, spec)
        self._next = 1
        self.shapes = {k: op.native.block_shape for k, op in self.ops.items()}
        self.tuning = {}

    def autotune(self, specs: list, limit: int = 48, use_cache: bool = True) -> dict:
        """Pick each kernel's blocked-loop tile shape on fresh copies of a sample of
        `specs` (before any instance is added) and rebuild its batch with it."""
        from . import autotune
        if self.members:
            raise RuntimeError("autotune before adding instances")
        for kind, op in self.ops.items():
            mine = [s for s in specs if s.kind == kind]
            if not mine:
                continue
            with precision_scope(self.precision):
                if not autotune.has_blocked_loop(op):
                    continue
                res = autotune.tune_batch(op, lambda s: W.instance_grids(s, Grid),
                                          autotune.sample(mine, limit, size=lambda s: s.n), self.precision,
                                          max_blocks=self.max_blocks, use_cache=use_cache)
                native = HipKernel(op, res["shape"])
            self.batches[kind].close()
            self.batches[kind] = Batch(op, graphs=self.graphs, max_blocks=self.max_blocks, native=native)
            self.shapes[kind] = res["shape"]
            self.tuning[kind] = res
        return self.tuning

    def grids_for(self, spec: W.InstanceSpec):
        with precision_scope(self.precision):
            return W.instance_grids(spec, Grid)

    def add(self, spec: W.InstanceSpec) -> int:
        grids, scalars = self.grids_for(spec)
        mid = self.batches[spec.kind].add(*grids, *scalars)
        h = self._next
        self._next += 1
        self.members[h] = (spec.kind, mid, grids, spec)
        return h

    def remove(self, handle: int) -> list:
        """Drop an instance; returns its grids (still valid, holding its fields)."""
        kind, mid, grids, _ = self.members.pop(handle)
        self.batches[kind].remove(mid)
        return grids

    def grids(self, handle: int) -> list:
        return self.members[handle][2]

    def step(self, steps: int = 1) -> None:
        for b in self.batches.values():
            b.step(steps)

    def set_options(self, graphs: bool = True, max_blocks: int = 0) -> None:
        for b in self.batches.values():
            b.set_options(graphs, max_blocks)

    def counters(self) -> dict:
        return {k: b.counters() for k, b in self.batches.items()}

    def plain_call(self, spec: W.InstanceSpec, grids, scalars) -> None:
        """One drop-in kernel call of an instance (its own grids, plain launches)."""
        self.ops[spec.kind](*grids, *scalars)

    def close(self) -> None:
        for b in self.batches.values():
            b.close()
        self.members.clear()
