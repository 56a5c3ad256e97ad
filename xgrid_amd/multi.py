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
    """Cavity and 2-D diffusion instances in batches, one precision.  Instances of
    a kernel with a blocked loop may be split by grid side into size classes
    (`size_classes` are the class boundaries), each class a batch of its own with
    its own tuned tile shape.  On the mixed 64-1024 batch the small-grid classes
    pick the 40 x 40 tile and the large ones the default, but the extra launches
    cost more than the better tiles gain (FP64 5.20-5.31 ms per step against
    5.13 for one batch; FP32 4.40-4.47 against 4.29), so by default there is one
    batch per kernel."""

    KINDS = ("cavity", "diffusion")

    def __init__(self, precision: str = "double", graphs: bool = True, max_blocks: int = 0,
                 size_classes: tuple = ()):
        from . import autotune
        self.precision = precision
        self.ops = {}
        self.blocked = {}
        with precision_scope(precision):
            for kind, make in (("cavity", W.make_cavity), ("diffusion", W.make_diffusion_2d)):
                op = make()
                op.native, op.depth = build_kernel(op)
                self.ops[kind] = op
                self.blocked[kind] = bool(op.native.lowered.loops)
        self.graphs, self.max_blocks = graphs, max_blocks
        self.size_classes = tuple(size_classes)
        self.natives = {}            # batch key -> the build its batch runs
        self.batches = {}            # batch key -> Batch (created on first use)
        self.members = {}            # handle -> (batch key, member id, grids, spec)
        self._next = 1
        self.tuning = {}

    def key(self, spec: W.InstanceSpec) -> str:
        if not self.blocked[spec.kind] or not self.size_classes:
            return spec.kind
        c = sum(1 for b in self.size_classes if spec.n >= b)
        return f"{spec.kind}/{c}"

    @property
    def shapes(self) -> dict:
        return {k: b.native.block_shape for k, b in self.batches.items()}

    def _batch(self, key: str) -> Batch:
        b = self.batches.get(key)
        if b is None:
            op = self.ops[key.split("/")[0]]
            b = Batch(op, graphs=self.graphs, max_blocks=self.max_blocks, native=self.natives.get(key))
            self.batches[key] = b
        return b

    def autotune(self, specs: list, limit: int = 48, use_cache: bool = True) -> dict:
        """Pick the blocked-loop tile shape of each batch (kernel and size class) on
        fresh copies of a sample of its `specs`, before any instance is added."""
        from . import autotune
        if self.members:
            raise RuntimeError("autotune before adding instances")
        groups = {}
        for s in specs:
            groups.setdefault(self.key(s), []).append(s)
        for key, mine in groups.items():
            kind = key.split("/")[0]
            if not self.blocked[kind]:
                continue
            op = self.ops[kind]
            with precision_scope(self.precision):
                res = autotune.tune_batch(op, lambda s: W.instance_grids(s, Grid),
                                          autotune.sample(mine, limit, size=lambda s: s.n), self.precision,
                                          max_blocks=self.max_blocks, use_cache=use_cache)
                self.natives[key] = HipKernel(op, res["shape"])
            if key in self.batches:
                self.batches.pop(key).close()
            self.tuning[key] = res
        return self.tuning

    def grids_for(self, spec: W.InstanceSpec):
        with precision_scope(self.precision):
            return W.instance_grids(spec, Grid)

    def add(self, spec: W.InstanceSpec) -> int:
        grids, scalars = self.grids_for(spec)
        key = self.key(spec)
        mid = self._batch(key).add(*grids, *scalars)
        h = self._next
        self._next += 1
        self.members[h] = (key, mid, grids, spec)
        return h

    def remove(self, handle: int) -> list:
        """Drop an instance; returns its grids (still valid, holding its fields)."""
        key, mid, grids, _ = self.members.pop(handle)
        self.batches[key].remove(mid)
        return grids

    def grids(self, handle: int) -> list:
        return self.members[handle][2]

    def step(self, steps: int = 1) -> None:
        for b in self.batches.values():
            if len(b):
                b.step(steps)

    def set_options(self, graphs: bool = True, max_blocks: int = 0) -> None:
        self.graphs, self.max_blocks = graphs, max_blocks
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
        self.batches.clear()
        self.members.clear()
