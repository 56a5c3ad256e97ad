"""Stencil analysis of a parsed kernel.

For every assignment whose target is a grid cell (a "stencil statement") this
records what the reference's ``StencilParser`` records
(``xgrid/lang/generator.py:50-76``) - the store grid, its boundary class and
whether the statement is *implicit* (it reads its own store grid at time 0
outside a boundary block, which the reference runs as a two-phase Jacobi nest,
``generator.py:312-352``) - plus what the device backend needs: the distinct
(grid, time level) arrays the statement loads, the scalar variables its
expression captures, and the kernel's time depth (``generator.py:428``).
"""
from __future__ import annotations

from dataclasses import dataclass, field

from . import ir
from .types import GridType, Pointer, Structure


class StencilError(RuntimeError):
    pass


@dataclass
class StencilStmt:
    sid: int
    node: ir.Assign
    grid: str
    gtype: GridType
    klass: int
    store_offsets: tuple
    implicit: bool
    self_offsets: list            # offsets of time-0 loads of the store grid
    refs: list                    # distinct (grid name, level) loaded, first-seen order
    scalars: list                 # captured scalar Vars (first-seen order)
    info_grids: list              # grids whose shape/dimension the expression reads
    calls: list = field(default_factory=list)   # user operators called by the expression

    @property
    def dimension(self) -> int:
        return self.gtype.dimension

    @property
    def reads_self_now(self) -> bool:
        return bool(self.self_offsets)

    @property
    def mode(self) -> str:
        """How the device executes the statement.

        ``inplace``: no cell reads a cell another cell of the statement writes.
        ``pingpong``: Jacobi with an unshifted store; computed into the grid's
        scratch buffer, cells outside the class copied, buffers swapped.
        ``twophase``: the reference's compute-then-store (temp per cell).
        """
        if self.implicit:
            if all(all(o == s for o, s in zip(off, self.store_offsets)) for off in self.self_offsets):
                return "inplace"      # only the cell itself is read: in place is identical
            if any(self.store_offsets):
                return "twophase"
            return "pingpong"
        return "inplace"

    @property
    def hazard_offsets(self) -> list:
        """Relative offsets (read - store) a same-class cell could race on."""
        out = []
        for off in self.self_offsets:
            d = tuple(o - s for o, s in zip(off, self.store_offsets))
            if any(d) and d not in out:
                out.append(d)
        return out


@dataclass
class KernelInfo:
    func: ir.Function
    stmts: list
    depth: int                     # number of time levels every grid arg gets
    written: list                  # grid args written by some statement
    grid_args: list                # [(index, name, GridType)]


def _captures(expr, scalars: list, info: list, calls: list, stencil_ok: bool):
    for n in ir.walk(expr):
        if isinstance(n, ir.Name):
            v = n.var
            if isinstance(v.type, GridType):
                raise StencilError(f"{n.loc}: grid '{v.name}' used as a value")
            if all(v is not s for s in scalars):
                scalars.append(v)
        elif isinstance(n, ir.GridInfo):
            if n.grid.name not in info:
                info.append(n.grid.name)
        elif isinstance(n, ir.Call) and n.callee is not None:
            if n.callee not in calls:
                calls.append(n.callee)
        elif isinstance(n, ir.Cell) and not stencil_ok:
            raise StencilError(
                f"{n.loc}: Unable to perform load operation to grid '{n.grid.name}' without stencil context")


def analyze(func: ir.Function) -> KernelInfo:
    stmts: list[StencilStmt] = []
    depth = 0

    def visit_block(block):
        nonlocal depth
        for s in block:
            if isinstance(s, ir.Assign) and isinstance(s.target, ir.Cell) and s.target.ctx == "store":
                tgt = s.target
                depth = max(depth, tgt.level)
                refs, self_offs = [], []
                for n in ir.walk(s.value):
                    if isinstance(n, ir.Cell):
                        if n.ctx != "load":
                            raise StencilError(f"{n.loc}: nested grid store")
                        depth = max(depth, n.level)
                        key = (n.grid.name, n.level)
                        if key not in refs:
                            refs.append(key)
                        if n.grid.name == tgt.grid.name and n.time == 0:
                            self_offs.append(n.offsets)
                implicit = any(n.klass == 0 for n in ir.walk(s.value)
                               if isinstance(n, ir.Cell) and n.grid.name == tgt.grid.name and n.time == 0)
                scalars, info, calls = [], [], []
                _captures(s.value, scalars, info, calls, stencil_ok=True)
                stmts.append(StencilStmt(len(stmts), s, tgt.grid.name, tgt.grid.type, tgt.klass,
                                         tgt.offsets, implicit, self_offs, refs, scalars, info, calls))
                continue
            # any other statement: grid cells may not appear
            if isinstance(s, ir.Assign):
                _captures(s.target, [], [], [], stencil_ok=False)
                _captures(s.value, [], [], [], stencil_ok=False)
            elif isinstance(s, (ir.Eval, ir.Return)):
                if getattr(s, "value", None) is not None:
                    _captures(s.value, [], [], [], stencil_ok=False)
            elif isinstance(s, ir.If):
                _captures(s.cond, [], [], [], stencil_ok=False)
                visit_block(s.body)
                visit_block(s.orelse)
            elif isinstance(s, ir.While):
                _captures(s.cond, [], [], [], stencil_ok=False)
                visit_block(s.body)
            elif isinstance(s, ir.For):
                for e in (s.start, s.stop, s.step):
                    _captures(e, [], [], [], stencil_ok=False)
                visit_block(s.body)

    visit_block(func.body)
    written = []
    for st in stmts:
        if st.grid not in written:
            written.append(st.grid)
    return KernelInfo(func, stmts, depth + 1, written, func.grid_args)


def captured_value_type(v: ir.Var):
    """Type a captured scalar has inside device code (pointers are dereferenced)."""
    return v.type.element if isinstance(v.type, Pointer) else v.type


def structures_in(func: ir.Function) -> list:
    """Every Structure type reachable from a function's arguments and locals."""
    out: list = []

    def add(t):
        if isinstance(t, Pointer):
            t = t.element
        if isinstance(t, GridType):
            t = t.element
        if isinstance(t, Structure):
            for _, et in t.elements:
                add(et)
            if t not in out:
                out.append(t)

    for v in func.scope.values():
        add(v.type)
    add(func.ret)
    for n in ir.walk_all(func.body):
        if isinstance(n, ir.Expr):
            add(n.type)
    return out
