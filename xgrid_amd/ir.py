"""Typed intermediate representation produced by the frontend.

Plays the role of the reference's ``xgrid/lang/ir/{statement,expression}.py``
(statements ``Assignment/For/If/While/Return/Inline``, expressions
``Binary/Unary/Condition/Constant/Cast/Access/Call/Stencil/GridInfo``).
The node set and field names are this package's own; what matters for parity
is that every construct the reference accepts has a node with the same
meaning.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Optional

from .types import BaseType, GridType, Structure


@dataclass
class Loc:
    file: str
    func: str
    line: int

    def __repr__(self) -> str:
        return f"File {self.file}, line {self.line}, in {self.func}"


@dataclass
class Var:
    name: str
    type: BaseType


# ---------------------------------------------------------------- expressions
@dataclass(eq=False)
class Expr:
    loc: Loc
    type: BaseType


@dataclass(eq=False)
class Const(Expr):
    value: Any = None


@dataclass(eq=False)
class Name(Expr):
    """A local or argument.  ``deref`` is set for ``ptr[T]`` arguments."""
    var: Var = None
    ctx: str = "load"

    @property
    def deref(self) -> bool:
        from .types import Pointer
        return isinstance(self.var.type, Pointer)


@dataclass(eq=False)
class Field(Expr):
    """``value.attr`` on a structure-typed terminal."""
    value: Expr = None
    attr: str = ""
    ctx: str = "load"


@dataclass(eq=False)
class Cell(Expr):
    """A stencil access ``g[o0, .., oD-1][t]`` (reference ``Stencil``).

    ``time`` keeps the sign the user wrote; level ``abs(time)`` is read.
    ``klass`` is the enclosing ``with boundary(k)`` (0 outside).
    """
    grid: Var = None
    time: int = 0
    offsets: tuple = ()
    klass: int = 0
    ctx: str = "load"

    @property
    def level(self) -> int:
        return abs(self.time)


BIN_OPS = {"+", "-", "*", "/", "**", "%", "==", "!=", ">", ">=", "<", "<=", "&&", "||"}


@dataclass(eq=False)
class Bin(Expr):
    op: str = "+"
    left: Expr = None
    right: Expr = None


@dataclass(eq=False)
class Un(Expr):
    op: str = "-"          # '+', '-', '!'
    operand: Expr = None


@dataclass(eq=False)
class Select(Expr):
    cond: Expr = None
    then: Expr = None
    orelse: Expr = None


@dataclass(eq=False)
class Cast(Expr):
    value: Expr = None


@dataclass(eq=False)
class Call(Expr):
    """Call of a user ``@function``/``@external`` (``callee`` is its Operator)
    or of a structure constructor (``ctor`` is the Structure type)."""
    callee: Any = None
    ctor: Optional[Structure] = None
    args: list = field(default_factory=list)
    # argument types the callee declares (pointer parameters take ``&arg``)
    param_types: list = field(default_factory=list)


@dataclass(eq=False)
class GridInfo(Expr):
    info: str = "shape"      # 'shape' | 'dimension'
    grid: Var = None
    dim: Optional[Expr] = None


# ----------------------------------------------------------------- statements
@dataclass(eq=False)
class Stmt:
    loc: Loc


@dataclass(eq=False)
class Assign(Stmt):
    target: Expr = None
    value: Expr = None


@dataclass(eq=False)
class Eval(Stmt):
    value: Expr = None


@dataclass(eq=False)
class If(Stmt):
    cond: Expr = None
    body: list = field(default_factory=list)
    orelse: list = field(default_factory=list)


@dataclass(eq=False)
class While(Stmt):
    cond: Expr = None
    body: list = field(default_factory=list)


@dataclass(eq=False)
class For(Stmt):
    var: Var = None
    start: Expr = None
    stop: Expr = None
    step: Expr = None
    body: list = field(default_factory=list)


@dataclass(eq=False)
class Return(Stmt):
    value: Optional[Expr] = None


@dataclass(eq=False)
class Break(Stmt):
    pass


@dataclass(eq=False)
class Continue(Stmt):
    pass


@dataclass(eq=False)
class Inline(Stmt):
    source: str = ""


@dataclass(eq=False)
class Function:
    """A parsed ``@kernel``/``@function``/``@external`` definition."""
    loc: Loc
    name: str
    mode: str
    args: list            # [(name, BaseType)]
    ret: BaseType
    scope: dict           # name -> Var (arguments and locals, in definition order)
    body: list
    includes: list = field(default_factory=list)

    @property
    def arg_names(self) -> list:
        return [n for n, _ in self.args]

    @property
    def locals(self) -> list:
        names = set(self.arg_names)
        return [v for n, v in self.scope.items() if n not in names]

    @property
    def grid_args(self) -> list:
        return [(i, n, t) for i, (n, t) in enumerate(self.args) if isinstance(t, GridType)]


# ----------------------------------------------------------------- traversal
def children(node):
    """Yield the direct IR children of a node (expressions and statements)."""
    if isinstance(node, (Const, Name, GridInfo)):
        if isinstance(node, GridInfo) and node.dim is not None:
            yield node.dim
        return
    if isinstance(node, Field):
        yield node.value
    elif isinstance(node, Cell):
        return
    elif isinstance(node, Bin):
        yield node.left
        yield node.right
    elif isinstance(node, Un):
        yield node.operand
    elif isinstance(node, Select):
        yield node.cond
        yield node.then
        yield node.orelse
    elif isinstance(node, Cast):
        yield node.value
    elif isinstance(node, Call):
        yield from node.args
    elif isinstance(node, Assign):
        yield node.target
        yield node.value
    elif isinstance(node, Eval):
        yield node.value
    elif isinstance(node, If):
        yield node.cond
        yield from node.body
        yield from node.orelse
    elif isinstance(node, While):
        yield node.cond
        yield from node.body
    elif isinstance(node, For):
        yield node.start
        yield node.stop
        yield node.step
        yield from node.body
    elif isinstance(node, Return):
        if node.value is not None:
            yield node.value


def walk(node):
    yield node
    for c in children(node):
        yield from walk(c)


def walk_all(stmts):
    for s in stmts:
        yield from walk(s)
