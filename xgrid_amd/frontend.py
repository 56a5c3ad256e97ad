"""The DSL's front end: the Python source of a decorated function -> typed IR.

The language is the reference's (``xgrid/lang/parser.py``), so that kernels run
unchanged:

* statements - assignment (a new local takes the value's type), augmented
  assignment of scalars, ``if``/``else``, ``while`` (no ``else``),
  ``for v in range(a, b[, s])``, ``return``, ``break``, ``continue``,
  ``pass``, ``import a.b`` (an ``#include "a/b.h"``), ``with
  xgrid.boundary(k)`` (statements inside store to boundary class k) and
  ``with xgrid.c()`` (string literals inside are C lines for the host);
* expressions - ``+ - * / ** %``, comparisons, ``and``/``or``/``not``,
  ``x if c else y``, ``typing.cast(T, x)``, dataclass constructors and
  methods, ``@function``/``@external`` calls, ``xgrid.shape(g, i)`` and
  ``xgrid.dimension(g)``, and global names, which are read once, when the
  kernel is first parsed (``parser.py:548-554``);
* grid cells - ``g[o0, ..., oD-1]`` with integer-literal offsets, optionally
  followed by a time subscript ``[t]``; without one, a load reads time -1 and
  a store writes time 0 (``parser.py:512-525``).

Typing (``parser.py:340-426``): operands of an arithmetic operator or a
comparison have one numeric type, with no implicit conversion; ``**`` gives
a double when its left operand is a double, else the configured float; a
comparison chain ``a < b < c`` compares ``a`` with every later operand and
joins the results with ``&&`` (``parser.py:413-414``); integer literals are
int32, float literals the configured float.

How this module is organised (its own design): ``_Rules`` holds the typing
rules as functions of types only; ``_Scope`` the arguments, locals and the
lookup of dotted global names; ``_Lowering`` turns statements and expressions
into IR through two dispatch tables, keeping the pragma state (the boundary
class in force, and a stack of the enclosing block kinds that says whether
a bare string is a line of C).  Every diagnostic is a
``DSLSyntaxError`` whose first line locates the construct and whose second
says what is wrong and, where it helps, what is accepted instead.
"""
from __future__ import annotations

import ast
import inspect
import textwrap
import typing
from functools import reduce

from . import ir
from .config import get_config
from .types import (BaseType, Boolean, Floating, GridType, Integer, Number, Pointer, Reference, Structure, Value,
                    Void, parse_annotation)


class DSLSyntaxError(SyntaxError):
    pass


# ------------------------------------------------------------------ typing rules
class _Rules:
    """Typing of the DSL's operators, as functions of operand types.  Each
    returns the result type, or a string saying why the operands do not fit."""

    ARITH = {ast.Add: "+", ast.Sub: "-", ast.Mult: "*", ast.Div: "/", ast.Pow: "**", ast.Mod: "%"}
    COMPARE = {ast.Eq: "==", ast.NotEq: "!=", ast.Gt: ">", ast.GtE: ">=", ast.Lt: "<", ast.LtE: "<="}
    LOGIC = {ast.And: "&&", ast.Or: "||"}
    UNARY = {ast.UAdd: "+", ast.USub: "-", ast.Not: "!"}

    @staticmethod
    def literal(value):
        """Type of a Python literal (bool before int: True is an int in Python)."""
        if type(value) is bool:
            return Boolean()
        if type(value) is int:
            return Integer(4)
        if type(value) is float:
            return Floating(get_config().fsize)
        return f"{value!r} ({type(value).__name__}) is not a DSL value: only int, float and bool literals are"

    @staticmethod
    def arith(op: str, lt, rt):
        if not (isinstance(lt, Number) and isinstance(rt, Number)):
            return f"'{op}' needs numeric operands, got {lt} and {rt}"
        if lt != rt:
            return f"'{op}' needs operands of one type (there is no implicit conversion), got {lt} and {rt}"
        if op == "**":
            wide = isinstance(lt, Floating) and lt.width_bytes == 8
            return Floating(8) if wide else Floating(get_config().fsize)
        return lt

    @staticmethod
    def unary(op: str, t):
        if op == "!":
            return t if isinstance(t, Boolean) else f"'not' needs a bool operand, got {t}"
        return t if isinstance(t, Number) else f"unary '{op}' needs a numeric operand, got {t}"

    @staticmethod
    def augmented(op: str, tt, vt):
        if op == "**":
            return "'**=' is not part of the language; write 'x = x ** y'"
        if not (isinstance(tt, Number) and isinstance(vt, Number)) or tt != vt:
            return f"'{op}=' needs a numeric target and value of one type, got {tt} and {vt}"
        return tt

    @staticmethod
    def loop_range(types: list):
        step = types[2]
        if not isinstance(step, Number) or any(t != step for t in types):
            return f"range() bounds and step must share one numeric type, got {', '.join(map(str, types))}"
        return step


def _int_literal(node):
    """k or -k for an integer literal k, else None."""
    negated = isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub)
    k = node.operand if negated else node
    if isinstance(k, ast.Constant) and type(k.value) is int:
        return -k.value if negated else k.value
    return None


def _is_store(node) -> bool:
    return isinstance(getattr(node, "ctx", None), ast.Store)


# ------------------------------------------------------------------ names
class _Scope:
    """Arguments and locals of the function being parsed (in definition
    order), and the global names it can see: its module's globals, then its
    closure's free variables, then the builtins int, float, bool and range."""

    BUILTINS = {"int": int, "float": float, "bool": bool, "range": range}

    def __init__(self, func):
        self.visible = dict(func.__globals__)
        try:
            self.visible.update(inspect.getclosurevars(func).nonlocals)
        except (TypeError, ValueError):
            pass
        self.vars: dict[str, ir.Var] = {}

    def declare(self, name: str, t) -> ir.Var:
        v = ir.Var(name, t)
        self.vars[name] = v
        return v

    def local(self, name: str):
        return self.vars.get(name)

    def dotted(self, node) -> list:
        """['a', 'b', 'c'] for the expression a.b.c, or None if it is not a dotted name."""
        parts = []
        while isinstance(node, ast.Attribute):
            parts.append(node.attr)
            node = node.value
        if not isinstance(node, ast.Name):
            return None
        parts.append(node.id)
        return parts[::-1]

    def global_object(self, path: list):
        """(object, None) for the object a dotted global name denotes, else (None,
        the first prefix that does not resolve).  A dict along the path is
        indexed by the next name, any other object gives its attribute."""
        head = path[0]
        if head in self.visible:
            obj = self.visible[head]
        elif head in self.BUILTINS:
            obj = self.BUILTINS[head]
        else:
            return None, head
        for i, attr in enumerate(path[1:], 1):
            if isinstance(obj, dict) and attr in obj:
                obj = obj[attr]
            elif not isinstance(obj, dict) and hasattr(obj, attr):
                obj = getattr(obj, attr)
            else:
                return None, ".".join(path[:i + 1])
        return obj, None


# ------------------------------------------------------------------ lowering
class _Lowering:
    def __init__(self, func, name: str, mode: str, self_type: BaseType | None):
        self.func, self.name, self.mode, self.self_type = func, name, mode, self_type
        self.file = inspect.getsourcefile(func) or "<unknown>"
        lines, first = inspect.getsourcelines(func)
        self.line0 = max(first, 1) - 1
        module = ast.parse(textwrap.dedent("".join(lines)))
        self.fdef = module.body[0]
        if not isinstance(self.fdef, ast.FunctionDef):
            raise DSLSyntaxError(f"{name}: a DSL operator must be a plain function definition")
        self.scope = _Scope(func)
        # pragma state, modelled as the reference keeps it: one boundary class in force
        # (0 again after any boundary block ends, parser.py:229-233), and a stack of
        # block kinds whose top says whether a bare string is C text ('c') or not
        self.klass = 0
        self.blocks: list = ["body"]
        self.includes: list = []

    # ---- diagnostics and locations
    def at(self, node) -> ir.Loc:
        return ir.Loc(self.file, self.name, self.line0 + getattr(node, "lineno", 1))

    def fail(self, node, message: str):
        loc = self.at(node)
        raise DSLSyntaxError(f"{loc.file}:{loc.line}: in '{self.name}'\n  {message}")

    def nested(self, kind: str, nodes) -> list:
        self.blocks.append(kind)
        try:
            return self.statements(nodes)
        finally:
            self.blocks.pop()

    # ---- the function
    def function(self) -> ir.Function:
        node = self.fdef
        args = []
        for pname, p in inspect.signature(self.func).parameters.items():
            if p.kind is not inspect.Parameter.POSITIONAL_OR_KEYWORD:
                self.fail(node, f"parameter '{pname}' is {p.kind.description}: "
                                f"operators take plain positional parameters only")
            if p.annotation is inspect.Parameter.empty:
                if self.self_type is None:
                    self.fail(node, f"parameter '{pname}' needs a type annotation")
                t = self.self_type
            else:
                t = parse_annotation(p.annotation, self.scope.visible)
            if t is None or isinstance(t, Void):
                self.fail(node, f"parameter '{pname}': {p.annotation!r} is not a value, grid or pointer type")
            self.scope.declare(pname, t)
            args.append((pname, t))
        ann = inspect.signature(self.func).return_annotation
        ret = None if ann is inspect.Signature.empty else parse_annotation(ann, self.scope.visible)
        if ret is None or isinstance(ret, Reference):
            self.fail(node, f"the return annotation must be a value type or None, not {ret!r}")
        self.ret = ret
        stmts = list(node.body)
        if self.mode == "external":
            stmts = []                           # declared elsewhere (a C header)
        elif stmts and isinstance(stmts[0], ast.Expr) and isinstance(stmts[0].value, ast.Constant) \
                and isinstance(stmts[0].value.value, str):
            stmts = stmts[1:]                    # docstring
        body = self.statements(stmts)
        return ir.Function(self.at(node), self.name, self.mode, args, ret, self.scope.vars, body, self.includes)

    def statements(self, nodes) -> list:
        out = []
        for n in nodes:
            rule = _STATEMENTS.get(type(n))
            if rule is None:
                self.fail(n, f"'{type(n).__name__}' statements are not part of the DSL")
            out.extend(rule(self, n))
        return out

    def expr(self, node) -> ir.Expr:
        rule = _EXPRESSIONS.get(type(node))
        if rule is None:
            self.fail(node, f"'{type(node).__name__}' expressions are not part of the DSL")
        return rule(self, node)

    def checked(self, node, result):
        """A rule's result type, or a diagnostic if it is a message."""
        if isinstance(result, str):
            self.fail(node, result)
        return result

    # ---- terminals: locals, struct fields, grid cells
    def terminal(self, node):
        """IR for a local name, a field of a local structure or a grid cell;
        None when the name is not local (it may be global)."""
        if isinstance(node, ast.Name):
            v = self.scope.local(node.id)
            if v is None:
                return None
            t = v.type.element if isinstance(v.type, Pointer) else v.type
            return ir.Name(self.at(node), t, v, "store" if _is_store(node) else "load")
        if isinstance(node, ast.Attribute):
            base = self.terminal(node.value)
            if base is None or not isinstance(base.type, Structure):
                return None
            ft = base.type.elements_map.get(node.attr)
            if ft is None:
                return None
            return ir.Field(self.at(node), ft, base, node.attr, "store" if _is_store(node) else "load")
        if isinstance(node, ast.Subscript):
            return self.cell(node)
        self.fail(node, f"'{type(node).__name__}' cannot be read or assigned here")

    def cell(self, node: ast.Subscript) -> ir.Cell:
        store = _is_store(node)
        if isinstance(node.value, ast.Subscript):       # g[offsets][t]
            if store:
                # (the reference's generator rejects such a store too: it takes the
                # subscripted cell for a load outside a stencil, generator.py:57-65)
                self.fail(node, "a store writes the current level: g[...][t] is for loads only; write g[...] = ...")
            t = _int_literal(node.slice)
            if t is None:
                self.fail(node, "a time subscript g[...][t] takes an integer literal")
            space = node.value
        else:
            t = 0 if store else -1
            space = node
        if not isinstance(space.value, ast.Name):
            self.fail(node, "only a grid argument can be subscripted")
        g = self.scope.local(space.value.id)
        if g is None:
            self.fail(node, f"'{space.value.id}' is not a grid argument of this operator")
        if not isinstance(g.type, GridType):
            self.fail(node, f"'{g.name}' is {g.type}, not a grid: only grids take subscripts")
        items = space.slice.elts if isinstance(space.slice, ast.Tuple) else [space.slice]
        offsets = []
        for it in items:
            o = _int_literal(it)
            if o is None:
                self.fail(node, f"grid offsets are integer literals relative to the cell, not {ast.unparse(it)!r}")
            offsets.append(o)
        if len(offsets) != g.type.dimension:
            self.fail(node, f"'{g.name}' has {g.type.dimension} dimension(s) but the subscript gives "
                            f"{len(offsets)} offset(s)")
        return ir.Cell(self.at(space), g.type.element, g, t, tuple(offsets), self.klass,
                       "store" if store else "load")

    def global_value(self, node):
        path = self.scope.dotted(node)
        if path is None:
            self.fail(node, f"'{ast.unparse(node)}' is not a name")
        obj, missing = self.scope.global_object(path)
        if missing is not None:
            self.fail(node, f"'{missing}' is not defined (neither a local nor a global visible to the kernel)")
        return obj

    def constant(self, node, value) -> ir.Const:
        return ir.Const(self.at(node), self.checked(node, _Rules.literal(value)), value)


# ---------------------------------------------------------------- statements
def _s_pass(lw, node):
    return []


def _s_break(lw, node):
    return [ir.Break(lw.at(node))]


def _s_continue(lw, node):
    return [ir.Continue(lw.at(node))]


def _s_return(lw, node):
    value = None if node.value is None else lw.expr(node.value)
    if value is not None and value.type != lw.ret:
        lw.fail(node, f"returns {value.type}, but the operator is declared to return {lw.ret}")
    return [ir.Return(lw.at(node), value)]


def _s_if(lw, node):
    cond = lw.expr(node.test)
    return [ir.If(lw.at(node), cond, lw.nested("if", node.body), lw.nested("if", node.orelse))]


def _s_while(lw, node):
    cond = lw.expr(node.test)
    body = lw.nested("while", node.body)
    if node.orelse:
        lw.fail(node, "'while ... else' is not part of the DSL")
    return [ir.While(lw.at(node), cond, body)]


def _s_expr(lw, node):
    if lw.blocks[-1] == "c":
        v = node.value
        if isinstance(v, ast.Constant) and isinstance(v.value, str):
            return [ir.Inline(lw.at(node), v.value)]
        return []                                   # only C text means something in a c() block
    return [ir.Eval(lw.at(node), lw.expr(node.value))]


def _s_with(lw, node):
    from . import lang
    if len(node.items) != 1 or not isinstance(node.items[0].context_expr, ast.Call):
        lw.fail(node, "'with' takes one pragma call: xgrid.boundary(k) or xgrid.c()")
    call = node.items[0].context_expr
    pragma = lw.global_value(call.func)
    if pragma is lang.c:
        return lw.nested("c", node.body)
    if pragma is lang.boundary:
        if len(call.args) != 1 or not (isinstance(call.args[0], ast.Constant) and type(call.args[0].value) is int):
            lw.fail(node, "xgrid.boundary takes exactly one argument, the class as an integer literal")
        lw.klass = call.args[0].value
        body = lw.statements(node.body)
        lw.klass = 0
        return body
    lw.fail(node, f"{pragma!r} is not a pragma: 'with' accepts xgrid.boundary(k) or xgrid.c()")


def _s_assign(lw, node):
    value = lw.expr(node.value)
    if isinstance(value.type, GridType):
        lw.fail(node, "a grid is not a value: it cannot be assigned to a name")
    if len(node.targets) != 1:
        lw.fail(node, "assign one target at a time ('a = b = x' is not part of the DSL)")
    target_node = node.targets[0]
    target = lw.terminal(target_node)
    if target is None:
        if not isinstance(target_node, ast.Name):
            lw.fail(node, f"cannot assign to '{ast.unparse(target_node)}': it is not a local, a field or a grid cell")
        target = ir.Name(lw.at(node), value.type, lw.scope.declare(target_node.id, value.type), "store")
    if target.type != value.type:
        lw.fail(node, f"cannot assign a {value.type} to '{ast.unparse(target_node)}' of type {target.type} "
                      f"(types must match exactly)")
    return [ir.Assign(lw.at(node), target, value)]


def _s_augassign(lw, node):
    value = lw.expr(node.value)
    target = lw.terminal(node.target)
    if target is None:
        lw.fail(node, f"'{ast.unparse(node.target)}' must exist before an augmented assignment")
    op = _Rules.ARITH.get(type(node.op))
    if op is None:
        lw.fail(node, f"'{type(node.op).__name__}=' is not part of the DSL")
    if isinstance(target, ir.Cell):
        lw.fail(node, "a grid cell cannot take an augmented assignment; write g[...] = g[...] op x")
    t = lw.checked(node, _Rules.augmented(op, target.type, value.type))
    loc = lw.at(node)
    return [ir.Assign(loc, target, ir.Bin(loc, t, op, target, value))]


def _s_for(lw, node):
    if not isinstance(node.target, ast.Name):
        lw.fail(node, "the loop variable must be a single name")
    rng = node.iter
    if not (isinstance(rng, ast.Call) and isinstance(rng.func, ast.Name) and rng.func.id == "range"):
        lw.fail(node, "'for' loops iterate over range(start, stop[, step]) only")
    if len(rng.args) not in (2, 3):
        lw.fail(node, "range() takes a start and a stop, and optionally a step")
    if node.orelse:
        lw.fail(node, "'for ... else' is not part of the DSL")
    bounds = [lw.expr(a) for a in rng.args]
    if len(bounds) == 2:
        bounds.append(ir.Const(lw.at(node), Integer(4), 1))
    t = lw.checked(node, _Rules.loop_range([b.type for b in bounds]))
    var = lw.scope.local(node.target.id) or lw.scope.declare(node.target.id, t)
    if var.type != t:
        lw.fail(node, f"loop variable '{var.name}' is {var.type} but the range is {t}")
    return [ir.For(lw.at(node), var, bounds[0], bounds[1], bounds[2], lw.statements(node.body))]


def _s_import(lw, node):
    for alias in node.names:
        if alias.asname is not None:
            lw.fail(node, f"'import {alias.name} as {alias.asname}': an include takes no alias")
        lw.includes.append(alias.name.replace(".", "/") + ".h")
    return []


_STATEMENTS = {
    ast.Pass: _s_pass, ast.Break: _s_break, ast.Continue: _s_continue, ast.Return: _s_return,
    ast.If: _s_if, ast.While: _s_while, ast.Expr: _s_expr, ast.With: _s_with, ast.Assign: _s_assign,
    ast.AugAssign: _s_augassign, ast.For: _s_for, ast.Import: _s_import,
}


# ---------------------------------------------------------------- expressions
def _e_constant(lw, node):
    return lw.constant(node, node.value)


def _e_name(lw, node):
    # a local, a field of a local structure, or else a global folded to a literal
    t = lw.terminal(node)
    return t if t is not None else lw.constant(node, lw.global_value(node))


def _e_subscript(lw, node):
    return lw.cell(node)


def _e_unary(lw, node):
    op = _Rules.UNARY.get(type(node.op))
    if op is None:
        lw.fail(node, f"unary '{type(node.op).__name__}' is not part of the DSL")
    x = lw.expr(node.operand)
    return ir.Un(lw.at(node), lw.checked(node, _Rules.unary(op, x.type)), op, x)


def _e_binary(lw, node):
    op = _Rules.ARITH.get(type(node.op))
    if op is None:
        lw.fail(node, f"'{type(node.op).__name__}' is not one of the DSL's operators + - * / ** %")
    a, b = lw.expr(node.left), lw.expr(node.right)
    return ir.Bin(lw.at(node), lw.checked(node, _Rules.arith(op, a.type, b.type)), op, a, b)


def _e_logic(lw, node):
    loc = lw.at(node)
    op = _Rules.LOGIC[type(node.op)]
    terms = [lw.expr(v) for v in node.values]
    for t in terms:
        if not isinstance(t.type, Boolean):
            lw.fail(node, f"'{'and' if op == '&&' else 'or'}' joins bool values, got {t.type}")
    return reduce(lambda x, y: ir.Bin(loc, Boolean(), op, x, y), terms)


def _e_compare(lw, node):
    loc = lw.at(node)
    first = lw.expr(node.left)
    if not isinstance(first.type, Number):
        lw.fail(node, f"comparisons need numeric operands, got {first.type}")
    others = [lw.expr(c) for c in node.comparators]
    tests = []
    for o, other in zip(node.ops, others):
        op = _Rules.COMPARE.get(type(o))
        if op is None:
            lw.fail(node, f"'{type(o).__name__}' is not one of the DSL's comparisons == != < <= > >=")
        if other.type != first.type:
            lw.fail(node, f"'{op}' compares values of one type, got {first.type} and {other.type}")
        # each later operand is compared with the first one (parser.py:413-414)
        tests.append(ir.Bin(loc, Boolean(), op, first, other))
    return reduce(lambda x, y: ir.Bin(loc, Boolean(), "&&", x, y), tests)


def _e_select(lw, node):
    cond = lw.expr(node.test)
    if not isinstance(cond.type, Boolean):
        lw.fail(node, f"the condition of 'x if c else y' must be a bool, got {cond.type}")
    a, b = lw.expr(node.body), lw.expr(node.orelse)
    if a.type != b.type or not isinstance(a.type, Value):
        lw.fail(node, f"both branches of 'x if c else y' must be values of one type, got {a.type} and {b.type}")
    return ir.Select(lw.at(node), a.type, cond, a, b)


def _e_call(lw, node):
    from .operator import Operator, method_operator
    loc = lw.at(node)
    args: list = []
    label = None
    receiver = lw.terminal(node.func.value) if isinstance(node.func, ast.Attribute) else None
    if receiver is not None:
        # s.method(...) on a local structure: a @function(method=True) of its dataclass
        if not isinstance(receiver.type, Structure):
            lw.fail(node, f"'{ast.unparse(node.func)}': {receiver.type} has no methods")
        raw = getattr(receiver.type.dataclass, node.func.attr, None)
        if getattr(raw, "__xgrid_method", None) is None:
            lw.fail(node, f"'{ast.unparse(node.func)}' is not a @xgrid.function(method=True) of "
                          f"{receiver.type.dataclass.__qualname__}")
        callee = method_operator(raw, receiver.type)
        label = f"{receiver.type.dataclass.__qualname__}.{node.func.attr}"
        args.append(receiver)
    else:
        callee = lw.global_value(node.func)

    if callee is typing.cast:
        if len(node.args) != 2:
            lw.fail(node, "cast(T, x) takes a type and a value")
        target = parse_annotation(lw.global_value(node.args[0]), lw.scope.visible)
        if target is None:
            lw.fail(node, f"cast target '{ast.unparse(node.args[0])}' is not a DSL type")
        return ir.Cast(loc, target, lw.expr(node.args[1]))

    if node.keywords:
        lw.fail(node, "arguments are passed by position only")
    args.extend(lw.expr(a) for a in node.args)

    if isinstance(callee, type):
        st = parse_annotation(callee, lw.scope.visible)
        if not isinstance(st, Structure):
            lw.fail(node, f"'{callee.__name__}' is not a dataclass the DSL can construct")
        if len(st.elements) != len(args):
            lw.fail(node, f"{st.name}(...) takes {len(st.elements)} field value(s), got {len(args)}")
        for (fname, ft), a in zip(st.elements, args):
            if a.type != ft:
                lw.fail(node, f"{st.name}(...): field '{fname}' is {ft}, got {a.type}")
        return ir.Call(loc, st, None, st, args, [t for _, t in st.elements])

    if not isinstance(callee, Operator):
        lw.fail(node, f"'{ast.unparse(node.func)}' is not callable in a kernel: call @xgrid.function, "
                      f"@xgrid.external operators, dataclass constructors or cast")
    label = label or callee.name
    params = callee.signature_args()
    if len(params) != len(args):
        lw.fail(node, f"'{label}' takes {len(params)} argument(s), got {len(args)}")
    if callee.mode == "external" and callee.typecheck_override is not None:
        try:
            rtype = callee.typecheck_override([a.type for a in args])
        except Exception as e:        # the override's own message says what is wrong
            lw.fail(node, f"'{label}': {e.args[0] if e.args else e}")
        if callee.intrinsic in ("shape", "dimension"):
            if not isinstance(args[0], ir.Name):
                lw.fail(node, f"xgrid.{callee.intrinsic} takes a grid argument by name")
            return ir.GridInfo(loc, Integer(4), callee.intrinsic, args[0].var,
                               args[1] if callee.intrinsic == "shape" else None)
    else:
        for (pname, pt), a in zip(params, args):
            if a.type != pt:
                lw.fail(node, f"'{label}': parameter '{pname}' is {pt}, got {a.type}")
        rtype = callee.signature_ret()
    return ir.Call(loc, rtype, callee, None, args, [t for _, t in params])


_EXPRESSIONS = {
    ast.Constant: _e_constant, ast.Name: _e_name, ast.Attribute: _e_name, ast.Subscript: _e_subscript,
    ast.UnaryOp: _e_unary, ast.BinOp: _e_binary, ast.BoolOp: _e_logic, ast.Compare: _e_compare,
    ast.IfExp: _e_select, ast.Call: _e_call,
}


def parse(func, name: str, mode: str, self_type: BaseType | None = None) -> ir.Function:
    """The typed IR of a @kernel / @function / @external Python function."""
    return _Lowering(func, name, mode, self_type).function()
